// FP8 (OCP e4m3) form of the projection B = U * (W (*) X) on gfx950's
// v_mfma_f32_32x32x64_f8f6f4 (64-deep e4m3 products) with FP32 accumulation.
//
// Reference: kilosort/template_matching.py:180-181 (computed there in FP32).
//
// Scales keep every operand inside e4m3's range (|v| <= 448; the conversion
// does not saturate):
//   correlation row (c, p):  s[c][p] = max_t |X[c, t]| * sum_k |W[p, k]| / 448,
//     a bound on max_t |Bc[c, p, t]|, so Bq[t][c P + p] = e4m3(Bc / s);
//   template m (peel order):  u[m] = max_r |U[m, r] s[r]| / 448,
//     Uq[m][r] = e4m3(U[m, r] s[r] / u[m]);
//   B[m, t] = u[m] * sum_r Uq[m][r] Bq[t][r]     (r = c P + p).
// Both operands are contiguous along r, as in the BF16 forms.  A template
// tile's reduction steps where Uq is all zero are skipped (exact zeros only;
// every step is kept when an input is not finite).
#include <cmath>

#include "ksm_common.h"
#include "ksm_profile.h"
#include "ksm_project_internal.h"
#include "ksm_bound.h"
#include "ksm_graph.h"

namespace {

typedef unsigned char fp8_t;
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr float kE4M3Max = 448.f;
// standard deviations of the modelled rounding error in the FP8 decision margin
constexpr double kFp8Sigmas = 6.0;

// Finite values beyond e4m3's range are clamped to +-448 (the conversion
// does not saturate); NaN and +-inf pass and convert to NaN, so a non-finite
// input stays non-finite in the product.
__device__ __forceinline__ float clamp_e4m3(float a) {
    const float m = fabsf(a);
    return (m > kE4M3Max && m < __builtin_inff()) ? __builtin_copysignf(kE4M3Max, a) : a;
}

// Two floats -> two e4m3 bytes (round to nearest even).
__device__ __forceinline__ unsigned pack_fp8x2(float a, float b) {
    return (unsigned)__builtin_amdgcn_cvt_pk_fp8_f32(clamp_e4m3(a), clamp_e4m3(b), 0, false) &
           0xffffu;
}

// amax[c] = max_t |X[c, t]| (a NaN anywhere raises the non-finite flag).
__global__ __launch_bounds__(256) void chan_absmax_kernel(const float *__restrict__ X, int ldx,
                                                          int T, float *__restrict__ amax,
                                                          int *__restrict__ nonfinite) {
    __shared__ float s_m[4];
    const int c = blockIdx.x;
    const float *xr = X + (size_t)c * ldx;
    float m = 0.f;
    bool bad = false;
    for (int t = threadIdx.x; t < T; t += 256) {
        const float v = xr[t];
        bad |= !__builtin_isfinite(v);
        m = fmaxf(m, fabsf(v));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (__builtin_amdgcn_ballot_w64(bad) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(nonfinite, 1);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) amax[c] = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
}

// s[c P + p] for c < C, 1 for padding rows.
__global__ void row_scales_kernel(const float *__restrict__ amax, const float *__restrict__ W,
                                  int C, int P, int nt, int ldr, float *__restrict__ s) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ldr) return;
    const int c = r / P, p = r % P;
    float v = 1.f;
    if (c < C) {
        float l1 = 0.f;
        for (int k = 0; k < nt; ++k) l1 += fabsf(W[p * nt + k]);
        const float b = amax[c] * l1 / kE4M3Max;
        v = b > 0.f && __builtin_isfinite(b) ? b : 1.f;
    }
    s[r] = v;
}

constexpr int kCvCh = 64;     // channels per correlation block (one per lane)
constexpr int kCvTT = 8;      // time samples per thread
constexpr int kCvT = 4 * kCvTT;

// Bq[t][c P + p] = e4m3(Bc[c, p, t] / s[c P + p]) for t < Tp (zero for t >= T
// or c >= C).
//
// stats (may be NULL): per 32-channel sub-block b (two per block) and
// column t, over the sub-block's (channel, PC) rows: the sum of squares of
// the FP32 correlation (j = 0), its largest square (j = 1) and the sum of
// squares of the scaled values Bc / s (j = 2), at
// stats[(j * nsb2 + b) * Tp + t] (inputs of the error bound, ksm_bound.h).
template <int P, int NTAP>
__global__ __launch_bounds__(256) void conv_pc_fp8_kernel(
    const float *__restrict__ X, int ldx, const float *__restrict__ W, int C, int T, int Tp,
    int ldr, const float *__restrict__ srow, fp8_t *__restrict__ Bq, float *__restrict__ stats) {
    constexpr int H = NTAP / 2;
    constexpr int SPAN = kCvT + NTAP - 1;
    constexpr int LD = SPAN | 1;
    __shared__ float xs[kCvCh * LD];
    const int t0 = blockIdx.x * kCvT;
    const int c0 = blockIdx.y * kCvCh;
    for (int e = threadIdx.x; e < kCvCh * SPAN; e += 256) {
        const int cl = e / SPAN, j = e % SPAN;
        const int c = c0 + cl, t = t0 - H + j;
        xs[cl * LD + j] = (c < C && t >= 0 && t < T) ? X[(size_t)c * ldx + t] : 0.f;
    }
    __syncthreads();
    const int cl = threadIdx.x & 63, tg = threadIdx.x >> 6;
    const float *xr = xs + cl * LD + tg * kCvTT;
    float win[kCvTT + NTAP - 1];
#pragma unroll
    for (int e = 0; e < kCvTT + NTAP - 1; ++e) win[e] = xr[e];
    float acc[kCvTT][P];
#pragma unroll
    for (int u = 0; u < kCvTT; ++u)
#pragma unroll
        for (int p = 0; p < P; ++p) acc[u][p] = 0.f;
#pragma unroll
    for (int k = 0; k < NTAP; ++k) {
        float w[P];
#pragma unroll
        for (int p = 0; p < P; ++p) w[p] = W[p * NTAP + k];
#pragma unroll
        for (int u = 0; u < kCvTT; ++u)
#pragma unroll
            for (int p = 0; p < P; ++p) acc[u][p] = __builtin_fmaf(win[u + k], w[p], acc[u][p]);
    }
    const int c = c0 + cl;
    float inv[P];
#pragma unroll
    for (int p = 0; p < P; ++p) inv[p] = c * P < ldr ? 1.f / srow[c * P + p] : 1.f;
    if (stats) {
        // per 32-channel half of the block (lanes 0-31, 32-63)
        const int nsb2 = 2 * gridDim.y, sb = 2 * blockIdx.y + (cl >> 5);
#pragma unroll
        for (int u = 0; u < kCvTT; ++u) {
            const int t = t0 + tg * kCvTT + u;
            const bool live = t < T && c < C;
            float sq = 0.f, mx = 0.f, sq2 = 0.f;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float a = live ? acc[u][p] : 0.f;
                const float q = a * inv[p];
                sq = __builtin_fmaf(a, a, sq);
                sq2 = __builtin_fmaf(q, q, sq2);
                mx = fmaxf(mx, a * a);
                if (a != a) mx = a;             // NaN reaches the bound
            }
            sq = ksm::half_sum_f(sq);
            sq2 = ksm::half_sum_f(sq2);
            mx = ksm::half_max_nan(mx);
            if ((cl & 31) == 0 && t < Tp) {
                stats[((size_t)0 * nsb2 + sb) * Tp + t] = sq;
                stats[((size_t)1 * nsb2 + sb) * Tp + t] = mx;
                stats[((size_t)2 * nsb2 + sb) * Tp + t] = sq2;
            }
        }
    }
    if (c * P >= ldr) return;
    static_assert(P % 2 == 0, "pairs of e4m3 bytes are stored as 16-bit words");
#pragma unroll
    for (int u = 0; u < kCvTT; ++u) {
        const int t = t0 + tg * kCvTT + u;
        if (t >= Tp) break;
        const bool live = t < T && c < C;
        unsigned short *dst = reinterpret_cast<unsigned short *>(Bq + (size_t)t * ldr + c * P);
#pragma unroll
        for (int p = 0; p < P; p += 2) {
            const float x0 = live ? acc[u][p] * inv[p] : 0.f;
            const float x1 = live ? acc[u][p + 1] * inv[p + 1] : 0.f;
            dst[p / 2] = (unsigned short)pack_fp8x2(x0, x1);
        }
    }
}

// Uq[m][r] and u[m] (one block per template position m of the peel order).
__global__ __launch_bounds__(256) void u_fp8_kernel(const float *__restrict__ U,
                                                    const int32_t *__restrict__ perm, int K,
                                                    int P, int C, int ldr,
                                                    const float *__restrict__ srow,
                                                    fp8_t *__restrict__ Uq,
                                                    float *__restrict__ uscale) {
    __shared__ float s_m[4];
    const int m = blockIdx.x;
    const int id = m < K ? (perm ? perm[m] : m) : 0;
    float mx = 0.f;
    for (int r = threadIdx.x; r < P * C; r += 256) {
        const int c = r / P, p = r % P;
        if (m < K) mx = fmaxf(mx, fabsf(U[((size_t)id * P + p) * C + c] * srow[r]));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = mx;
    __syncthreads();
    const float um = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3])) / kE4M3Max;
    const float us = um > 0.f && __builtin_isfinite(um) ? um : 1.f;
    if (threadIdx.x == 0) uscale[m] = us;
    const float inv = 1.f / us;
    unsigned short *row = reinterpret_cast<unsigned short *>(Uq + (size_t)m * ldr);
    for (int r2 = threadIdx.x; r2 < ldr / 2; r2 += 256) {
        float v[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int r = 2 * r2 + e, c = r / P, p = r % P;
            v[e] = (m < K && c < C) ? U[((size_t)id * P + p) * C + c] * srow[r] * inv : 0.f;
        }
        row[r2] = (unsigned short)pack_fp8x2(v[0], v[1]);
    }
}

constexpr int QBM = 128, QBN = 128, QBK = 128;   // tile (m, n) and depth step (bytes)
constexpr int QLD = QBK + 16;                     // padded LDS row (bytes)

// Steps (QBK columns of r) where a tile of QBM rows of Uq is not all zero;
// every step when an input is not finite.  steps[tm][0 .. counts[tm]).
constexpr int kQStepsMax = 4096;
__global__ __launch_bounds__(256) void tile_steps_fp8_kernel(
    const fp8_t *__restrict__ Uq, int ldr, int nk, const float *__restrict__ W, int PW,
    const int *__restrict__ nonfinite, int *__restrict__ steps, int *__restrict__ counts) {
    __shared__ unsigned char s_need[kQStepsMax];
    __shared__ int s_dense;
    const int tm = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_dense = *nonfinite != 0 || nk > kQStepsMax;
    __syncthreads();
    bool wbad = false;
    for (int e = tid; e < PW; e += 256) wbad |= !__builtin_isfinite(W[e]);
    if (wbad) s_dense = 1;
    __syncthreads();
    const bool dense = s_dense != 0;
    if (!dense) {
        const int rsub = lane >> 3, chunk = lane & 7;      // 8 rows x 8 chunks of 16 B
        for (int st = wave; st < nk; st += 4) {
            bool nz = false;
            for (int i = 0; i < QBM / 8; ++i) {
                const uint4 v = *reinterpret_cast<const uint4 *>(
                    Uq + (size_t)(tm * QBM + rsub + 8 * i) * ldr + st * QBK + chunk * 16);
                const unsigned z = 0x7f7f7f7fu;            // e4m3 +0 and -0 are zero
                nz |= ((v.x & z) | (v.y & z) | (v.z & z) | (v.w & z)) != 0u;
            }
            const bool any = __builtin_amdgcn_ballot_w64(nz) != 0ull;
            if (lane == 0) s_need[st] = any;
        }
    }
    __syncthreads();
    if (wave == 0) {
        int cnt = 0;
        for (int c0 = 0; c0 < nk; c0 += 64) {
            const int st = c0 + lane;
            const bool f = st < nk && (dense || s_need[st]);
            const unsigned long long m = __builtin_amdgcn_ballot_w64(f);
            const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (f) steps[(size_t)tm * nk + pos] = st;
            cnt += __popcll(m);
        }
        if (lane == 0) counts[tm] = cnt;
    }
}

__device__ __forceinline__ int xcd_remap_q(int bid, int nb) {
    const int xcd = bid & 7, pos = bid >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// TN GEMM Bout[n][m] = u[m] * sum_r Uq[m][r] Bq[n][r] over the tile's steps.
// Block tile 128 x 128, 4 waves each 64 x 64 as 2 x 2 tiles of 32x32x64.
__global__ __launch_bounds__(256, 2) void gemm_fp8_mfma(
    const fp8_t *__restrict__ Uq, const fp8_t *__restrict__ Bq, int ldr,
    const float *__restrict__ uscale, const int *__restrict__ steps,
    const int *__restrict__ counts, float *__restrict__ Bout, int ldb, int K, int nTm, int nTn) {
    __shared__ fp8_t As[2][QBM * QLD];
    __shared__ fp8_t Bs[2][QBN * QLD];
    const int logical = xcd_remap_q(blockIdx.x, nTm * nTn);
    const int tm = logical % nTm, tn = logical / nTm;
    const int m0 = tm * QBM, n0 = tn * QBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int grow = tid >> 1, gcol = (tid & 1) * 64;   // 64 bytes per thread per step

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    const int nk_all = ldr / QBK;
    const int nk = counts[tm];
    const int *lst = steps + (size_t)tm * nk_all;
    const int k00 = nk > 0 ? lst[0] * QBK : 0;
    const uint4 *pa = reinterpret_cast<const uint4 *>(Uq + (size_t)(m0 + grow) * ldr + k00 + gcol);
    const uint4 *pb = reinterpret_cast<const uint4 *>(Bq + (size_t)(n0 + grow) * ldr + k00 + gcol);
    uint4 ra0 = pa[0], ra1 = pa[1], ra2 = pa[2], ra3 = pa[3];
    uint4 rb0 = pb[0], rb1 = pb[1], rb2 = pb[2], rb3 = pb[3];
    {
        uint4 *la = reinterpret_cast<uint4 *>(&As[0][grow * QLD + gcol]);
        uint4 *lb = reinterpret_cast<uint4 *>(&Bs[0][grow * QLD + gcol]);
        la[0] = ra0; la[1] = ra1; la[2] = ra2; la[3] = ra3;
        lb[0] = rb0; lb[1] = rb1; lb[2] = rb2; lb[3] = rb3;
    }
    __syncthreads();
    // v_mfma_f32_32x32x64_f8f6f4 operands: lane l holds row l % 32 and the 32
    // consecutive depth bytes 32 (l / 32) .. + 31 of the 64-byte step
    const int kr = (lane >> 5) * 32, kc = lane & 31;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            const int k0 = lst[kt + 1] * QBK;
            pa = reinterpret_cast<const uint4 *>(Uq + (size_t)(m0 + grow) * ldr + k0 + gcol);
            pb = reinterpret_cast<const uint4 *>(Bq + (size_t)(n0 + grow) * ldr + k0 + gcol);
            ra0 = pa[0]; ra1 = pa[1]; ra2 = pa[2]; ra3 = pa[3];
            rb0 = pb[0]; rb1 = pb[1]; rb2 = pb[2]; rb3 = pb[3];
        }
#pragma unroll
        for (int kk = 0; kk < QBK; kk += 64) {
            i32x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const i32x8 *>(
                    &As[buf][(wm * 64 + i * 32 + kc) * QLD + kk + kr]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const i32x8 *>(
                    &Bs[buf][(wn * 64 + j * 32 + kc) * QLD + kk + kr]);
            // formats 0, 0: both operands e4m3; scales 0: the unscaled form
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        a[i], b[j], acc[i][j], 0, 0, 0, 0, 0, 0);
        }
        if (more) {
            uint4 *la = reinterpret_cast<uint4 *>(&As[buf ^ 1][grow * QLD + gcol]);
            uint4 *lb = reinterpret_cast<uint4 *>(&Bs[buf ^ 1][grow * QLD + gcol]);
            la[0] = ra0; la[1] = ra1; la[2] = ra2; la[3] = ra3;
            lb[0] = rb0; lb[1] = rb1; lb[2] = rb2; lb[3] = rb3;
        }
        __syncthreads();
    }
    // D layout of 32x32 MFMA: lane l, register r -> row 8(r/4) + 4(l/32) + r%4,
    // column l%32; stored transposed, Bout[n][m], scaled per template row.
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int m = m0 + wm * 64 + i * 32 + 8 * g + 4 * (lane >> 5);
                const int n = n0 + wn * 64 + j * 32 + (lane & 31);
                float *dst = Bout + (size_t)n * ldb + m;
                if (m + 3 < K) {
                    float4 v;
                    v.x = acc[i][j][4 * g + 0] * uscale[m + 0];
                    v.y = acc[i][j][4 * g + 1] * uscale[m + 1];
                    v.z = acc[i][j][4 * g + 2] * uscale[m + 2];
                    v.w = acc[i][j][4 * g + 3] * uscale[m + 3];
                    *reinterpret_cast<float4 *>(dst) = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (m + e < K) dst[e] = acc[i][j][4 * g + e] * uscale[m + e];
                }
            }
}

__global__ void set_flag_q_kernel(int *__restrict__ p) { *p = 0; }

struct QLayout {
    int ldr, Kp, Tp, nTm, nTn, nk;
};

QLayout q_layout(int C, int T, int K, int P) {
    QLayout L;
    L.ldr = (int)ksm::align_up((size_t)P * C, (size_t)P * QBK);   // multiple of P and QBK
    L.Kp = (int)ksm::align_up(K, QBM);
    L.Tp = (int)ksm::align_up(T, QBN);
    L.nTm = L.Kp / QBM;
    L.nTn = L.Tp / QBN;
    L.nk = L.ldr / QBK;
    return L;
}

// Absolute (subnormal) error terms of the FP8 product per unit template
// norm: each scaled operand is e4m3(x) = x (1 + d) + h with |h| <= 2^-10, so
// beyond the relative part a term of r = (c, p) errs by at most
// 2^-10 (1 + 2^-4) (|U_r| s_r + u |Bc_r| / s_r) + 2^-20 u; with
// u <= ||U|| max_r s_r / 448 and Cauchy-Schwarz over the n = ldr terms:
//   a0 = 2^-10 (1 + 2^-4) ||s|| + 2^-20 n smax / 448,
//   a1 = 2^-10 (1 + 2^-4) sqrt(n) smax / 448   (times ||Bc / s|| per column).
__global__ __launch_bounds__(256) void fp8_abs_terms_kernel(const float *__restrict__ srow,
                                                            int n, float *__restrict__ out) {
    __shared__ double s_sq[4], s_mx[4];
    double sq = 0.0, mx = 0.0;
    for (int r = threadIdx.x; r < n; r += 256) {
        const double v = srow[r];
        sq += v * v;
        mx = v > mx ? v : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sq += __shfl_xor(sq, o, 64);
        const double w = __shfl_xor(mx, o, 64);
        mx = w > mx ? w : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        s_sq[threadIdx.x >> 6] = sq;
        s_mx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = 0.0, m = 0.0;
        for (int w = 0; w < 4; ++w) {
            q += s_sq[w];
            m = s_mx[w] > m ? s_mx[w] : m;
        }
        const double h = 0x1p-10 * (1.0 + 0x1p-4) * (1.0 + 0x1p-16);
        out[0] = (float)((h * sqrt(q) + 0x1p-20 * n * m / 448.0) * (1.0 + 0x1p-16));
        out[1] = (float)(h * sqrt((double)n) * m / 448.0 * (1.0 + 0x1p-16));
    }
}

}  // namespace

namespace ksm {

size_t project_fp8_workspace(int C, int T, int K, int P) {
    const QLayout L = q_layout(C, T, K, P);
    return align_up((size_t)L.Tp * L.ldr, 256) + align_up((size_t)L.Kp * L.ldr, 256) +
           align_up((size_t)C * 4, 256) + align_up((size_t)L.ldr * 4, 256) +
           align_up((size_t)L.Kp * 4, 256) + align_up((size_t)L.nTm * L.nk * 4, 256) +
           align_up((size_t)L.nTm * 4, 256) +
           align_up(3 * 2 * (size_t)((L.ldr / P + kCvCh - 1) / kCvCh) * L.Tp * 4, 256) + 256 + 256 +
           512;
}

int project_fp8(const float *X, int ldx, const float *W, const float *U, const int32_t *perm,
                int C, int T, int K, int P, int nt, float *Bt, int ldbt, void *bound,
                void *workspace, hipStream_t s) {
    if (P != 6 || nt != 61)
        return fail(KSM_ERR_INVALID, "ksm_project: the FP8 form is built for P=6, nt=61");
    const QLayout L = q_layout(C, T, K, P);
    Carver cv(workspace);
    fp8_t *Bq = cv.take<fp8_t>((size_t)L.Tp * L.ldr);
    fp8_t *Uq = cv.take<fp8_t>((size_t)L.Kp * L.ldr);
    float *amax = cv.take<float>(C);
    float *srow = cv.take<float>(L.ldr);
    float *uscale = cv.take<float>(L.Kp);
    int *steps = cv.take<int>((size_t)L.nTm * L.nk);
    int *counts = cv.take<int>(L.nTm);
    int *nonfinite = cv.take<int>(1);
    const int nblk = (L.ldr / P + kCvCh - 1) / kCvCh;
    float *stats = cv.take<float>((size_t)3 * 2 * nblk * L.Tp);
    float *abs_terms = cv.take<float>(2);
    ksm::launch(set_flag_q_kernel, 1, 1, 0, s, nonfinite);
    KSM_LAUNCH_CHECK("set_flag_q_kernel");
    ksm::launch(chan_absmax_kernel, C, 256, 0, s, X, ldx, T, amax, nonfinite);
    KSM_LAUNCH_CHECK("chan_absmax_kernel");
    ksm::launch(row_scales_kernel, (L.ldr + 255) / 256, 256, 0, s, amax, W, C, P, nt, L.ldr, srow);
    KSM_LAUNCH_CHECK("row_scales_kernel");
    {
        ProfileScope prof("conv_pc", kProfProject, s);
        dim3 grid(L.Tp / kCvT, nblk);
        prof.launch(conv_pc_fp8_kernel<6, 61>, grid, dim3(256), 0u, X, ldx, W, C, T, L.Tp, L.ldr,
                    static_cast<const float *>(srow), Bq, bound ? stats : nullptr);
    }
    KSM_LAUNCH_CHECK("conv_pc_fp8_kernel");
    if (bound) {
        // Not a worst-case bound: e4m3 keeps 4 significant bits, and the
        // worst case 2^-3 sum |U||Bc| would leave nearly every column
        // undecided.  The relative rounding errors d (|d| <= 2^-4, variance
        // <= 2^-8 / 3 each, independent, zero mean) give a score error of
        // variance <= (2/3) 2^-8 sum_r U_r^2 Bc_r^2 <= (2/3) 2^-8 sum_b E_b
        // max_b Bc^2 (ksm_bound.h); the margin is kFp8Sigmas standard
        // deviations of that, plus the worst-case accumulation and subnormal
        // terms.
        ksm::launch(fp8_abs_terms_kernel, 1, 256, 0, s, srow, L.ldr, abs_terms);
        KSM_LAUNCH_CHECK("fp8_abs_terms_kernel");
        const double e = 0x1p-4;
        const int depth = L.ldr;
        const double g = depth * 0x1p-24 / (1.0 - depth * 0x1p-24);
        const double zc = kFp8Sigmas * e * std::sqrt(2.0 / 3.0) * (1.0 + 0x1p-10);
        const int rc = bound_finish(stats, 2 * nblk, L.Tp, U, W, perm, C, T, K, P, nt, kBoundSqrt,
                                    0.0, zc, g * (1 + e) * (1 + e), abs_terms, bound, s);
        if (rc != KSM_OK) return rc;
    }
    ksm::launch(u_fp8_kernel, L.Kp, 256, 0, s, U, perm, K, P, C, L.ldr, srow, Uq, uscale);
    KSM_LAUNCH_CHECK("u_fp8_kernel");
    ksm::launch(tile_steps_fp8_kernel, L.nTm, 256, 0, s, Uq, L.ldr, L.nk, W, P * nt, nonfinite,
                steps, counts);
    KSM_LAUNCH_CHECK("tile_steps_fp8_kernel");
    {
        ProfileScope prof("gemm_fp8", kProfProject, s);
        prof.launch(gemm_fp8_mfma, dim3(L.nTm * L.nTn), dim3(256), 0u,
                    static_cast<const fp8_t *>(Uq), static_cast<const fp8_t *>(Bq), L.ldr,
                    static_cast<const float *>(uscale), static_cast<const int *>(steps),
                    static_cast<const int *>(counts), Bt, ldbt, K, L.nTm, L.nTn);
    }
    KSM_LAUNCH_CHECK("gemm_fp8_mfma");
    return KSM_OK;
}

}  // namespace ksm
