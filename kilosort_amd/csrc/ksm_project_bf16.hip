// BF16 forms of the projection B = U * (W (*) X) on v_mfma_f32_32x32x16_bf16
// with FP32 accumulation.
//
// Reference: kilosort/template_matching.py:180-181 (computed there in FP32).
//
// The temporal-PC correlation is computed in FP32 (VALU) and stored as
// BF16, time-major and channel-major: Bh[t][c*P + p] (and the rounding
// remainder Bl = bf16(x - Bh) for the split form), so that both GEMM operands
// are contiguous along the contraction index r = c*P + p.  U is stored the
// same way per template position: Uh[m][c*P + p] = bf16(U[perm[m], p, c]).
//
//   KSM_PREC_BF16     B ~ Uh.Bh                       (one product)
//   KSM_PREC_BF16X3   B ~ Uh.Bh + Uh.Bl + Ul.Bh       (split operands; the
//                     dropped Ul.Bl term is ~2^-16 relative)
// The three products of the split form are one GEMM over a contraction three
// times as deep: [Uh | Uh | Ul] . [Bh | Bl | Bh].
#include <cmath>

#include "ksm_common.h"
#include "ksm_profile.h"
#include "ksm_project_internal.h"
#include "ksm_bound.h"
#include "ksm_graph.h"

namespace {

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ bf16_t to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

constexpr int kCvCh = 64;     // channels per correlation block (one per lane)
constexpr int kCvTT = 8;      // time samples per thread
constexpr int kCvT = 4 * kCvTT;  // 32 time samples per block (4 waves)

// Bh[t][c*P + p] (and Bl) for t < Tp, c < Cp (zero for t >= T or c >= C).
//
// stats (may be NULL): per 32-channel sub-block b (two per block) and column
// t, over the sub-block's (channel, PC) rows, the FP32 correlation's sum of
// squares at stats[(0 * nsb2 + b) * Tp + t] and its largest square at
// stats[(1 * nsb2 + b) * Tp + t] (inputs of the error bound, ksm_bound.h).
template <int P, int NTAP, bool SPLIT>
__global__ __launch_bounds__(256) void conv_pc_bf16_kernel(
    const float *__restrict__ X, int ldx, const float *__restrict__ W, int C, int T, int Tp,
    int ldr, bf16_t *__restrict__ Bh, bf16_t *__restrict__ Bl, int *__restrict__ nonfinite,
    float *__restrict__ stats) {
    constexpr int H = NTAP / 2;
    constexpr int SPAN = kCvT + NTAP - 1;
    constexpr int LD = SPAN | 1;                 // odd row stride: no bank conflicts
    __shared__ float xs[kCvCh * LD];
    const int t0 = blockIdx.x * kCvT;
    const int c0 = blockIdx.y * kCvCh;
    bool bad = false;
    for (int e = threadIdx.x; e < kCvCh * SPAN; e += 256) {
        const int cl = e / SPAN, j = e % SPAN;
        const int c = c0 + cl, t = t0 - H + j;
        const float v = (c < C && t >= 0 && t < T) ? X[(size_t)c * ldx + t] : 0.f;
        bad |= !__builtin_isfinite(v);
        xs[cl * LD + j] = v;
    }
    if (__builtin_amdgcn_ballot_w64(bad) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(nonfinite, 1);
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
    if (stats) {
        // per 32-channel half of the block (lanes 0-31, 32-63): sum and largest
        // square of the correlation values of each column
        const int nsb2 = 2 * gridDim.y, sb = 2 * blockIdx.y + (cl >> 5);
#pragma unroll
        for (int u = 0; u < kCvTT; ++u) {
            const int t = t0 + tg * kCvTT + u;
            const bool live = t < T && c < C;
            float sq = 0.f, mx = 0.f;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float a = live ? acc[u][p] : 0.f;
                sq = __builtin_fmaf(a, a, sq);
                mx = fmaxf(mx, a * a);
                if (a != a) mx = a;             // NaN reaches the bound
            }
            sq = ksm::half_sum_f(sq);
            mx = ksm::half_max_nan(mx);
            if ((cl & 31) == 0 && t < Tp) {
                stats[((size_t)0 * nsb2 + sb) * Tp + t] = sq;
                stats[((size_t)1 * nsb2 + sb) * Tp + t] = mx;
            }
        }
    }
    if (c * P >= ldr) return;
    static_assert(P % 2 == 0, "pairs of bf16 are stored as 32-bit words");
    // the P values of (t, c) are 2P consecutive bytes at a 4-byte aligned
    // offset (c P 2 bytes, P even): stored as P/2 32-bit words in one go
#pragma unroll
    for (int u = 0; u < kCvTT; ++u) {
        const int t = t0 + tg * kCvTT + u;
        if (t >= Tp) break;
        const bool live = t < T && c < C;
        unsigned wh[P / 2], wl[P / 2];
#pragma unroll
        for (int p = 0; p < P; p += 2) {
            const float x0 = live ? acc[u][p] : 0.f, x1 = live ? acc[u][p + 1] : 0.f;
            const bf16_t h0 = to_bf16(x0), h1 = to_bf16(x1);
            wh[p / 2] = (unsigned)h0 | ((unsigned)h1 << 16);
            if (SPLIT) {
                const bf16_t l0 = to_bf16(x0 - from_bf16(h0)), l1 = to_bf16(x1 - from_bf16(h1));
                wl[p / 2] = (unsigned)l0 | ((unsigned)l1 << 16);
            }
        }
        if constexpr (P == 6) {
            *reinterpret_cast<uint3 *>(Bh + (size_t)t * ldr + c * P) =
                make_uint3(wh[0], wh[1], wh[2]);
            if (SPLIT)
                *reinterpret_cast<uint3 *>(Bl + (size_t)t * ldr + c * P) =
                    make_uint3(wl[0], wl[1], wl[2]);
        } else {
            unsigned *dh = reinterpret_cast<unsigned *>(Bh + (size_t)t * ldr + c * P);
#pragma unroll
            for (int k = 0; k < P / 2; ++k) dh[k] = wh[k];
            if (SPLIT) {
                unsigned *dl = reinterpret_cast<unsigned *>(Bl + (size_t)t * ldr + c * P);
#pragma unroll
                for (int k = 0; k < P / 2; ++k) dl[k] = wl[k];
            }
        }
    }
}

// Uh[m][c*P + p] = bf16(U[perm[m], p, c]) (Ul = remainder), zero for m >= K or
// c >= C.
__global__ void u_bf16_kernel(const float *__restrict__ U, const int32_t *__restrict__ perm,
                              int K, int P, int C, int Kp, int ldr, bf16_t *__restrict__ Uh,
                              bf16_t *__restrict__ Ul) {
    const size_t n = (size_t)Kp * ldr;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(e / ldr), r = (int)(e % ldr);
        const int c = r / P, p = r % P;
        float x = 0.f;
        if (m < K && c < C) x = U[((size_t)(perm ? perm[m] : m) * P + p) * C + c];
        const bf16_t h = to_bf16(x);
        Uh[e] = h;
        if (Ul) Ul[e] = to_bf16(x - from_bf16(h));
    }
}

// ---------------------------------------------------------------------------
// TN GEMM: Bout[n][m] = sum_r A[m][r] * B[n][r] over nseg contraction segments
// (segment s uses A = As_[s], B = Bs_[s]).  Block tile 128 (m) x 128 (n),
// depth step 64; 4 waves each own 64x64 as 2x2 tiles of 32x32x16 MFMAs.
// LDS rows are padded by 16 bytes so that the 16-byte operand reads of 32
// lanes fall in distinct banks.

constexpr int HBM_ = 128, HBN = 128, HBK = 64;
constexpr int HLD = HBK + 8;      // padded LDS row (bf16 elements)

struct Segs {
    const bf16_t *a[3];
    const bf16_t *b[3];
};

__global__ void set_flag_h_kernel(int *__restrict__ p) { *p = 0; }

// The depth steps (HBK columns of r = c*P + p) where a tile of MT template
// rows of Uh is not all zero, in increasing order: steps[tm][0 .. counts[tm]).
// A skipped step would add only exact zeros (templates are zero outside their
// channel footprint); with a non-finite value in X, W or U every step is kept
// (0 * inf must give NaN as in the dense product).  Ul is zero wherever Uh is.
constexpr int kHStepsMax = 4096;
template <int MT>
__global__ __launch_bounds__(256) void tile_steps_bf16_kernel(
    const bf16_t *__restrict__ Uh, int ldr, int nk, const float *__restrict__ W, int PW,
    const int *__restrict__ nonfinite, int *__restrict__ steps, int *__restrict__ counts) {
    __shared__ unsigned char s_need[kHStepsMax];
    __shared__ int s_dense;
    const int tm = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_dense = *nonfinite != 0 || nk > kHStepsMax;
    __syncthreads();
    bool wbad = false;
    for (int e = tid; e < PW; e += 256) wbad |= !__builtin_isfinite(W[e]);
    if (wbad) s_dense = 1;                 // benign race: any writer stores 1
    __syncthreads();
    const bool dense = s_dense != 0;
    if (!dense) {
        // lane: rows lane / 8 + 8 i of the tile, 16-byte chunk lane % 8 of the step
        const int rsub = lane >> 3, chunk = lane & 7;
        for (int st = wave; st < nk; st += 4) {
            bool nz = false;
            for (int i = 0; i < MT / 8; ++i) {
                const uint4 v = *reinterpret_cast<const uint4 *>(
                    Uh + (size_t)(tm * MT + rsub + 8 * i) * ldr + st * HBK + chunk * 8);
                // bf16 +0 and -0 are zero; any other pattern (NaN included) is not
                const unsigned m = 0x7fff7fffu;
                nz |= ((v.x & m) | (v.y & m) | (v.z & m) | (v.w & m)) != 0u;
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

__device__ __forceinline__ int xcd_remap_h(int bid, int nb) {
    const int xcd = bid & 7, pos = bid >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// MT template rows per block tile (128: 4 waves of 64 x 64; 64: 4 waves of
// 32 x 64), 128 time columns.
template <int MT>
__global__ __launch_bounds__(256, 2) void gemm_bf16_mfma(Segs sg, int nseg, int ldr,
                                                      const int *__restrict__ steps,
                                                      const int *__restrict__ counts,
                                                      float *__restrict__ Bout, int ldb, int K,
                                                      int nTm, int nTn) {
    constexpr int MI = MT / 64;                  // 32-row MFMA tiles per wave
    constexpr int AE = MT * HBK / 256;           // A elements each thread loads per step
    __shared__ bf16_t As[2][MT * HLD];
    __shared__ bf16_t Bs[2][HBN * HLD];
    const int logical = xcd_remap_h(blockIdx.x, nTm * nTn);
    const int tm = logical % nTm, tn = logical / nTm;
    const int m0 = tm * MT, n0 = tn * HBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // global -> LDS: B: thread owns row tid/2, 32 consecutive elements (4 x 16
    // B); A: row tid / (HBK / AE), AE consecutive elements
    const int grow = tid >> 1, gcol = (tid & 1) * 32;
    const int arow = tid / (HBK / AE), acol = (tid % (HBK / AE)) * AE;

    f32x16 acc[MI][2];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    const int nk_all = ldr / HBK;
    const int nk_seg = counts[tm];              // the tile's steps in each segment
    const int *lst = steps + (size_t)tm * nk_all;
    const int nk = nk_seg * nseg;
    // k step kt reads segment kt / nk_seg at depth lst[kt % nk_seg] * HBK
    const int k00 = nk_seg > 0 ? lst[0] * HBK : 0;
    const bf16_t *ap = sg.a[0] + (size_t)(m0 + arow) * ldr + k00 + acol;
    const bf16_t *bp = sg.b[0] + (size_t)(n0 + grow) * ldr + k00 + gcol;
    // the A registers as named variables (an array here was left in scratch
    // memory by the compiler): AE / 8 = 2 or 4 of them
    static_assert(AE / 8 == 2 || AE / 8 == 4, "A loads per thread");
    uint4 ra0, ra1, ra2, ra3;
    auto load_a = [&](const uint4 *p) {
        ra0 = p[0];
        ra1 = p[1];
        if constexpr (AE / 8 == 4) {
            ra2 = p[2];
            ra3 = p[3];
        }
    };
    auto store_a = [&](uint4 *la) {
        la[0] = ra0;
        la[1] = ra1;
        if constexpr (AE / 8 == 4) {
            la[2] = ra2;
            la[3] = ra3;
        }
    };
    load_a(reinterpret_cast<const uint4 *>(ap));
    uint4 rb0 = reinterpret_cast<const uint4 *>(bp)[0], rb1 = reinterpret_cast<const uint4 *>(bp)[1];
    uint4 rb2 = reinterpret_cast<const uint4 *>(bp)[2], rb3 = reinterpret_cast<const uint4 *>(bp)[3];
    {
        uint4 *la = reinterpret_cast<uint4 *>(&As[0][arow * HLD + acol]);
        uint4 *lb = reinterpret_cast<uint4 *>(&Bs[0][grow * HLD + gcol]);
        store_a(la);
        lb[0] = rb0; lb[1] = rb1; lb[2] = rb2; lb[3] = rb3;
    }
    __syncthreads();
    const int kr = (lane >> 5) * 8, kc = lane & 31;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            const int s = (kt + 1) / nk_seg, k0 = lst[(kt + 1) % nk_seg] * HBK;
            const bf16_t *sa = s == 0 ? sg.a[0] : (s == 1 ? sg.a[1] : sg.a[2]);
            const bf16_t *sb = s == 0 ? sg.b[0] : (s == 1 ? sg.b[1] : sg.b[2]);
            const uint4 *pa = reinterpret_cast<const uint4 *>(sa + (size_t)(m0 + arow) * ldr + k0 + acol);
            const uint4 *pb = reinterpret_cast<const uint4 *>(sb + (size_t)(n0 + grow) * ldr + k0 + gcol);
            load_a(pa);
            rb0 = pb[0]; rb1 = pb[1]; rb2 = pb[2]; rb3 = pb[3];
        }
#pragma unroll
        for (int kk = 0; kk < HBK; kk += 16) {
            bf16x8 a[MI], b[2];
#pragma unroll
            for (int i = 0; i < MI; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(
                    &As[buf][(wm * (MT / 2) + i * 32 + kc) * HLD + kk + kr]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(
                    &Bs[buf][(wn * 64 + j * 32 + kc) * HLD + kk + kr]);
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0,
                                                                        0, 0);
        }
        if (more) {
            uint4 *la = reinterpret_cast<uint4 *>(&As[buf ^ 1][arow * HLD + acol]);
            uint4 *lb = reinterpret_cast<uint4 *>(&Bs[buf ^ 1][grow * HLD + gcol]);
            store_a(la);
            lb[0] = rb0; lb[1] = rb1; lb[2] = rb2; lb[3] = rb3;
        }
        __syncthreads();
    }
    // transposed store Bout[n][m]: registers 4g..4g+3 are 4 consecutive m
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int m = m0 + wm * (MT / 2) + i * 32 + 8 * g + 4 * (lane >> 5);
                const int n = n0 + wn * 64 + j * 32 + (lane & 31);
                float *dst = Bout + (size_t)n * ldb + m;
                if (m + 3 < K) {
                    float4 v;
                    v.x = acc[i][j][4 * g + 0];
                    v.y = acc[i][j][4 * g + 1];
                    v.z = acc[i][j][4 * g + 2];
                    v.w = acc[i][j][4 * g + 3];
                    *reinterpret_cast<float4 *>(dst) = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (m + e < K) dst[e] = acc[i][j][4 * g + e];
                }
            }
}

struct HLayout {
    int ldr, Kp, Tp;
    size_t bc_elems, u_elems;
};

HLayout h_layout(int C, int T, int K, int P) {
    HLayout L;
    L.ldr = (int)ksm::align_up((size_t)P * C, 192);   // multiple of P=6 and of 64
    if (L.ldr % P) L.ldr = (int)ksm::align_up((size_t)P * C, (size_t)P * HBK);
    L.Kp = (int)ksm::align_up(K, HBM_);
    L.Tp = (int)ksm::align_up(T, HBN);
    L.bc_elems = (size_t)L.Tp * L.ldr;
    L.u_elems = (size_t)L.Kp * L.ldr;
    return L;
}

}  // namespace

namespace ksm {

size_t project_bf16_workspace(int C, int T, int K, int P, int precision) {
    const HLayout L = h_layout(C, T, K, P);
    const int terms = precision == KSM_PREC_BF16X3 ? 2 : 1;
    const size_t nTm = L.Kp / 64, nk = L.ldr / HBK;     // the finer (64-row) tiling
    const size_t nblk = (L.ldr / P + kCvCh - 1) / kCvCh;
    return terms * (align_up(L.bc_elems * 2, 256) + align_up(L.u_elems * 2, 256)) +
           align_up(nTm * nk * sizeof(int), 256) + align_up(nTm * sizeof(int), 256) +
           align_up(2 * 2 * nblk * L.Tp * sizeof(float), 256) + 256 + 512;
}

int project_bf16(const float *X, int ldx, const float *W, const float *U, const int32_t *perm,
                 int C, int T, int K, int P, int nt, float *Bt, int ldbt, int precision,
                 int tile_m, void *bound, void *workspace, hipStream_t s) {
    if (P != 6 || nt != 61)
        return fail(KSM_ERR_INVALID, "ksm_project: BF16 forms are built for P=6, nt=61");
    const HLayout L = h_layout(C, T, K, P);
    const bool split = precision == KSM_PREC_BF16X3;
    Carver cv(workspace);
    bf16_t *Bh = cv.take<bf16_t>(L.bc_elems);
    bf16_t *Uh = cv.take<bf16_t>(L.u_elems);
    bf16_t *Bl = split ? cv.take<bf16_t>(L.bc_elems) : nullptr;
    bf16_t *Ul = split ? cv.take<bf16_t>(L.u_elems) : nullptr;
    // template rows per GEMM tile: 64 or 128 (Kp is a multiple of both)
    const int MT = tile_m == 64 ? 64 : 128;
    const int nTm = (int)align_up(K, MT) / MT, nTn = L.Tp / HBN, nk = L.ldr / HBK;
    int *steps = cv.take<int>((size_t)nTm * nk);
    int *counts = cv.take<int>(nTm);
    int *nonfinite = cv.take<int>(1);
    const int nblk = (L.ldr / P + kCvCh - 1) / kCvCh;
    float *stats = cv.take<float>((size_t)2 * 2 * nblk * L.Tp);
    ksm::launch(set_flag_h_kernel, 1, 1, 0, s, nonfinite);
    KSM_LAUNCH_CHECK("set_flag_h_kernel");
    {
        ProfileScope prof("conv_pc", kProfProject, s);
        dim3 grid(L.Tp / kCvT, nblk);
        float *st_out = bound ? stats : nullptr;
        if (split)
            prof.launch(conv_pc_bf16_kernel<6, 61, true>, grid, dim3(256), 0u, X, ldx, W, C, T,
                        L.Tp, L.ldr, Bh, Bl, nonfinite, st_out);
        else
            prof.launch(conv_pc_bf16_kernel<6, 61, false>, grid, dim3(256), 0u, X, ldx, W, C, T,
                        L.Tp, L.ldr, Bh, Bl, nonfinite, st_out);
    }
    KSM_LAUNCH_CHECK("conv_pc_bf16_kernel");
    if (bound) {
        // |B~ - B| <= kappa sum_r |U_r| |Bc_r| with kappa = 2u + u^2 + gamma_n for
        // one BF16 product (u = 2^-8 per rounded operand), or the dropped
        // lo * lo term and the split remainders (3 u^2 (1 + u)^3) for the split
        // form, plus FP32 accumulation over the n products (gamma_n)
        const double e = std::ldexp(1.0, -8);
        const int depth = (split ? 3 : 1) * L.ldr;
        const double g = depth * std::ldexp(1.0, -24) / (1.0 - depth * std::ldexp(1.0, -24));
        const double kappa = split ? 3.0 * e * e * (1 + e) * (1 + e) * (1 + e) + g * (1 + e) * (1 + e)
                                   : (2 * e + e * e) + g * (1 + e) * (1 + e);
        const int rc = bound_finish(stats, 2 * nblk, L.Tp, U, W, perm, C, T, K, P, nt,
                                    kBoundLinear, kappa, 0.0, 0.0, nullptr, bound, s);
        if (rc != KSM_OK) return rc;
    }
    ksm::launch(u_bf16_kernel, 512, 256, 0, s, U, perm, K, P, C, L.Kp, L.ldr, Uh, Ul);
    KSM_LAUNCH_CHECK("u_bf16_kernel");
    if (MT == 64)
        ksm::launch(tile_steps_bf16_kernel<64>, nTm, 256, 0, s, Uh, L.ldr, nk, W, P * nt,
                    nonfinite, steps, counts);
    else
        ksm::launch(tile_steps_bf16_kernel<128>, nTm, 256, 0, s, Uh, L.ldr, nk, W, P * nt,
                    nonfinite, steps, counts);
    KSM_LAUNCH_CHECK("tile_steps_bf16_kernel");
    Segs sg;
    sg.a[0] = Uh;
    sg.b[0] = Bh;
    sg.a[1] = Uh;
    sg.b[1] = split ? Bl : Bh;
    sg.a[2] = split ? Ul : Uh;
    sg.b[2] = Bh;
    {
        ProfileScope prof("gemm_bf16", kProfProject, s);
        if (MT == 64)
            prof.launch(gemm_bf16_mfma<64>, dim3(nTm * nTn), dim3(256), 0u, sg, split ? 3 : 1,
                        L.ldr, static_cast<const int *>(steps), static_cast<const int *>(counts),
                        Bt, ldbt, K, nTm, nTn);
        else
            prof.launch(gemm_bf16_mfma<128>, dim3(nTm * nTn), dim3(256), 0u, sg, split ? 3 : 1,
                        L.ldr, static_cast<const int *>(steps), static_cast<const int *>(counts),
                        Bt, ldbt, K, nTm, nTn);
    }
    KSM_LAUNCH_CHECK("gemm_bf16_mfma");
    return KSM_OK;
}

}  // namespace ksm
