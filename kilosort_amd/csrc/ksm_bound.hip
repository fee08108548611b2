// Error bound of the reduced-precision projection forms, per template and
// column, for the peel's decisions (no reference counterpart: the reference
// scores templates in FP32 only, template_matching.py:180-181).
//
// Scores per unit template norm: beta_r(t) = B[r, t] / sqrt(nm_r), so that
// sqrt(Cf) = relu(beta).  With the channels cut into sub-blocks b of 32 and
// E_rb = sum over the sub-block's (channel, PC) rows of U[r]^2:
//   BF16 forms (a proof): |B~ - B| <= kappa sum_rows |U||Bc|
//       <= kappa sum_b sqrt(E_rb) ||Bc_b(t)||   (Cauchy-Schwarz per sub-block),
//   FP8 (a model): the rounding error has variance <= (2/3) 2^-8 sum U^2 Bc^2
//       <= (2/3) 2^-8 sum_b E_rb max_b Bc(t)^2, and the margin is zc times
//       its standard deviation.
// The buffer also carries each template's significant channels and its
// waveforms on them, T_r[c][k] = sum_p U[r,p,c] W[p,k], from which the peel
// re-scores columns in FP32.
#include <algorithm>
#include <cmath>

#include "ksm_bound.h"
#include "ksm_graph.h"

namespace {

using ksm::BoundLayout;
using ksm::kBoundSub;

__global__ void bound_header_kernel(int *__restrict__ hdr, int mode, int nsb) {
    hdr[0] = mode;
    hdr[1] = nsb;
}

// Per template position r: Wt[b][r], the significant channel range and the
// waveform rows.  Significant channels: every channel whose energy
// sum_p U[r,p,c]^2 is at least 2^-70 of the template's largest, closed into
// one range; the channels left out carry at most C 2^-70 of the template's
// energy, so a re-score without them differs by less than 2^-30 ||X window||.
__global__ __launch_bounds__(256) void bound_weights_kernel(
    const float *__restrict__ U, const float *__restrict__ W, const int32_t *__restrict__ perm,
    int C, int K, int P, int nt, int nsb, int mode, double kappa, double zc,
    float *__restrict__ Wt, int *__restrict__ srng, float *__restrict__ Ttab) {
    __shared__ double s_e[ksm::kBoundMaxC];
    __shared__ double s_red[4];
    __shared__ int s_lo[4], s_hi[4];
    const int r = blockIdx.x;
    const int id = perm ? perm[r] : r;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float *u = U + (size_t)id * P * C;
    double mx = 0.0;
    for (int c = tid; c < C; c += 256) {
        double e = 0.0;
        for (int p = 0; p < P; ++p) {
            const double v = u[(size_t)p * C + c];
            e += v * v;
        }
        s_e[c] = e;
        mx = e > mx ? e : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(mx, o, 64);
        mx = w > mx ? w : mx;
    }
    if (lane == 0) s_red[wave] = mx;
    __syncthreads();
    mx = s_red[0];
    for (int w = 1; w < 4; ++w) mx = s_red[w] > mx ? s_red[w] : mx;
    const double thr = mx * 0x1p-70;
    int lo = C, hi = 0;
    for (int c = tid; c < C; c += 256)
        if (mx > 0.0 && s_e[c] >= thr) {
            lo = c < lo ? c : lo;
            hi = c + 1 > hi ? c + 1 : hi;
        }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    __syncthreads();
    if (lane == 0) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    __syncthreads();
    lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
    hi = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
    if (hi <= lo) lo = hi = 0;
    // total energy in a fixed order (sub-block sums, then their sum)
    double nm = 0.0;
    for (int b = 0; b < nsb; ++b) {
        double eb = 0.0;
        for (int c = b * kBoundSub; c < min(C, (b + 1) * kBoundSub); ++c) eb += s_e[c];
        nm += eb;
    }
    for (int b = tid; b < nsb; b += 256) {
        double eb = 0.0;
        for (int c = b * kBoundSub; c < min(C, (b + 1) * kBoundSub); ++c) eb += s_e[c];
        double w = 0.0;
        if (nm > 0.0)
            w = mode == ksm::kBoundLinear ? kappa * sqrt(eb / nm) : zc * zc * (eb / nm);
        Wt[(size_t)b * K + r] = (float)(w * (1.0 + 0x1p-20));
    }
    if (tid == 0) {
        srng[2 * r] = lo;
        srng[2 * r + 1] = hi;
    }
    float *tr = Ttab + (size_t)r * C * nt;
    const int n = (hi - lo) * nt;
    for (int e = tid; e < n; e += 256) {
        const int c = lo + e / nt, k = e % nt;
        float v = 0.f;
        for (int p = 0; p < P; ++p) v = __builtin_fmaf(u[(size_t)p * C + c], W[p * nt + k], v);
        tr[e] = v;
    }
}

// Per column t: S[b][t], colA[t] and colD[t] = max_r D_r(t) + colA[t].  A
// block takes 64 columns: wave 0 reads the column statistics and forms
// colA; then the four waves split the templates, each lane (column) forming
// D_r = sum_b Wt[b][r] S[b][t] (summed in b order) with the column's
// statistics in registers and the template's weights read from the block's
// LDS copy of Wt (template-major rows padded to 4 floats, read 16 bytes at a
// time, the same address on every lane), and the waves' maxima are combined
// at the end.
constexpr int kColThreads = 256;
constexpr int kWtTile = 8192;           // floats of Wt staged per tile
constexpr int kRegSub = 16;             // sub-blocks held in registers

__global__ __launch_bounds__(kColThreads) void bound_columns_kernel(
    const float *__restrict__ stats, int nsb2, int ldst, int T, int K, int nsb, int mode,
    double acc, const float *__restrict__ abs_terms, const float *__restrict__ Wt,
    float *__restrict__ S, float *__restrict__ colA, float *__restrict__ colD) {
    extern __shared__ float4 b_smem4[];
    const int nsbp = (nsb + 3) & ~3;          // padded row of one template's weights
    float *s_v = reinterpret_cast<float *>(b_smem4);            // [nsb][64]
    float *s_w = reinterpret_cast<float *>(b_smem4) + nsbp * 64; // [tile][nsbp]
    __shared__ float s_dm[kColThreads / 64][64];
    __shared__ int s_bad[64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int t = blockIdx.x * 64 + lane;
    const bool live = t < T;
    const int tt = live ? t : T - 1;
    double a = 0.0;
    if (wave == 0) {
        double s0 = 0.0, s2 = 0.0;
        bool bad = false;
        auto sub = [&](int b, float q, float m, float q2) {
            bad |= !__builtin_isfinite(q) || !__builtin_isfinite(m);
            s0 += q;
            if (mode == ksm::kBoundSqrt) {
                bad |= !__builtin_isfinite(q2);
                s2 += q2;
            }
            // sqrt rounded up by one part in 2^20 (q carries its own float rounding)
            const float v = mode == ksm::kBoundLinear ? sqrtf(q) * (1.f + 0x1p-20f) : m;
            s_v[b * 64 + lane] = v;
            if (live) S[(size_t)b * T + t] = bad ? __builtin_inff() : v;
        };
        auto ld = [&](int b, int k) { return stats[((size_t)k * nsb2 + b) * ldst + tt]; };
        const bool sq = mode == ksm::kBoundSqrt;
        // the first kRegSub sub-blocks' loads issued together, then the rest
        float qv[kRegSub], mv[kRegSub], q2v[kRegSub];
#pragma unroll
        for (int b = 0; b < kRegSub; ++b) {
            qv[b] = b < nsb ? ld(b, 0) : 0.f;
            mv[b] = b < nsb ? ld(b, 1) : 0.f;
            q2v[b] = b < nsb && sq ? ld(b, 2) : 0.f;
        }
#pragma unroll
        for (int b = 0; b < kRegSub; ++b)
            if (b < nsb) sub(b, qv[b], mv[b], q2v[b]);
        for (int b = kRegSub; b < nsb; ++b) sub(b, ld(b, 0), ld(b, 1), sq ? ld(b, 2) : 0.f);
        const double up = 1.0 + 0x1p-18;
        a = (0x1p-14 + acc) * sqrt(s0 * up);
        if (mode == ksm::kBoundSqrt && abs_terms)
            a += (double)abs_terms[0] + (double)abs_terms[1] * sqrt(s2 * up);
        a *= up;
        s_bad[lane] = bad ? 1 : 0;
    }
    __syncthreads();
    float sv[kRegSub];
#pragma unroll
    for (int b = 0; b < kRegSub; ++b) sv[b] = b < nsb ? s_v[b * 64 + lane] : 0.f;
    const int tile = max(1, min(K, kWtTile / nsbp));
    float dmax = 0.f;
    for (int r0 = 0; r0 < K; r0 += tile) {
        const int n = min(tile, K - r0);
        __syncthreads();                      // the previous tile is consumed
        // 16 loads per thread in flight, then their stores (pads stay 0)
        for (int e0 = tid; e0 < n * nsbp; e0 += 16 * kColThreads) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = e0 + u * kColThreads;
                const int bb = e / n, rr = e % n;   // coalesced over templates
                v[u] = e < n * nsbp && bb < nsb ? Wt[(size_t)bb * K + r0 + rr] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = e0 + u * kColThreads;
                if (e < n * nsbp) s_w[(e % n) * nsbp + e / n] = v[u];
            }
        }
        __syncthreads();
        for (int rr = wave; rr < n; rr += kColThreads / 64) {
            const float4 *w4 = reinterpret_cast<const float4 *>(s_w + rr * nsbp);
            float d = 0.f;
            if (nsb <= kRegSub) {
#pragma unroll
                for (int b4 = 0; b4 < kRegSub / 4; ++b4) {
                    if (4 * b4 >= nsb) break;
                    const float4 w = w4[b4];
                    const float wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (4 * b4 + u < nsb) d = __builtin_fmaf(wv[u], sv[4 * b4 + u], d);
                }
            } else {
                for (int b = 0; b < nsb; ++b)
                    d = __builtin_fmaf(s_w[rr * nsbp + b], s_v[b * 64 + lane], d);
            }
            dmax = fmaxf(dmax, d);
        }
    }
    s_dm[wave][lane] = dmax;
    __syncthreads();
    if (wave != 0 || !live) return;
#pragma unroll
    for (int w = 1; w < kColThreads / 64; ++w) dmax = fmaxf(dmax, s_dm[w][lane]);
    const double up = 1.0 + 0x1p-18;
    // float sums of nsb nonnegative products: relative rounding < nsb 2^-24
    double dm = (double)dmax * (1.0 + (nsb + 2) * 0x1p-24);
    if (mode == ksm::kBoundSqrt) dm = sqrt(dm);
    const float inf = __builtin_inff();
    const bool bad = s_bad[lane] != 0;
    colA[t] = bad ? inf : (float)a * (1.f + 0x1p-20f);
    colD[t] = bad ? inf : (float)((dm + a) * up) * (1.f + 0x1p-20f);
}

}  // namespace

namespace ksm {

int bound_finish(const float *stats, int nsb2, int ldst, const float *U, const float *W,
                 const int32_t *perm, int C, int T, int K, int P, int nt, int mode, double kappa,
                 double zc, double acc, const float *abs_terms, void *bound, hipStream_t s) {
    if (C > kBoundMaxC)
        return fail(KSM_ERR_INVALID, "ksm_project_bounded: at most %d channels", kBoundMaxC);
    const BoundLayout L = bound_layout(C, T, K, nt);
    if (nsb2 < L.nsb) return fail(KSM_ERR_INVALID, "ksm_project_bounded: statistics layout");
    char *b = static_cast<char *>(bound);
    float *colD = reinterpret_cast<float *>(b + L.o_colD);
    float *colA = reinterpret_cast<float *>(b + L.o_colA);
    float *S = reinterpret_cast<float *>(b + L.o_S);
    float *Wt = reinterpret_cast<float *>(b + L.o_Wt);
    int *srng = reinterpret_cast<int *>(b + L.o_srng);
    float *Ttab = reinterpret_cast<float *>(b + L.o_T);
    ksm::launch(bound_header_kernel, 1, 1, 0, s, reinterpret_cast<int *>(b + L.o_hdr), mode, L.nsb);
    KSM_LAUNCH_CHECK("bound_header_kernel");
    ksm::launch(bound_weights_kernel, K, 256, 0, s, U, W, perm, C, K, P, nt, L.nsb, mode, kappa,
                zc, Wt, srng, Ttab);
    KSM_LAUNCH_CHECK("bound_weights_kernel");
    const int nsbp = (L.nsb + 3) & ~3;
    const int tile = std::max(1, std::min(K, kWtTile / nsbp));
    const unsigned col_lds = (unsigned)((nsbp * 64 + tile * nsbp) * sizeof(float));
    ksm::launch(bound_columns_kernel, (T + 63) / 64, kColThreads, col_lds, s, stats, nsb2, ldst,
                T, K, L.nsb, mode, acc, abs_terms, Wt, S, colA, colD);
    KSM_LAUNCH_CHECK("bound_columns_kernel");
    return KSM_OK;
}

}  // namespace ksm

extern "C" size_t ksm_bound_bytes(int C, int T, int K, int P, int nt) {
    (void)P;
    if (C <= 0 || T <= 0 || K <= 0 || nt <= 0) return 0;
    return ksm::bound_layout(C, T, K, nt).bytes;
}
