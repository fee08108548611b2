// Per-template set-up of the matching path: norms, peak channels, the
// float64 template Gram (ctc) and its nonzero-row lists.
//
// Reference: kilosort/template_matching.py:14-53 (prepare_extract),
// :155-167 (prepare_matching), :176 (template norms).
#include "ksm_common.h"

namespace {

__device__ inline double wave_sum_f64(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// nm[i] = sum_p sum_c U[i,p,c]^2; one wave per template, float64 sums.
__global__ __launch_bounds__(256) void norms_kernel(const float *__restrict__ U, int K,
                                                    int PC, float *__restrict__ nm) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= K) return;
    const float *u = U + (size_t)wave * PC;
    double acc = 0.0;
    for (int e = lane; e < PC; e += 64) {
        double x = (double)u[e];
        acc = fma(x, x, acc);
    }
    acc = wave_sum_f64(acc);
    if (lane == 0) nm[wave] = (float)acc;
}

// iU[i] = argmax_c sum_p U[i,p,c]^2 (float32 sum over p in order, first index
// on ties) and Ucc[n,i,p] = U[i,p,iCC[n,iU[i]]].
struct PeakKey {
    float best;
    int arg;                              // 0x7fffffff: no channel seen
    bool nan;
};

// The sequential rule (first NaN if any, else the first maximum) as an
// associative choice between two partial results.
__device__ __forceinline__ PeakKey peak_pick(PeakKey a, PeakKey b) {
    if (b.arg == 0x7fffffff) return a;
    if (a.arg == 0x7fffffff) return b;
    if (a.nan || b.nan) {
        if (a.nan && b.nan) return a.arg < b.arg ? a : b;
        return a.nan ? a : b;
    }
    if (b.best > a.best || (b.best == a.best && b.arg < a.arg)) return b;
    return a;
}

// One wave per template: lanes take channels (coalesced loads), each keeps the
// sequential rule over its channels, then a wave reduction.
__global__ __launch_bounds__(64) void peak_channel_kernel(const float *__restrict__ U, int K,
                                                         int P, int C,
                                                         const int64_t *__restrict__ iCC, int nC,
                                                         int64_t *__restrict__ iU,
                                                         float *__restrict__ Ucc) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const float *u = U + (size_t)i * P * C;
    PeakKey k{0.f, 0x7fffffff, false};
    for (int c = lane; c < C; c += 64) {
        float s = 0.f;
        for (int p = 0; p < P; ++p) {
            const float v = u[(size_t)p * C + c];
            s = s + v * v;
        }
        const bool sn = s != s;
        if (k.arg == 0x7fffffff || (!k.nan && (sn || s > k.best))) k = PeakKey{s, c, sn};
    }
    for (int o = 32; o > 0; o >>= 1) {
        PeakKey o2;
        o2.best = __shfl_xor(k.best, o, 64);
        o2.arg = __shfl_xor(k.arg, o, 64);
        o2.nan = __shfl_xor((int)k.nan, o, 64) != 0;
        k = peak_pick(k, o2);
    }
    const int arg = k.arg;
    if (lane == 0) iU[i] = arg;
    for (int e = lane; e < nC * P; e += 64) {
        const int n = e / P, p = e - n * P;
        const int64_t ch = iCC[(size_t)n * C + arg];
        Ucc[((size_t)n * K + i) * P + p] = u[(size_t)p * C + ch];
    }
}

// WtW[p][q][d+nt] = sum_k W[p,k] W[q,k+d], float64.
__global__ void wtw_kernel(const float *__restrict__ W, int P, int nt,
                           double *__restrict__ wtw) {
    const int L = 2 * nt + 1;
    const int n = P * P * L;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int l = e % L;
        const int pq = e / L;
        const int p = pq / P, q = pq % P;
        const int d = l - nt;
        const int k0 = d < 0 ? -d : 0;
        const int k1 = d > 0 ? nt - d : nt;
        double acc = 0.0;
        for (int k = k0; k < k1; ++k)
            acc = fma((double)W[p * nt + k], (double)W[q * nt + k + d], acc);
        wtw[e] = acc;
    }
}

// Ut[(p*C + c)*K + i] = U[i,p,c]: template index fastest, for coalesced reads.
__global__ void transpose_u_kernel(const float *__restrict__ U, int K, int PC,
                                   float *__restrict__ Ut) {
    const size_t n = (size_t)K * PC;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e % K, r = e / K;
        Ut[e] = U[i * PC + r];
    }
}

// One block per (firing template j, 64 templates i): UtU[i,j,:,:] in float64
// registers, then ctc[j][i][l] = sum_pq UtU[pq] * WtW[pq][l] staged in LDS and
// written as 64 contiguous rows.
constexpr int kCtcChunk = 32;      // lags staged per output step

// ctc[j][i][l] for the 64 templates i of the block: utu[p][q] = the float64
// sum over channels of U[i][p][c] U[j][q][c] in registers, then per lag the
// sum over (p, q) of utu * WtW (the table read with wave-uniform loads), the
// rows staged through LDS 32 lags at a time for coalesced stores.
template <int P>
__global__ __launch_bounds__(64) void ctc_kernel(const float *__restrict__ Ut,
                                                 const float *__restrict__ U,
                                                 const double *__restrict__ wtw,
                                                 int K, int C, int nt, float *__restrict__ ctc,
                                                 int ldl) {
    __shared__ float s_out[64][kCtcChunk + 1];
    const int L = 2 * nt + 1;
    const int j = blockIdx.x;
    const int i0 = blockIdx.y * 64;
    const int lane = threadIdx.x;
    const int i = i0 + lane;

    double utu[P * P];
#pragma unroll
    for (int e = 0; e < P * P; ++e) utu[e] = 0.0;
    const float *uj = U + (size_t)j * P * C;
    if (i < K) {
        for (int c = 0; c < C; ++c) {
            double ui[P], uq[P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                ui[p] = (double)Ut[((size_t)p * C + c) * K + i];
                uq[p] = (double)uj[(size_t)p * C + c];
            }
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int q = 0; q < P; ++q) utu[p * P + q] = fma(ui[p], uq[q], utu[p * P + q]);
        }
    }
    const int nrow = min(64, K - i0);
    float *dst = ctc + ((size_t)j * K + i0) * ldl;
    for (int l0 = 0; l0 < ldl; l0 += kCtcChunk) {
        const int nl = min(kCtcChunk, ldl - l0);
        for (int u = 0; u < nl; ++u) {
            const int l = l0 + u;
            double acc = 0.0;
            if (l < L) {
#pragma unroll
                for (int e = 0; e < P * P; ++e) acc = fma(utu[e], wtw[e * L + l], acc);
            }
            s_out[lane][u] = (float)acc;
        }
        __syncthreads();
        for (int e = lane; e < nrow * nl; e += 64) {
            const int r = e / nl, u = e - r * nl;
            dst[(size_t)r * ldl + l0 + u] = s_out[r][u];
        }
        __syncthreads();
    }
}

// Generic-P fallback of ctc_kernel (runtime P <= 8).
__global__ __launch_bounds__(64) void ctc_kernel_generic(
    const float *__restrict__ Ut, const float *__restrict__ U, const double *__restrict__ wtw,
    int K, int P, int C, int nt, float *__restrict__ ctc, int ldl) {
    __shared__ float s_out[64][kCtcChunk + 1];
    const int L = 2 * nt + 1;
    const int j = blockIdx.x;
    const int i0 = blockIdx.y * 64;
    const int lane = threadIdx.x;
    const int i = i0 + lane;
    double utu[64];
    for (int e = 0; e < 64; ++e) utu[e] = 0.0;
    const float *uj = U + (size_t)j * P * C;
    if (i < K) {
        for (int c = 0; c < C; ++c)
            for (int p = 0; p < P; ++p) {
                const double a = (double)Ut[((size_t)p * C + c) * K + i];
                for (int q = 0; q < P; ++q)
                    utu[p * P + q] = fma(a, (double)uj[(size_t)q * C + c], utu[p * P + q]);
            }
    }
    const int nrow = min(64, K - i0);
    float *dst = ctc + ((size_t)j * K + i0) * ldl;
    for (int l0 = 0; l0 < ldl; l0 += kCtcChunk) {
        const int nl = min(kCtcChunk, ldl - l0);
        for (int u = 0; u < nl; ++u) {
            const int l = l0 + u;
            double acc = 0.0;
            if (l < L)
                for (int e = 0; e < P * P; ++e) acc = fma(utu[e], wtw[e * L + l], acc);
            s_out[lane][u] = (float)acc;
        }
        __syncthreads();
        for (int e = lane; e < nrow * nl; e += 64) {
            const int r = e / nl, u = e - r * nl;
            dst[(size_t)r * ldl + l0 + u] = s_out[r][u];
        }
        __syncthreads();
    }
}

// Nonzero rows of ctc[j][:, :L]: one block of 256 per firing template j.
__global__ __launch_bounds__(256) void ctc_rows_kernel(const float *__restrict__ ctc, int K,
                                                       int L, int ldl,
                                                       int32_t *__restrict__ rows,
                                                       int32_t *__restrict__ nrows) {
    extern __shared__ unsigned char smem[];
    unsigned char *flag = smem;                                  // [K]
    int *s_cnt = reinterpret_cast<int *>(smem + ksm::align_up(K, 16));  // [257]
    const int j = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float *base = ctc + (size_t)j * K * ldl;
    for (int i = wave; i < K; i += 4) {
        const float *row = base + (size_t)i * ldl;
        bool nz = false;
        for (int l = lane; l < L; l += 64) nz |= (row[l] != 0.f);
        const unsigned long long m = __ballot(nz);
        if (lane == 0) flag[i] = m != 0ull;
    }
    __syncthreads();
    // ordered compaction: thread t owns rows [t*seg, (t+1)*seg)
    const int seg = (K + 255) / 256;
    const int r0 = threadIdx.x * seg, r1 = min(K, r0 + seg);
    int cnt = 0;
    for (int r = r0; r < r1; ++r) cnt += flag[r];
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < 256; ++t) {
            const int v = s_cnt[t];
            s_cnt[t] = run;
            run += v;
        }
        s_cnt[256] = run;
    }
    __syncthreads();
    int pos = s_cnt[threadIdx.x];
    int32_t *out = rows + (size_t)j * K;
    for (int r = r0; r < r1; ++r)
        if (flag[r]) out[pos++] = r;
    if (threadIdx.x == 0) nrows[j] = s_cnt[256];
}

}  // namespace

extern "C" {

int ksm_template_norms(const float *U, int K, int P, int C, float *nm, void *stream) {
    if (!U || !nm || K <= 0 || P <= 0 || C <= 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_template_norms: bad arguments");
    const int blocks = (K * 64 + 255) / 256;
    norms_kernel<<<blocks, 256, 0, ksm::as_stream(stream)>>>(U, K, P * C, nm);
    KSM_LAUNCH_CHECK("norms_kernel");
    return KSM_OK;
}

int ksm_prepare_extract(const float *U, int K, int P, int C, const int64_t *iCC, int nC,
                        int64_t *iU, float *Ucc, void *stream) {
    if (!U || !iCC || !iU || !Ucc || K <= 0 || P <= 0 || C <= 0 || nC <= 0 || nC > C)
        return ksm::fail(KSM_ERR_INVALID, "ksm_prepare_extract: bad arguments");
    peak_channel_kernel<<<K, 64, 0, ksm::as_stream(stream)>>>(U, K, P, C, iCC, nC, iU, Ucc);
    KSM_LAUNCH_CHECK("peak_channel_kernel");
    return KSM_OK;
}

size_t ksm_prepare_matching_workspace(int K, int P, int C, int nt) {
    const size_t L = 2 * (size_t)nt + 1;
    return ksm::align_up((size_t)P * P * L * sizeof(double), 256) +
           ksm::align_up((size_t)K * P * C * sizeof(float), 256) + 512;
}

int ksm_prepare_matching(const float *U, const float *W, int K, int P, int C, int nt,
                         float *ctc, int ldl, void *workspace, size_t ws_bytes,
                         void *stream) {
    const int L = 2 * nt + 1;
    if (!U || !W || !ctc || K <= 0 || P <= 0 || P > 8 || C <= 0 || nt <= 0 || ldl < L)
        return ksm::fail(KSM_ERR_INVALID, "ksm_prepare_matching: bad arguments");
    if (!workspace || ws_bytes < ksm_prepare_matching_workspace(K, P, C, nt))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_prepare_matching: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    ksm::Carver cv(workspace);
    double *wtw = cv.take<double>((size_t)P * P * L);
    float *Ut = cv.take<float>((size_t)K * P * C);
    wtw_kernel<<<(P * P * L + 255) / 256, 256, 0, s>>>(W, P, nt, wtw);
    KSM_LAUNCH_CHECK("wtw_kernel");
    transpose_u_kernel<<<1024, 256, 0, s>>>(U, K, P * C, Ut);
    KSM_LAUNCH_CHECK("transpose_u_kernel");
    dim3 grid(K, (K + 63) / 64);
    if (P == 6) {
        ctc_kernel<6><<<grid, 64, 0, s>>>(Ut, U, wtw, K, C, nt, ctc, ldl);
    } else {
        ctc_kernel_generic<<<grid, 64, 0, s>>>(Ut, U, wtw, K, P, C, nt, ctc, ldl);
    }
    KSM_LAUNCH_CHECK("ctc_kernel");
    return KSM_OK;
}

int ksm_ctc_rows(const float *ctc, int K, int L, int ldl, int32_t *rows, int32_t *nrows,
                 void *stream) {
    if (!ctc || !rows || !nrows || K <= 0 || L <= 0 || ldl < L || K > 65536)
        return ksm::fail(KSM_ERR_INVALID, "ksm_ctc_rows: bad arguments");
    const size_t shm = ksm::align_up(K, 16) + 257 * sizeof(int);
    ctc_rows_kernel<<<K, 256, shm, ksm::as_stream(stream)>>>(ctc, K, L, ldl, rows, nrows);
    KSM_LAUNCH_CHECK("ctc_rows_kernel");
    return KSM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Peel-side template order and the compact Gram.
//
// The peel keeps the scores column-major (Bt[t][r]) over templates sorted by
// peak channel, so that the rows a spike's ctc update touches form one short
// contiguous range.  The compact Gram stores, for the template at position r,
// only the rows [lo_r, hi_r) of that range, lag-major:
//   cc[off[r] + l*(hi_r - lo_r) + (q - lo_r)] = ctc[q', perm[r], l],
// with q' = perm[q] the template at position q.

namespace {

// Wave per template: peak channel = argmax_c sum_p U[j,p,c]^2 (first index).
__global__ __launch_bounds__(1024) void order_kernel(const float *__restrict__ U, int K, int P,
                                                     int C, int32_t *__restrict__ perm,
                                                     int32_t *__restrict__ rank) {
    extern __shared__ int s_mem[];
    int *s_key = s_mem;          // [K]
    int *s_hist = s_mem + K;     // [C]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (int j = wave; j < K; j += nwave) {
        const float *u = U + (size_t)j * P * C;
        float best = -1.f;
        int arg = 0;
        for (int c0 = 0; c0 < C; c0 += 64) {
            const int c = c0 + lane;
            float s = -1.f;
            if (c < C) {
                s = 0.f;
                for (int p = 0; p < P; ++p) {
                    const float v = u[(size_t)p * C + c];
                    s = s + v * v;
                }
            }
            float v = s;
            int a = c;
            for (int o = 32; o > 0; o >>= 1) {
                const float ov = __shfl_xor(v, o, 64);
                const int oa = __shfl_xor(a, o, 64);
                if (ov > v || (ov == v && oa < a)) {
                    v = ov;
                    a = oa;
                }
            }
            if (v > best) {
                best = v;
                arg = a;
            }
        }
        if (lane == 0) s_key[j] = arg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // stable counting sort by peak channel
        for (int c = 0; c < C; ++c) s_hist[c] = 0;
        for (int j = 0; j < K; ++j) s_hist[s_key[j]]++;
        int run = 0;
        for (int c = 0; c < C; ++c) {
            const int v = s_hist[c];
            s_hist[c] = run;
            run += v;
        }
        for (int j = 0; j < K; ++j) {
            const int pos = s_hist[s_key[j]]++;
            perm[pos] = j;
            rank[j] = pos;
        }
    }
}

// Block per firing template j (original id): range of positions of its
// nonzero Gram rows.
__global__ __launch_bounds__(256) void gram_range_kernel(const int32_t *__restrict__ rows,
                                                         const int32_t *__restrict__ nrows,
                                                         const int32_t *__restrict__ rank, int K,
                                                         int32_t *__restrict__ lohi) {
    __shared__ int s_lo[256], s_hi[256];
    const int j = blockIdx.x;
    const int nr = nrows[j];
    int lo = K, hi = 0;
    for (int q = threadIdx.x; q < nr; q += 256) {
        const int r = rank[rows[(size_t)j * K + q]];
        lo = min(lo, r);
        hi = max(hi, r + 1);
    }
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_lo[threadIdx.x] = min(s_lo[threadIdx.x], s_lo[threadIdx.x + o]);
            s_hi[threadIdx.x] = max(s_hi[threadIdx.x], s_hi[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // widened to multiples of 4 positions so that the peel can read the
        // compact Gram with aligned 16-byte loads (the extra rows are zero)
        const int jr = rank[j];
        const bool empty = s_lo[0] >= s_hi[0];
        lohi[2 * jr] = empty ? 0 : (s_lo[0] & ~3);
        lohi[2 * jr + 1] = empty ? 0 : ((s_hi[0] + 3) & ~3);
    }
}

// off[r] = sum_{r' < r} L * (hi - lo), off[K] = total (one block).
__global__ __launch_bounds__(1024) void gram_offsets_kernel(const int32_t *__restrict__ lohi,
                                                            int K, int L,
                                                            int64_t *__restrict__ off) {
    __shared__ int64_t s_part[1024];
    const int seg = (K + 1023) / 1024;
    const int r0 = threadIdx.x * seg, r1 = min(K, r0 + seg);
    int64_t sum = 0;
    for (int r = r0; r < r1; ++r) sum += (int64_t)L * (lohi[2 * r + 1] - lohi[2 * r]);
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int t = 0; t < 1024; ++t) {
            const int64_t v = s_part[t];
            s_part[t] = run;
            run += v;
        }
        off[K] = run;
    }
    __syncthreads();
    int64_t run = s_part[threadIdx.x];
    for (int r = r0; r < r1; ++r) {
        off[r] = run;
        run += (int64_t)L * (lohi[2 * r + 1] - lohi[2 * r]);
    }
}

// Block per position r: gather its row range of the Gram, lag-major.
__global__ __launch_bounds__(256) void gram_compact_kernel(const float *__restrict__ ctc, int ldl,
                                                           const int32_t *__restrict__ perm,
                                                           const int32_t *__restrict__ lohi,
                                                           const int64_t *__restrict__ off, int K,
                                                           int L, float *__restrict__ cc) {
    const int r = blockIdx.x;
    const int j = perm[r];
    const int lo = lohi[2 * r], w = lohi[2 * r + 1] - lo;
    float *dst = cc + off[r];
    const float *src = ctc + (size_t)j * K * ldl;
    const int n = L * w;
    for (int e = threadIdx.x; e < n; e += 256) {
        const int l = e / w, q = lo + e % w;
        dst[e] = q < K ? src[(size_t)perm[q] * ldl + l] : 0.f;
    }
}

}  // namespace

extern "C" {

int ksm_template_order(const float *U, int K, int P, int C, int32_t *perm, int32_t *rank,
                       void *stream) {
    if (!U || !perm || !rank || K <= 0 || P <= 0 || C <= 0 || (size_t)(K + C) * 4 > 150 * 1024)
        return ksm::fail(KSM_ERR_INVALID, "ksm_template_order: bad arguments");
    order_kernel<<<1, 1024, (size_t)(K + C) * sizeof(int), ksm::as_stream(stream)>>>(U, K, P, C,
                                                                                    perm, rank);
    KSM_LAUNCH_CHECK("order_kernel");
    return KSM_OK;
}

int ksm_gram_ranges(const int32_t *rows, const int32_t *nrows, const int32_t *rank, int K, int L,
                    int32_t *lohi, int64_t *off, void *stream) {
    if (!rows || !nrows || !rank || !lohi || !off || K <= 0 || L <= 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_gram_ranges: bad arguments");
    hipStream_t s = ksm::as_stream(stream);
    gram_range_kernel<<<K, 256, 0, s>>>(rows, nrows, rank, K, lohi);
    KSM_LAUNCH_CHECK("gram_range_kernel");
    gram_offsets_kernel<<<1, 1024, 0, s>>>(lohi, K, L, off);
    KSM_LAUNCH_CHECK("gram_offsets_kernel");
    return KSM_OK;
}

int ksm_gram_compact(const float *ctc, int ldl, const int32_t *perm, const int32_t *lohi,
                     const int64_t *off, int K, int L, float *cc, void *stream) {
    if (!ctc || !perm || !lohi || !off || !cc || K <= 0 || L <= 0 || ldl < L)
        return ksm::fail(KSM_ERR_INVALID, "ksm_gram_compact: bad arguments");
    gram_compact_kernel<<<K, 256, 0, ksm::as_stream(stream)>>>(ctc, ldl, perm, lohi, off, K, L,
                                                               cc);
    KSM_LAUNCH_CHECK("gram_compact_kernel");
    return KSM_OK;
}

}  // extern "C"
