// Template Gram consumers: normalised template similarity over lags and the
// alignment of templates to the universal waveforms' peaks.
//
// Reference: kilosort/CCG.py:94-102 (similarity), kilosort/template_matching.py
// :128-141 (align_U) and :247-358 (merging_function, whose per-step rows are
// rows of the similarity).
//
// similarity: U[k][p][c] = Wall[k][c][p] / (1e-6 + ||Wall[k]||) (norm in
// double, rounded once), the float64 Gram of ksm_prepare_matching over all
// lags, S[i][j] = max over lags, lag[i][j] = the first lag attaining it.
// align: Y = Wall @ wPCA (the templates' waveforms per channel), laid out as
// one sequence per channel (template after template, as the reference's
// reshape), correlated with the universal waveforms (conv_pc), the peak
// |value| over channels and waveforms per sample, its first time index per
// template, and the templates rolled to centre it and projected back on wPCA.
#include "ksm_common.h"
#include "ksm_profile.h"
#include "ksm_project_internal.h"

namespace {

// U[k][p][c] = Wall[k][c][p] / (1e-6 + mu_k); one block per template.
__global__ __launch_bounds__(256) void normalize_kernel(const float *__restrict__ Wall, int C,
                                                        int P, float *__restrict__ U) {
    __shared__ double s_part[4];
    __shared__ float s_mu;
    const int k = blockIdx.x, n = C * P;
    const float *w = Wall + (size_t)k * n;
    double acc = 0.0;
    for (int e = threadIdx.x; e < n; e += 256) acc += (double)w[e] * (double)w[e];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        s_mu = 1e-6f + (float)sqrt(((s_part[0] + s_part[1]) + s_part[2]) + s_part[3]);
    __syncthreads();
    const float d = s_mu;
    for (int e = threadIdx.x; e < n; e += 256) {
        const int c = e / P, p = e - c * P;
        U[((size_t)k * P + p) * C + c] = w[e] / d;
    }
}

// S[i][j] = max over l < L of store[j][i][l] (the Gram of ksm_prepare_matching
// is stored firing-template major); lag[i][j] = the first l attaining it.
// S[i][j] = max over l < L of store[j][i][l] (the Gram of ksm_prepare_matching
// is stored firing-template major); lag[i][j] = the first l attaining it.  One
// wave per (j, i) row: lanes read consecutive lags, then a wave reduction
// (larger value first, then smaller lag).
__global__ __launch_bounds__(256) void maxlag_kernel(const float *__restrict__ store, int K,
                                                     int L, int ldl, float *__restrict__ S,
                                                     int32_t *__restrict__ lag) {
    const int lane = threadIdx.x & 63;
    const long long nrow = (long long)K * K;
    const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long row = w0; row < nrow; row += nw) {
        const float *r = store + (size_t)row * ldl;      // row = j K + i
        float m = -INFINITY;
        int bl = 0x7fffffff;
        for (int l = lane; l < L; l += 64) {
            const float v = r[l];
            if (v > m || (v == m && l < bl)) {
                m = v;
                bl = l;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float v2 = __shfl_xor(m, o, 64);
            const int l2 = __shfl_xor(bl, o, 64);
            if (v2 > m || (v2 == m && l2 < bl)) {
                m = v2;
                bl = l2;
            }
        }
        if (lane == 0) {
            const int j = (int)(row / K), i = (int)(row - (long long)j * K);
            const long long e = (long long)i * K + j;
            S[e] = m;
            if (lag) lag[e] = bl;
        }
    }
}

// X[c][k nt + t] = Y[(k C + c) nt + t].
__global__ void sequence_kernel(const float *__restrict__ Y, int K, int C, int nt,
                                float *__restrict__ X) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)K * C * nt;
    if (e >= n) return;
    const int t = (int)(e % nt);
    const long long r = e / nt;
    const int c = (int)(r % C), k = (int)(r / C);
    X[(size_t)c * K * nt + (size_t)k * nt + t] = Y[e];
}

// imax[k] = the first t < nt maximising max_{q, c} |Bc[q][c][k nt + t]|.
__global__ __launch_bounds__(256) void peak_kernel(const float *__restrict__ Bc, int nq, int C,
                                                   int ldc, int nt, int32_t *__restrict__ imax) {
    __shared__ float s_m[4][64];
    __shared__ float s_col[256];
    const int k = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int t0 = 0; t0 < nt; t0 += 64) {
        // lane: column t0 + lane; waves split the (q, c) rows
        const int t = t0 + lane;
        float m = 0.f;
        if (t < nt)
            for (int r = wave; r < nq * C; r += 4)
                m = fmaxf(m, fabsf(Bc[(size_t)r * ldc + (size_t)k * nt + t]));
        s_m[wave][lane] = m;
        __syncthreads();
        if (wave == 0 && t0 + lane < 256)
            s_col[t0 + lane] = fmaxf(fmaxf(s_m[0][lane], s_m[1][lane]),
                                     fmaxf(s_m[2][lane], s_m[3][lane]));
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float best = s_col[0];
        int bt = 0;
        for (int t = 1; t < nt; ++t)
            if (s_col[t] > best) {
                best = s_col[t];
                bt = t;
            }
        imax[k] = bt;
    }
}

// Unew[k][p][c] = sum_t wPCA[p][t] Y[(k C + c) nt + (t - s_k) mod nt],
// s_k = nt/2 - imax[k] (torch.roll by s_k along time).  One block per
// (template, 64 channels): the 64 waveforms are copied to LDS with coalesced
// loads (rows padded to nt + 1), then thread (p, c) sums over t in order.
constexpr int kRollCh = 64;
constexpr int kMaxRollNt = 128;

__global__ __launch_bounds__(256) void roll_project_kernel(const float *__restrict__ Y,
                                                           const float *__restrict__ wPCA,
                                                           const int32_t *__restrict__ imax,
                                                           int K, int C, int P, int nt,
                                                           float *__restrict__ Unew) {
    __shared__ float s_y[kRollCh * (kMaxRollNt + 1)];
    const int k = blockIdx.y;
    const int c0 = blockIdx.x * kRollCh;
    const int nc = min(kRollCh, C - c0);
    const float *y0 = Y + ((size_t)k * C + c0) * nt;
    for (int e = threadIdx.x; e < nc * nt; e += blockDim.x) {
        const int q = e / nt;
        s_y[q * (nt + 1) + (e - q * nt)] = y0[e];
    }
    __syncthreads();
    const int s = nt / 2 - imax[k];
    for (int e = threadIdx.x; e < P * kRollCh; e += blockDim.x) {
        const int p = e / kRollCh, cl = e - p * kRollCh;
        if (cl >= nc) continue;
        const float *y = s_y + cl * (nt + 1);
        float acc = 0.f;
        for (int t = 0; t < nt; ++t) {
            int u = (t - s) % nt;
            if (u < 0) u += nt;
            acc = __builtin_fmaf(wPCA[p * nt + t], y[u], acc);
        }
        Unew[((size_t)k * P + p) * C + c0 + cl] = acc;
    }
}

// Spike positions (postprocessing.py:33-62): per spike s of template id with
// peak channel u = iU[id], m_j = ||tF[s, j, :]||_2 * mask[j, u],
// w_j = m_j / sum_j m_j, (xs, ys) = sum_j w_j (xc, yc)[iCC[j, u]]
// (iCC, mask [nC][C] over the C channels); amp[s] = ||tF[s]||_F when amp is
// given (the Phy amplitudes, io.py:330).
template <typename Row>
__device__ __forceinline__ void position_of(Row f0, long long s, const int64_t *__restrict__ ids,
                                            int nC, int P, const int64_t *__restrict__ iU,
                                            const int64_t *__restrict__ iCC,
                                            const uint8_t *__restrict__ mask, int C,
                                            const float *__restrict__ xc,
                                            const float *__restrict__ yc, float *__restrict__ xs,
                                            float *__restrict__ ys, float *__restrict__ amp) {
    const int u = (int)iU[ids[s]];
    float m[32];
    float tot = 0.f, all = 0.f;
    for (int j = 0; j < nC; ++j) {
        Row f = f0 + j * P;
        float a = 0.f;
        for (int p = 0; p < P; ++p) a = __builtin_fmaf(f[p], f[p], a);
        all += a;
        const float v = sqrtf(a) * (mask[(size_t)j * C + u] ? 1.f : 0.f);
        m[j] = v;
        tot += v;
    }
    if (amp) amp[s] = sqrtf(all);
    float x = 0.f, y = 0.f;
    for (int j = 0; j < nC; ++j) {
        const float w = m[j] / tot;
        const int c = (int)iCC[(size_t)j * C + u];
        x += xc[c] * w;
        y += yc[c] * w;
    }
    xs[s] = x;
    ys[s] = y;
}

// One thread per spike; the block's rows (contiguous in tF) are first copied
// to LDS with coalesced loads, each row padded to an odd stride R + 1.
__global__ void positions_staged_kernel(const float *__restrict__ tF,
                                        const int64_t *__restrict__ ids, long long n, int nC,
                                        int P, const int64_t *__restrict__ iU,
                                        const int64_t *__restrict__ iCC,
                                        const uint8_t *__restrict__ mask, int C,
                                        const float *__restrict__ xc,
                                        const float *__restrict__ yc, float *__restrict__ xs,
                                        float *__restrict__ ys, float *__restrict__ amp) {
    extern __shared__ float s_rows[];
    const int R = nC * P;
    const long long s0 = (long long)blockIdx.x * blockDim.x;
    const long long nb = n - s0 < (long long)blockDim.x ? n - s0 : (long long)blockDim.x;
    const float *src = tF + (size_t)s0 * R;
    const int total = (int)nb * R;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int q = e / R;
        s_rows[q * (R + 1) + (e - q * R)] = src[e];
    }
    __syncthreads();
    if (threadIdx.x >= nb) return;
    position_of(s_rows + threadIdx.x * (R + 1), s0 + threadIdx.x, ids, nC, P, iU, iCC, mask, C,
                xc, yc, xs, ys, amp);
}

// The same from global memory, for rows too long to stage (nC P > 255).
__global__ void positions_kernel(const float *__restrict__ tF, const int64_t *__restrict__ ids,
                                 long long n, int nC, int P, const int64_t *__restrict__ iU,
                                 const int64_t *__restrict__ iCC, const uint8_t *__restrict__ mask,
                                 int C, const float *__restrict__ xc, const float *__restrict__ yc,
                                 float *__restrict__ xs, float *__restrict__ ys,
                                 float *__restrict__ amp) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    position_of(tF + (size_t)s * nC * P, s, ids, nC, P, iU, iCC, mask, C, xc, yc, xs, ys, amp);
}

__global__ void clear_flag_t_kernel(int *__restrict__ f) { *f = 0; }

}  // namespace

extern "C" {

size_t ksm_similarity_workspace(int K, int C, int P, int nt) {
    if (K <= 0 || C <= 0 || P <= 0 || nt <= 0) return 0;
    const int ldl = (int)ksm::align_up(2 * nt + 1, 4);
    return ksm::align_up((size_t)K * P * C * sizeof(float), 256) +
           ksm::align_up((size_t)K * K * ldl * sizeof(float), 256) +
           ksm::align_up(ksm_prepare_matching_workspace(K, P, C, nt), 256) + 256;
}

int ksm_similarity(const float *Wall, const float *W, int K, int C, int P, int nt, float *S,
                   int32_t *lag, void *workspace, size_t ws_bytes, void *stream) {
    if (!Wall || !W || !S || K <= 0 || C <= 0 || P <= 0 || nt <= 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_similarity: bad arguments");
    if (!workspace || ws_bytes < ksm_similarity_workspace(K, C, P, nt))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_similarity: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const int L = 2 * nt + 1, ldl = (int)ksm::align_up(L, 4);
    ksm::Carver cv(workspace);
    float *U = cv.take<float>((size_t)K * P * C);
    float *store = cv.take<float>((size_t)K * K * ldl);
    const size_t nb = ksm_prepare_matching_workspace(K, P, C, nt);
    void *ws2 = cv.take<unsigned char>(nb);
    normalize_kernel<<<K, 256, 0, s>>>(Wall, C, P, U);
    KSM_LAUNCH_CHECK("normalize_kernel");
    int rc = ksm_prepare_matching(U, W, K, P, C, nt, store, ldl, ws2, nb, stream);
    if (rc != KSM_OK) return rc;
    const long long n = (long long)K * K;
    maxlag_kernel<<<(unsigned)std::min<long long>((n + 3) / 4, 65536), 256, 0, s>>>(
        store, K, L, ldl, S, lag);
    KSM_LAUNCH_CHECK("maxlag_kernel");
    return KSM_OK;
}

size_t ksm_align_workspace(int K, int C, int P, int nt, int nk) {
    if (K <= 0 || C <= 0 || P <= 0 || nt <= 0 || nk <= 0) return 0;
    const size_t T = (size_t)K * nt, ldc = ksm::align_up(T, 256);
    return 2 * ksm::align_up(T * C * sizeof(float), 256) +
           ksm::align_up((size_t)nk * C * ldc * sizeof(float), 256) + 512;
}

int ksm_align_templates(const float *Wall, const float *wPCA, const float *wTEMP, int K, int C,
                        int P, int nt, int nk, float *Unew, int32_t *imax, void *workspace,
                        size_t ws_bytes, void *stream) {
    if (!Wall || !wPCA || !wTEMP || !Unew || !imax || K <= 0 || C <= 0 || P <= 0 || nt <= 0 ||
        (nt % 2) == 0 || nt > 256 || nk <= 0 || (long long)K * nt > (1ll << 30))
        return ksm::fail(KSM_ERR_INVALID, "ksm_align_templates: bad arguments");
    if (!workspace || ws_bytes < ksm_align_workspace(K, C, P, nt, nk))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_align_templates: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const int T = K * nt, ldc = (int)ksm::align_up(T, 256);
    ksm::Carver cv(workspace);
    float *Y = cv.take<float>((size_t)T * C);
    float *X = cv.take<float>((size_t)T * C);
    float *Bc = cv.take<float>((size_t)nk * C * ldc);
    int *flag = cv.take<int>(1);
    // Y[(k C + c)][t] = sum_p Wall[k][c][p] wPCA[p][t]
    int rc = ksm_gemm_f32(Wall, P, wPCA, nt, Y, nt, K * C, nt, P, stream);
    if (rc != KSM_OK) return rc;
    const long long n = (long long)T * C;
    sequence_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(Y, K, C, nt, X);
    KSM_LAUNCH_CHECK("sequence_kernel");
    clear_flag_t_kernel<<<1, 1, 0, s>>>(flag);
    KSM_LAUNCH_CHECK("clear_flag_t_kernel");
    rc = ksm::conv_pc(X, T, wTEMP, C, T, nk, nt, Bc, ldc, flag, s);
    if (rc != KSM_OK) return rc;
    peak_kernel<<<K, 256, 0, s>>>(Bc, nk, C, ldc, nt, imax);
    KSM_LAUNCH_CHECK("peak_kernel");
    if (nt > kMaxRollNt)
        return ksm::fail(KSM_ERR_INVALID, "ksm_align_templates: nt <= %d", kMaxRollNt);
    roll_project_kernel<<<dim3((C + kRollCh - 1) / kRollCh, K), 256, 0, s>>>(
        Y, wPCA, imax, K, C, P, nt, Unew);
    KSM_LAUNCH_CHECK("roll_project_kernel");
    return KSM_OK;
}

int ksm_spike_positions(const float *tF, const int64_t *ids, long long n, int nC, int P,
                        const int64_t *iU, const int64_t *iCC, const uint8_t *mask, int C,
                        const float *xc, const float *yc, float *xs, float *ys, float *amp,
                        void *stream) {
    if (n < 0 || (n > 0 && (!tF || !ids || !iU || !iCC || !mask || !xc || !yc || !xs || !ys)) ||
        nC <= 0 || nC > 32 || P <= 0 || C <= 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_spike_positions: bad arguments (nC <= 32)");
    if (n == 0) return KSM_OK;
    const int R = nC * P;
    if (R <= 255) {
        const int spb = (R + 1) * 256 * 4 <= 65536 ? 256 : ((R + 1) * 128 * 4 <= 65536 ? 128 : 64);
        positions_staged_kernel<<<(unsigned)((n + spb - 1) / spb), spb,
                                  (size_t)spb * (R + 1) * sizeof(float), ksm::as_stream(stream)>>>(
            tF, ids, n, nC, P, iU, iCC, mask, C, xc, yc, xs, ys, amp);
        KSM_LAUNCH_CHECK("positions_staged_kernel");
    } else {
        positions_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ksm::as_stream(stream)>>>(
            tF, ids, n, nC, P, iU, iCC, mask, C, xc, yc, xs, ys, amp);
        KSM_LAUNCH_CHECK("positions_kernel");
    }
    return KSM_OK;
}

}  // extern "C"
