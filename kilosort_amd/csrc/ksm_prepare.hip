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
__global__ void peak_channel_kernel(const float *__restrict__ U, int K, int P, int C,
                                    const int64_t *__restrict__ iCC, int nC,
                                    int64_t *__restrict__ iU, float *__restrict__ Ucc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const float *u = U + (size_t)i * P * C;
    float best = 0.f;
    int arg = 0;
    for (int c = 0; c < C; ++c) {
        float s = 0.f;
        for (int p = 0; p < P; ++p) {
            float v = u[(size_t)p * C + c];
            s = s + v * v;
        }
        if (c == 0 || s > best || (s != s && best == best)) {
            best = s;
            arg = c;
        }
    }
    iU[i] = arg;
    for (int n = 0; n < nC; ++n) {
        const int64_t ch = iCC[(size_t)n * C + arg];
        for (int p = 0; p < P; ++p)
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
template <int P>
__global__ __launch_bounds__(64) void ctc_kernel(const float *__restrict__ Ut,
                                                 const float *__restrict__ U,
                                                 const double *__restrict__ wtw,
                                                 int K, int C, int nt, float *__restrict__ ctc,
                                                 int ldl) {
    extern __shared__ unsigned char smem[];
    const int L = 2 * nt + 1;
    double *s_wtw = reinterpret_cast<double *>(smem);           // [P*P][L]
    float *s_out = reinterpret_cast<float *>(s_wtw + P * P * L);  // [64][ldl]
    const int j = blockIdx.x;
    const int i0 = blockIdx.y * 64;
    const int lane = threadIdx.x;
    const int i = i0 + lane;

    for (int e = lane; e < P * P * L; e += 64) s_wtw[e] = wtw[e];

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
    __syncthreads();
    for (int l = 0; l < ldl; ++l) {
        double acc = 0.0;
        if (l < L) {
#pragma unroll
            for (int e = 0; e < P * P; ++e) acc = fma(utu[e], s_wtw[e * L + l], acc);
        }
        s_out[lane * ldl + l] = (float)acc;
    }
    __syncthreads();
    const int nrow = min(64, K - i0);
    float *dst = ctc + ((size_t)j * K + i0) * ldl;
    for (int e = lane; e < nrow * ldl; e += 64) dst[e] = s_out[e];
}

// Generic-P fallback of ctc_kernel (runtime P <= 8).
__global__ __launch_bounds__(64) void ctc_kernel_generic(
    const float *__restrict__ Ut, const float *__restrict__ U, const double *__restrict__ wtw,
    int K, int P, int C, int nt, float *__restrict__ ctc, int ldl) {
    extern __shared__ unsigned char smem[];
    const int L = 2 * nt + 1;
    double *s_wtw = reinterpret_cast<double *>(smem);
    float *s_out = reinterpret_cast<float *>(s_wtw + P * P * L);
    const int j = blockIdx.x;
    const int i0 = blockIdx.y * 64;
    const int lane = threadIdx.x;
    const int i = i0 + lane;
    for (int e = lane; e < P * P * L; e += 64) s_wtw[e] = wtw[e];
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
    __syncthreads();
    for (int l = 0; l < ldl; ++l) {
        double acc = 0.0;
        if (l < L)
            for (int e = 0; e < P * P; ++e) acc = fma(utu[e], s_wtw[e * L + l], acc);
        s_out[lane * ldl + l] = (float)acc;
    }
    __syncthreads();
    const int nrow = min(64, K - i0);
    float *dst = ctc + ((size_t)j * K + i0) * ldl;
    for (int e = lane; e < nrow * ldl; e += 64) dst[e] = s_out[e];
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
    peak_channel_kernel<<<(K + 63) / 64, 64, 0, ksm::as_stream(stream)>>>(U, K, P, C, iCC, nC,
                                                                          iU, Ucc);
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
    const size_t shm = (size_t)P * P * L * sizeof(double) + (size_t)64 * ldl * sizeof(float);
    if (shm > 160 * 1024)
        return ksm::fail(KSM_ERR_INVALID, "ksm_prepare_matching: nt too large (%d)", nt);
    dim3 grid(K, (K + 63) / 64);
    if (P == 6) {
        ctc_kernel<6><<<grid, 64, shm, s>>>(Ut, U, wtw, K, C, nt, ctc, ldl);
    } else {
        ctc_kernel_generic<<<grid, 64, shm, s>>>(Ut, U, wtw, K, P, C, nt, ctc, ldl);
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
