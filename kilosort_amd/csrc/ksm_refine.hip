// Residual of a fixed spike set and its gradients (template refinement).
//
// The model of one padded batch is the matching path's own subtraction
// (kilosort/template_matching.py:235-236): spike s at column x_s with
// template id_s and amplitude a_s contributes
//     a_s * T_id[c, k],   T_i[c, k] = sum_p U[i, p, c] * W[p, k],
// at sample t = x_s - nt/2 + k, k = 0 .. nt-1.  The reference refines
// templates by averaging spike features (clustering_qr.py:524-527) and has no
// gradients; these kernels give the residual
//     R = X - sum_s a_s T_{id_s}(. - x_s),   loss = sum R^2
// and, with spike times and ids held fixed, the gradients of scale * loss / 2
// (scale = 2 for the plain loss):
//     dU[i, p, c] = -scale * sum_{s: id_s = i} a_s * sum_k R[c, x_s - h + k] W[p, k]
//     dW[p, k]    = -scale * sum_s a_s * sum_c U[id_s, p, c] R[c, x_s - h + k]
//     da_s        = -scale * sum_c sum_k R[c, x_s - h + k] T_{id_s}[c, k]
// The backward pass reads the residual windows of each spike again from R
// (lag-window recomputation) instead of keeping per-spike windows.
//
// Both passes are deterministic: sums over spikes follow a fixed order (time
// order for the residual, per template then time for the gradients) and
// cross-block partial sums are reduced in a fixed order.  Every kernel exists
// in float and double; the double form serves gradient checks.
#include "ksm_common.h"

namespace {

constexpr int kResT = 256;        // samples per residual block (one per thread)
constexpr int kResC = 16;         // channels per residual block
constexpr int kResSpk = 64;       // spikes staged in LDS at a time
constexpr int kBwdC = 64;         // channels per backward block (one per lane)
constexpr int kMaxNtR = 255;
constexpr int kMaxPR = 8;

// Spike q of order[] (x, id) from st [nsp][2].
__device__ __forceinline__ long long spike_x(const int64_t *st, const int32_t *order, int q) {
    return st[2 * (size_t)order[q]];
}

// First index q in [0, nsp] with x(order[q]) >= v (times non-decreasing).
__device__ int lower_bound_x(const int64_t *st, const int32_t *order, int nsp, long long v) {
    int lo = 0, hi = nsp;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (spike_x(st, order, mid) < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// R[c, t] = X[c, t] - sum over the spikes covering t (time order) of
// a_s * sum_p U[id, p, c] W[p, t - x_s + h]; per-block partial sums of R^2.
template <typename real>
__global__ __launch_bounds__(kResT) void residual_kernel(
    const real *__restrict__ X, int ldx, const real *__restrict__ U, const real *__restrict__ W,
    const int64_t *__restrict__ st, const real *__restrict__ amps,
    const int32_t *__restrict__ order, int nsp, int P, int C, int T, int nt,
    real *__restrict__ R, int ldr, double *__restrict__ partial) {
    __shared__ real s_w[kMaxPR * kMaxNtR];
    __shared__ long long s_x[kResSpk];
    __shared__ int s_id[kResSpk];
    __shared__ real s_a[kResSpk];
    __shared__ int s_rng[2];
    __shared__ double s_red[kResT / 64];
    const int h = nt / 2;
    const int t0 = blockIdx.x * kResT, c0 = blockIdx.y * kResC;
    const int tid = threadIdx.x, t = t0 + tid;
    for (int e = tid; e < P * nt; e += kResT) s_w[e] = W[e];
    if (tid == 0) {
        // spikes whose window [x - h, x + h] meets [t0, t0 + kResT)
        s_rng[0] = lower_bound_x(st, order, nsp, (long long)t0 - h);
        s_rng[1] = lower_bound_x(st, order, nsp, (long long)t0 + kResT + h);
    }
    __syncthreads();
    const int q0 = s_rng[0], q1 = s_rng[1];
    const int c1 = min(C, c0 + kResC);
    real acc[kResC];
#pragma unroll
    for (int i = 0; i < kResC; ++i) acc[i] = real(0);
    for (int qb = q0; qb < q1; qb += kResSpk) {
        const int n = min(kResSpk, q1 - qb);
        __syncthreads();
        if (tid < n) {
            const int s = order[qb + tid];
            s_x[tid] = st[2 * (size_t)s];
            s_id[tid] = (int)st[2 * (size_t)s + 1];
            s_a[tid] = amps[s];
        }
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            const long long k = (long long)t - s_x[j] + h;
            if (k < 0 || k >= nt) continue;
            const real a = s_a[j];
            const real *u = U + (size_t)s_id[j] * P * C;
#pragma unroll
            for (int i = 0; i < kResC; ++i) {
                const int c = c0 + i;
                if (c >= c1) break;
                real tv = real(0);
                for (int p = 0; p < P; ++p) tv += u[(size_t)p * C + c] * s_w[p * nt + (int)k];
                acc[i] += a * tv;
            }
        }
    }
    double sq = 0.0;
    if (t < T) {
#pragma unroll
        for (int i = 0; i < kResC; ++i) {
            const int c = c0 + i;
            if (c >= c1) break;
            const real r = X[(size_t)c * ldx + t] - acc[i];
            if (R) R[(size_t)c * ldr + t] = r;
            sq += (double)r * (double)r;
        }
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
    if ((tid & 63) == 0) s_red[tid >> 6] = sq;
    __syncthreads();
    if (tid == 0) {
        double b = 0.0;
        for (int w = 0; w < kResT / 64; ++w) b += s_red[w];
        partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = b;
    }
}

// loss[0] = sum of the partials, in index order.
__global__ void sum_partials_kernel(const double *__restrict__ partial, int n,
                                    double *__restrict__ loss) {
    __shared__ double s[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a += partial[i];
    s[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int i = 0; i < 256; ++i) b += s[i];
        loss[0] = b;
    }
}

// Block (template k, channels [c0, c0 + 64)): over the template's spikes in
// time order, with G the residual window of the spike on these channels,
//   F[c, p] = sum_j G[c, j] W[p, j]     -> dU[k, p, c] and the spike's dot
//   H[p, j] = sum_c U[k, p, c] G[c, j]  -> this block's share of dW.
// Wave 0: lane = channel (F, dU, dot); wave 1: lane = lag (H).
template <typename real>
__global__ __launch_bounds__(128) void residual_bwd_kernel(
    const real *__restrict__ R, int ldr, const real *__restrict__ U, const real *__restrict__ W,
    const int64_t *__restrict__ st, const real *__restrict__ amps,
    const int32_t *__restrict__ order_k, const int32_t *__restrict__ seg, int P, int C, int T,
    int nt, real scale, real *__restrict__ dU, real *__restrict__ dWpart,
    real *__restrict__ dotpart, int nchunk) {
    __shared__ real s_w[kMaxPR * kMaxNtR];
    __shared__ real s_u[kMaxPR * kBwdC];
    extern __shared__ char s_dyn[];                     // G window [kBwdC][nt | 1]
    real *s_g = reinterpret_cast<real *>(s_dyn);
    const int k = blockIdx.x, chunk = blockIdx.y;
    const int c0 = chunk * kBwdC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = nt / 2;
    const int ld = nt | 1;                              // odd row stride
    for (int e = tid; e < P * nt; e += 128) s_w[e] = W[e];
    for (int e = tid; e < P * kBwdC; e += 128) {
        const int p = e / kBwdC, l = e % kBwdC;
        s_u[e] = c0 + l < C ? U[((size_t)k * P + p) * C + c0 + l] : real(0);
    }
    real accU[kMaxPR], accW[kMaxPR];
#pragma unroll
    for (int p = 0; p < kMaxPR; ++p) accU[p] = accW[p] = real(0);
    const int q0 = seg[k], q1 = seg[k + 1];
    for (int q = q0; q < q1; ++q) {
        const int s = order_k[q];
        const long long x = st[2 * (size_t)s];
        const real a = amps[s];
        __syncthreads();            // the previous window is no longer read
        for (int e = tid; e < kBwdC * nt; e += 128) {
            const int l = e / nt, j = e % nt;
            const long long t = x - h + j;
            const int c = c0 + l;
            s_g[l * ld + j] = (c < C && t >= 0 && t < T) ? R[(size_t)c * ldr + t] : real(0);
        }
        __syncthreads();
        if (wave == 0) {
            real f[kMaxPR];
#pragma unroll
            for (int p = 0; p < kMaxPR; ++p) f[p] = real(0);
            const real *g = s_g + lane * ld;
            for (int j = 0; j < nt; ++j) {
                const real gv = g[j];
#pragma unroll
                for (int p = 0; p < kMaxPR; ++p)
                    if (p < P) f[p] += gv * s_w[p * nt + j];
            }
            real dot = real(0);
#pragma unroll
            for (int p = 0; p < kMaxPR; ++p)
                if (p < P) {
                    accU[p] += a * f[p];
                    dot += s_u[p * kBwdC + lane] * f[p];
                }
            for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, 64);
            if (lane == 0) dotpart[(size_t)q * nchunk + chunk] = dot;
        } else if (lane < nt || lane + 64 < nt) {
            // lags lane, lane + 64, ... (nt <= 255): this wave covers them in turn
            for (int j = lane; j < nt; j += 64) {
                real hv[kMaxPR];
#pragma unroll
                for (int p = 0; p < kMaxPR; ++p) hv[p] = real(0);
                for (int l = 0; l < kBwdC; ++l) {
                    const real gv = s_g[l * ld + j];
#pragma unroll
                    for (int p = 0; p < kMaxPR; ++p)
                        if (p < P) hv[p] += s_u[p * kBwdC + l] * gv;
                }
                // lane-held accumulators cover lag `lane` only; further lags
                // (nt > 64) go straight to their partial slot
                if (j == lane) {
#pragma unroll
                    for (int p = 0; p < kMaxPR; ++p) accW[p] += a * hv[p];
                } else {
                    real *dw = dWpart + ((size_t)k * gridDim.y + chunk) * P * nt;
                    for (int p = 0; p < P; ++p) dw[p * nt + j] += a * hv[p];
                }
            }
        }
    }
    if (wave == 0) {
        if (c0 + lane < C)
            for (int p = 0; p < P; ++p) dU[((size_t)k * P + p) * C + c0 + lane] = -scale * accU[p];
    } else if (lane < nt) {
        real *dw = dWpart + ((size_t)k * gridDim.y + chunk) * P * nt;
        for (int p = 0; p < P; ++p) dw[p * nt + lane] += accW[p];
    }
}

// damp[s] = -scale * sum over channel chunks of the dots (chunk order).
template <typename real>
__global__ void dot_reduce_kernel(const real *__restrict__ dotpart, const int32_t *__restrict__ order_k,
                                  int nsp, int nchunk, real scale, real *__restrict__ damp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsp) return;
    real a = real(0);
    for (int c = 0; c < nchunk; ++c) a += dotpart[(size_t)q * nchunk + c];
    damp[order_k[q]] = -scale * a;
}

// dW[p, j] = -scale * sum over (template, chunk) partials (index order).
template <typename real>
__global__ void dw_reduce_kernel(const real *__restrict__ dWpart, int nparts, int PN, real scale,
                                 real *__restrict__ dW) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= PN) return;
    real a = real(0);
    for (int i = 0; i < nparts; ++i) a += dWpart[(size_t)i * PN + e];
    dW[e] = -scale * a;
}

template <typename real>
__global__ void zero_kernel(real *__restrict__ p, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        p[e] = real(0);
}

bool bad_sizes(int nsp, int K, int P, int C, int T, int nt) {
    return nsp < 0 || K <= 0 || P <= 0 || P > kMaxPR || C <= 0 || T <= 0 || nt <= 0 ||
           nt > kMaxNtR || (nt % 2) == 0;
}

size_t elem_size(int dtype) { return dtype == KSM_DTYPE_F64 ? 8 : 4; }

template <typename real>
int residual_impl(const void *X, int ldx, const void *U, const void *W, const int64_t *st,
                  const void *amps, const int32_t *order, int nsp, int P, int C, int T, int nt,
                  void *R, int ldr, double *loss, void *workspace, hipStream_t s) {
    dim3 grid((T + kResT - 1) / kResT, (C + kResC - 1) / kResC);
    double *partial = static_cast<double *>(workspace);
    residual_kernel<real><<<grid, kResT, 0, s>>>(
        static_cast<const real *>(X), ldx, static_cast<const real *>(U),
        static_cast<const real *>(W), st, static_cast<const real *>(amps), order, nsp, P, C, T,
        nt, static_cast<real *>(R), ldr, partial);
    KSM_LAUNCH_CHECK("residual_kernel");
    if (loss) {
        sum_partials_kernel<<<1, 256, 0, s>>>(partial, (int)(grid.x * grid.y), loss);
        KSM_LAUNCH_CHECK("sum_partials_kernel");
    }
    return KSM_OK;
}

template <typename real>
int backward_impl(const void *R, int ldr, const void *U, const void *W, const int64_t *st,
                  const void *amps, const int32_t *order_k, const int32_t *seg, int nsp, int K,
                  int P, int C, int T, int nt, double scale, void *dU, void *dW, void *damp,
                  void *workspace, hipStream_t s) {
    const int nchunk = (C + kBwdC - 1) / kBwdC;
    ksm::Carver cv(workspace);
    real *dWpart = cv.take<real>((size_t)K * nchunk * P * nt);
    real *dotpart = cv.take<real>((size_t)(nsp > 0 ? nsp : 1) * nchunk);
    const size_t nparts = (size_t)K * nchunk * P * nt;
    zero_kernel<real><<<256, 256, 0, s>>>(dWpart, nparts);
    KSM_LAUNCH_CHECK("zero_kernel");
    const size_t gdyn = (size_t)kBwdC * (nt | 1) * sizeof(real);
    residual_bwd_kernel<real><<<dim3(K, nchunk), 128, gdyn, s>>>(
        static_cast<const real *>(R), ldr, static_cast<const real *>(U),
        static_cast<const real *>(W), st, static_cast<const real *>(amps), order_k, seg, P, C, T,
        nt, (real)scale, static_cast<real *>(dU), dWpart, dotpart, nchunk);
    KSM_LAUNCH_CHECK("residual_bwd_kernel");
    if (nsp > 0) {
        dot_reduce_kernel<real><<<(nsp + 255) / 256, 256, 0, s>>>(
            dotpart, order_k, nsp, nchunk, (real)scale, static_cast<real *>(damp));
        KSM_LAUNCH_CHECK("dot_reduce_kernel");
    }
    dw_reduce_kernel<real><<<(P * nt + 255) / 256, 256, 0, s>>>(
        dWpart, K * nchunk, P * nt, (real)scale, static_cast<real *>(dW));
    KSM_LAUNCH_CHECK("dw_reduce_kernel");
    return KSM_OK;
}

}  // namespace

extern "C" {

size_t ksm_residual_workspace(int nsp, int K, int P, int C, int T, int nt, int dtype) {
    if (bad_sizes(nsp, K, P, C, T, nt) || (dtype != KSM_DTYPE_F32 && dtype != KSM_DTYPE_F64))
        return 0;
    const size_t blocks = (size_t)((T + kResT - 1) / kResT) * ((C + kResC - 1) / kResC);
    const size_t nchunk = (C + kBwdC - 1) / kBwdC;
    const size_t fwd = ksm::align_up(blocks * sizeof(double), 256);
    const size_t bwd = ksm::align_up((size_t)K * nchunk * P * nt * elem_size(dtype), 256) +
                       ksm::align_up((size_t)(nsp > 0 ? nsp : 1) * nchunk * elem_size(dtype), 256);
    return (fwd > bwd ? fwd : bwd) + 256;
}

int ksm_residual(const void *X, int ldx, const void *U, const void *W, const int64_t *st,
                 const void *amps, const int32_t *order, int nsp, int K, int P, int C, int T,
                 int nt, void *R, int ldr, double *loss, int dtype, void *workspace,
                 size_t ws_bytes, void *stream) {
    if (bad_sizes(nsp, K, P, C, T, nt) || !X || !U || !W || ldx < T ||
        (nsp > 0 && (!st || !amps || !order)) || (R && ldr < T) || (!R && !loss) ||
        (dtype != KSM_DTYPE_F32 && dtype != KSM_DTYPE_F64))
        return ksm::fail(KSM_ERR_INVALID, "ksm_residual: bad arguments");
    if (!workspace || ws_bytes < ksm_residual_workspace(nsp, K, P, C, T, nt, dtype))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_residual: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    if (dtype == KSM_DTYPE_F64)
        return residual_impl<double>(X, ldx, U, W, st, amps, order, nsp, P, C, T, nt, R, ldr,
                                     loss, workspace, s);
    return residual_impl<float>(X, ldx, U, W, st, amps, order, nsp, P, C, T, nt, R, ldr, loss,
                                workspace, s);
}

int ksm_residual_backward(const void *R, int ldr, const void *U, const void *W,
                          const int64_t *st, const void *amps, const int32_t *order_k,
                          const int32_t *seg, int nsp, int K, int P, int C, int T, int nt,
                          double scale, void *dU, void *dW, void *damp, int dtype,
                          void *workspace, size_t ws_bytes, void *stream) {
    if (bad_sizes(nsp, K, P, C, T, nt) || !R || ldr < T || !U || !W || !seg || !dU || !dW ||
        (nsp > 0 && (!st || !amps || !order_k || !damp)) ||
        (dtype != KSM_DTYPE_F32 && dtype != KSM_DTYPE_F64))
        return ksm::fail(KSM_ERR_INVALID, "ksm_residual_backward: bad arguments");
    if (!workspace || ws_bytes < ksm_residual_workspace(nsp, K, P, C, T, nt, dtype))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_residual_backward: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    if (dtype == KSM_DTYPE_F64)
        return backward_impl<double>(R, ldr, U, W, st, amps, order_k, seg, nsp, K, P, C, T, nt,
                                     scale, dU, dW, damp, workspace, s);
    return backward_impl<float>(R, ldr, U, W, st, amps, order_k, seg, nsp, K, P, C, T, nt, scale,
                                dU, dW, damp, workspace, s);
}

}  // extern "C"
