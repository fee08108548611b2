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
// (lag-window recomputation) instead of keeping per-spike windows: one wave
// per spike forms its window's projections F (on W) and H (on U[id]), which
// fixed-order reductions turn into dU, dW and the amplitude gradients.
//
// Both passes are deterministic: sums over spikes follow a fixed order (time
// order for the residual, per template then time for the gradients) and
// cross-block partial sums are reduced in a fixed order.  Every kernel exists
// in float and double; the double form serves gradient checks.
#include "ksm_common.h"

namespace {

constexpr int kResT = 256;        // samples per residual block (one per thread)
constexpr int kResC = 16;         // channels per residual block
constexpr int kResSpk = 32;       // spikes staged in LDS at a time
constexpr int kBwdC = 64;         // channels per step of a spike's wave (one per lane)
constexpr int kBwdWaves = 4;      // spikes per backward block (one wave each)
constexpr int kMaxNtR = 255;
constexpr int kMaxJB = (kMaxNtR + 63) / 64;   // lag blocks of 64
constexpr int kMaxPR = 8;
constexpr int kSlices = 256;      // first-level slices of the dW reduction

// rng[2 b], rng[2 b + 1] = spikes (positions in `order`, times
// non-decreasing) whose window [x - h, x + h] meets samples
// [b kResT, (b + 1) kResT).
__global__ void tile_ranges_kernel(const int64_t *__restrict__ st, const int32_t *__restrict__ order,
                                   int nsp, int ntile, int h, int *__restrict__ rng) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 2 * ntile) return;
    const long long v = b % 2 == 0 ? (long long)(b / 2) * kResT - h
                                   : (long long)(b / 2 + 1) * kResT + h;
    int lo = 0, hi = nsp;                       // first q with x(order[q]) >= v
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (st[2 * (size_t)order[mid]] < v) lo = mid + 1;
        else hi = mid;
    }
    rng[b] = lo;
}

// R[c, t] = X[c, t] - sum over the spikes covering t (time order) of
// a_s sum_p U[id_s, p, c] W[p, t - x_s + h]; per-block partial sums of R^2.
// The block's channel slice of each staged spike's U sits in LDS.
template <typename real, int PT>
__global__ __launch_bounds__(kResT) void residual_kernel(
    const real *__restrict__ X, int ldx, const real *__restrict__ U, const real *__restrict__ W,
    const int64_t *__restrict__ st, const real *__restrict__ amps,
    const int32_t *__restrict__ order, const int *__restrict__ rng, int P, int C, int T, int nt,
    real *__restrict__ R, int ldr, double *__restrict__ partial) {
    __shared__ real s_w[PT * kMaxNtR];
    __shared__ real s_u[kResSpk][PT][kResC];
    __shared__ long long s_x[kResSpk];
    __shared__ real s_a[kResSpk];
    __shared__ double s_red[kResT / 64];
    const int h = nt / 2;
    const int t0 = blockIdx.x * kResT, c0 = blockIdx.y * kResC;
    const int tid = threadIdx.x, t = t0 + tid;
    for (int e = tid; e < PT * nt; e += kResT) s_w[e] = e < P * nt ? W[e] : real(0);
    const int q0 = rng[2 * blockIdx.x], q1 = rng[2 * blockIdx.x + 1];
    const int nc = min(C - c0, kResC);
    real acc[kResC];
#pragma unroll
    for (int i = 0; i < kResC; ++i) acc[i] = real(0);
    for (int qb = q0; qb < q1; qb += kResSpk) {
        const int n = min(kResSpk, q1 - qb);
        __syncthreads();
        // thread: spike j = e / kResC, channel i = e % kResC of the slice
        for (int e = tid; e < n * kResC; e += kResT) {
            const int j = e / kResC, i = e % kResC;
            const int sidx = order[qb + j];
            const int id = (int)st[2 * (size_t)sidx + 1];
            if (i == 0) {
                s_x[j] = st[2 * (size_t)sidx];
                s_a[j] = amps[sidx];
            }
            const int c = c0 + i;
#pragma unroll
            for (int p = 0; p < PT; ++p)
                s_u[j][p][i] = (p < P && c < C) ? U[((size_t)id * P + p) * C + c] : real(0);
        }
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            const long long k = (long long)t - s_x[j] + h;
            if (k < 0 || k >= nt) continue;
            const real a = s_a[j];
            real w[PT];
#pragma unroll
            for (int p = 0; p < PT; ++p) w[p] = s_w[p * nt + (int)k];
#pragma unroll
            for (int i = 0; i < kResC; ++i) {
                real tv = real(0);
#pragma unroll
                for (int p = 0; p < PT; ++p) tv += s_u[j][p][i] * w[p];
                acc[i] += a * tv;
            }
        }
    }
    double sq = 0.0;
    if (t < T) {
#pragma unroll
        for (int i = 0; i < kResC; ++i) {
            if (i >= nc) break;
            const int c = c0 + i;
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

// One wave per spike q (template order), over its residual window
// G[c, j] = R[c, x - h + j] (zero outside the batch), 64 channels at a time
// staged in the wave's own LDS rows (coalesced loads along the lags):
//   F[q][p][c] = sum_j G[c, j] W[p, j]          (lane = channel)
//   H[q][p][j] = sum_c U[id, p, c] G[c, j]      (lane = lag)
//   dot[q]     = sum_c sum_p U[id, p, c] F[c, p]
// Every wave runs the same channel steps (waves past the last spike idle
// through them), so the block can synchronise between steps.
template <typename real, int PT>
__global__ __launch_bounds__(64 * kBwdWaves) void spike_windows_kernel(
    const real *__restrict__ R, int ldr, const real *__restrict__ U, const real *__restrict__ W,
    const int64_t *__restrict__ st, const int32_t *__restrict__ order_k, int nsp, int P, int C,
    int T, int nt, real *__restrict__ F, real *__restrict__ H, real *__restrict__ dot) {
    __shared__ real s_w[PT * kMaxNtR];
    extern __shared__ char s_dyn[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ld = nt | 1;                              // odd row stride
    real *s_g = reinterpret_cast<real *>(s_dyn) + (size_t)wave * (kBwdC * ld + kMaxPR * kBwdC);
    real *s_u = s_g + kBwdC * ld;                       // [p][channel]
    for (int e = tid; e < PT * nt; e += 64 * kBwdWaves) s_w[e] = e < P * nt ? W[e] : real(0);
    const int q = blockIdx.x * kBwdWaves + wave;
    const bool live = q < nsp;
    const int h = nt / 2;
    const int s = live ? order_k[q] : 0;
    const long long x = live ? st[2 * (size_t)s] : 0;
    const int id = live ? (int)st[2 * (size_t)s + 1] : 0;
    const int njb = (nt + 63) / 64;
    real hacc[kMaxJB][PT];
#pragma unroll
    for (int b = 0; b < kMaxJB; ++b)
#pragma unroll
        for (int p = 0; p < PT; ++p) hacc[b][p] = real(0);
    real dot_l = real(0);
    for (int c0 = 0; c0 < C; c0 += kBwdC) {
        const int ncs = min(C - c0, kBwdC);
        __syncthreads();                                // previous step's reads done
        if (live) {
            for (int row = 0; row < ncs; ++row) {
                const real *rr = R + (size_t)(c0 + row) * ldr;
                for (int j = lane; j < nt; j += 64) {
                    const long long t = x - h + j;
                    s_g[row * ld + j] = (t >= 0 && t < T) ? rr[t] : real(0);
                }
            }
#pragma unroll
            for (int p = 0; p < PT; ++p)
                s_u[p * kBwdC + lane] =
                    (p < P && lane < ncs) ? U[((size_t)id * P + p) * C + c0 + lane] : real(0);
        }
        __syncthreads();
        if (!live) continue;
        // F: lane = channel
        if (lane < ncs) {
            real f[PT];
#pragma unroll
            for (int p = 0; p < PT; ++p) f[p] = real(0);
            const real *g = s_g + lane * ld;
            for (int j = 0; j < nt; ++j) {
                const real gv = g[j];
#pragma unroll
                for (int p = 0; p < PT; ++p) f[p] += gv * s_w[p * nt + j];
            }
            const int c = c0 + lane;
#pragma unroll
            for (int p = 0; p < PT; ++p) {
                dot_l += s_u[p * kBwdC + lane] * f[p];
                if (p < P) F[((size_t)q * P + p) * C + c] = f[p];
            }
        }
        // H: lane = lag, channels of this step in order
#pragma unroll
        for (int b = 0; b < kMaxJB; ++b) {
            if (b >= njb) break;
            const int j = b * 64 + lane;
            if (j < nt) {
                for (int cc = 0; cc < ncs; ++cc) {
                    const real gv = s_g[cc * ld + j];
#pragma unroll
                    for (int p = 0; p < PT; ++p) hacc[b][p] += s_u[p * kBwdC + cc] * gv;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) dot_l += __shfl_xor(dot_l, o, 64);
    if (!live) return;
    if (lane == 0) dot[q] = dot_l;
#pragma unroll
    for (int b = 0; b < kMaxJB; ++b) {
        const int j = b * 64 + lane;
        if (b < njb && j < nt) {
#pragma unroll
            for (int p = 0; p < PT; ++p)
                if (p < P) H[((size_t)q * P + p) * nt + j] = hacc[b][p];
        }
    }
}

// dU[k][p][c] = -scale * sum over template k's spikes (template order) of
// a_q F[q][p][c].
template <typename real>
__global__ void du_reduce_kernel(const real *__restrict__ F, const real *__restrict__ amps,
                                 const int32_t *__restrict__ order_k,
                                 const int32_t *__restrict__ seg, int P, int C, real scale,
                                 real *__restrict__ dU) {
    const int k = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;        // p * C + c
    if (e >= P * C) return;
    real a = real(0);
    for (int q = seg[k]; q < seg[k + 1]; ++q) a += amps[order_k[q]] * F[(size_t)q * P * C + e];
    dU[(size_t)k * P * C + e] = -scale * a;
}

// damp[s] = -scale * dot[q] for s = order_k[q].
template <typename real>
__global__ void damp_kernel(const real *__restrict__ dot, const int32_t *__restrict__ order_k,
                            int nsp, real scale, real *__restrict__ damp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nsp) damp[order_k[q]] = -scale * dot[q];
}

// First level of dW: slice b sums a_q H[q] over its contiguous range of q.
template <typename real>
__global__ void dw_slices_kernel(const real *__restrict__ H, const real *__restrict__ amps,
                                 const int32_t *__restrict__ order_k, int nsp, int PN,
                                 real *__restrict__ part) {
    const int b = blockIdx.x;
    const int per = (nsp + kSlices - 1) / kSlices;
    const int q0 = b * per, q1 = min(nsp, q0 + per);
    for (int e = threadIdx.x; e < PN; e += blockDim.x) {
        real a = real(0);
        for (int q = q0; q < q1; ++q) a += amps[order_k[q]] * H[(size_t)q * PN + e];
        part[(size_t)b * PN + e] = a;
    }
}

// dW[e] = -scale * sum of the slices, in slice order.
template <typename real>
__global__ void dw_final_kernel(const real *__restrict__ part, int PN, real scale,
                                real *__restrict__ dW) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= PN) return;
    real a = real(0);
    for (int b = 0; b < kSlices; ++b) a += part[(size_t)b * PN + e];
    dW[e] = -scale * a;
}

bool bad_sizes(int nsp, int K, int P, int C, int T, int nt) {
    return nsp < 0 || K <= 0 || P <= 0 || P > kMaxPR || C <= 0 || T <= 0 || nt <= 0 ||
           nt > kMaxNtR || (nt % 2) == 0;
}

size_t elem_size(int dtype) { return dtype == KSM_DTYPE_F64 ? 8 : 4; }

int n_tiles(int T) { return (T + kResT - 1) / kResT; }

template <typename real>
int residual_impl(const void *X, int ldx, const void *U, const void *W, const int64_t *st,
                  const void *amps, const int32_t *order, int nsp, int K, int P, int C, int T,
                  int nt, void *R, int ldr, double *loss, void *workspace, hipStream_t s) {
    dim3 grid(n_tiles(T), (C + kResC - 1) / kResC);
    ksm::Carver cv(workspace);
    double *partial = cv.take<double>((size_t)grid.x * grid.y);
    int *rng = cv.take<int>(2 * (size_t)grid.x);
    tile_ranges_kernel<<<(2 * grid.x + 255) / 256, 256, 0, s>>>(st, order, nsp, grid.x, nt / 2,
                                                                rng);
    KSM_LAUNCH_CHECK("tile_ranges_kernel");
    // P = 6 (Kilosort's n_pcs) compiled exactly; other P padded to kMaxPR
    auto launch = [&](auto kern) {
        kern<<<grid, kResT, 0, s>>>(static_cast<const real *>(X), ldx,
                                    static_cast<const real *>(U), static_cast<const real *>(W),
                                    st, static_cast<const real *>(amps), order, rng, P, C, T, nt,
                                    static_cast<real *>(R), ldr, partial);
    };
    if (P == 6) launch(residual_kernel<real, 6>);
    else launch(residual_kernel<real, kMaxPR>);
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
    ksm::Carver cv(workspace);
    const size_t n1 = nsp > 0 ? nsp : 1;
    real *F = cv.take<real>(n1 * P * C);
    real *H = cv.take<real>(n1 * P * nt);
    real *dot = cv.take<real>(n1);
    real *part = cv.take<real>((size_t)kSlices * P * nt);
    const real sc = (real)scale;
    const real *a = static_cast<const real *>(amps);
    if (nsp > 0) {
        const size_t wdyn =
            (size_t)kBwdWaves * (kBwdC * (nt | 1) + kMaxPR * kBwdC) * sizeof(real);
        auto launch = [&](auto kern) {
            kern<<<(nsp + kBwdWaves - 1) / kBwdWaves, 64 * kBwdWaves, wdyn, s>>>(
                static_cast<const real *>(R), ldr, static_cast<const real *>(U),
                static_cast<const real *>(W), st, order_k, nsp, P, C, T, nt, F, H, dot);
        };
        if (P == 6) launch(spike_windows_kernel<real, 6>);
        else launch(spike_windows_kernel<real, kMaxPR>);
        KSM_LAUNCH_CHECK("spike_windows_kernel");
        damp_kernel<real><<<(nsp + 255) / 256, 256, 0, s>>>(dot, order_k, nsp, sc,
                                                             static_cast<real *>(damp));
        KSM_LAUNCH_CHECK("damp_kernel");
    }
    du_reduce_kernel<real><<<dim3((P * C + 255) / 256, K), 256, 0, s>>>(
        F, a, order_k, seg, P, C, sc, static_cast<real *>(dU));
    KSM_LAUNCH_CHECK("du_reduce_kernel");
    dw_slices_kernel<real><<<kSlices, 256, 0, s>>>(H, a, order_k, nsp, P * nt, part);
    KSM_LAUNCH_CHECK("dw_slices_kernel");
    dw_final_kernel<real><<<(P * nt + 255) / 256, 256, 0, s>>>(part, P * nt, sc,
                                                                static_cast<real *>(dW));
    KSM_LAUNCH_CHECK("dw_final_kernel");
    return KSM_OK;
}

}  // namespace

extern "C" {

size_t ksm_residual_workspace(int nsp, int K, int P, int C, int T, int nt, int dtype) {
    if (bad_sizes(nsp, K, P, C, T, nt) || (dtype != KSM_DTYPE_F32 && dtype != KSM_DTYPE_F64))
        return 0;
    const size_t es = elem_size(dtype);
    const size_t tiles = n_tiles(T);
    const size_t blocks = tiles * ((C + kResC - 1) / kResC);
    const size_t n1 = nsp > 0 ? nsp : 1;
    const size_t fwd = ksm::align_up(blocks * sizeof(double), 256) +
                       ksm::align_up(2 * tiles * sizeof(int), 256);
    const size_t bwd = ksm::align_up(n1 * P * C * es, 256) + ksm::align_up(n1 * P * nt * es, 256) +
                       ksm::align_up(n1 * es, 256) + ksm::align_up((size_t)kSlices * P * nt * es, 256);
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
        return residual_impl<double>(X, ldx, U, W, st, amps, order, nsp, K, P, C, T, nt, R, ldr,
                                     loss, workspace, s);
    return residual_impl<float>(X, ldx, U, W, st, amps, order, nsp, K, P, C, T, nt, R, ldr, loss,
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
