// The matching-pursuit peel and the per-spike PC features.
//
// Reference: kilosort/template_matching.py:183-244 (peel loop of
// run_matching) and :85-86 (features in extract).
//
// Device state per batch (carved from the caller's workspace):
//   cfmax[T], imax[T]   best template score and id of every column
//                        (Cf = relu(B)^2/nm, columns < nt and >= T-nt are 0)
//   slots               accepted spikes of the current iteration, per chunk
//                        of kChunk columns, in increasing time
//   state               done flag, status, iteration count, and kbase[it]
//                        (spikes recorded before iteration it)
//
// One iteration = peel_select (window max, threshold, ordered compaction)
// + peel_update (each block owns a slab of kSlab columns of B for all
// templates: it applies the ctc updates of the spikes whose +-nt window
// touches the slab -- even-ranked spikes first, as the reference's two
// interleaved phases do -- then recomputes cfmax/imax of its columns, and
// subtracts the templates of the spikes it owns from Xres).  Once an
// iteration accepts nothing the done flag makes every later launch return
// at once, so the host never waits inside the loop.
#include "ksm_common.h"
#include "ksm_profile.h"

namespace {

constexpr int kChunk = 1024;        // columns per select block (4 per thread)
constexpr int kSlotCap = 32;        // accepted spikes per chunk (>= 1024/62)
constexpr int kSlab = 64;           // columns per update block (one per lane)
constexpr int kMaxTouch = 16;       // spikes touching one slab (<= 3 without ties)
constexpr int kMaxNt = 255;

enum { ST_DONE = 0, ST_STATUS = 1, ST_ITERS = 2, ST_KBASE = 4 };

struct Slot {
    int x;       // column in the padded batch
    int iy;      // template id
    float amp;   // B[iy, x] / nm[iy]
    float tha;   // sqrt(Cmax[x])
};

__device__ __forceinline__ int wave_sum_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// torch.max semantics for one reduction step: NaN wins, then larger value,
// then the smaller index on equal values.
__device__ __forceinline__ bool beats(float v, int i, float bv, int bi) {
    const bool vn = v != v, bn = bv != bv;
    if (vn || bn) return vn && (!bn || i < bi);
    return v > bv || (v == bv && i < bi);
}

// cfmax/imax of columns [s0, s0 + kSlab) from all K rows of B.
// 256 threads: lane = column, wave w scans rows w, w+4, ...
__device__ void slab_colmax(const float *__restrict__ B, int ldb, const float *__restrict__ nm,
                            int K, int s0, int T, int nt, float *__restrict__ cfmax,
                            int *__restrict__ imax, float *s_val, int *s_idx) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = s0 + lane;
    const bool live = col >= nt && col < T - nt;
    float best = 0.f;
    int bi = 0x7fffffff;
    if (live) {
        const float *bp = B + col;
        int i = wave;
        if (i < K) {
            const float b = bp[(size_t)i * ldb];
            const float r = b > 0.f ? b : 0.f;
            best = (r * r) / nm[i];
            bi = i;
            i += 4;
        }
        for (; i < K; i += 4) {
            const float b = bp[(size_t)i * ldb];
            const float r = b > 0.f ? b : 0.f;
            const float cf = (r * r) / nm[i];
            if (cf > best || (cf != cf && best == best)) {
                best = cf;
                bi = i;
            }
        }
    }
    s_val[threadIdx.x] = best;
    s_idx[threadIdx.x] = bi;
    __syncthreads();
    if (wave == 0 && col < T) {
        float v = s_val[lane];
        int vi = s_idx[lane];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            const float o = s_val[w * 64 + lane];
            const int oi = s_idx[w * 64 + lane];
            if (beats(o, oi, v, vi)) {
                v = o;
                vi = oi;
            }
        }
        if (!live) {
            v = 0.f;
            vi = 0;
        }
        cfmax[col] = v;
        imax[col] = vi;
    }
}

__global__ void peel_init_kernel(int *state, int n_state) {
    for (int e = threadIdx.x; e < n_state; e += blockDim.x) state[e] = 0;
}

__global__ __launch_bounds__(256) void colmax_all_kernel(const float *__restrict__ B, int ldb,
                                                         const float *__restrict__ nm, int K,
                                                         int T, int nt,
                                                         float *__restrict__ cfmax,
                                                         int *__restrict__ imax) {
    __shared__ float s_val[256];
    __shared__ int s_idx[256];
    slab_colmax(B, ldb, nm, K, blockIdx.x * kSlab, T, nt, cfmax, imax, s_val, s_idx);
}

// Accepts columns whose score is the maximum of its +-nt window and exceeds
// Th^2: (Cmax > Th^2) & (|Cmax - Cfmax| < 1e-9) with Cmax the window max
// (template_matching.py:205-212), and records them in time order.
__global__ __launch_bounds__(256) void peel_select_kernel(
    const float *__restrict__ cfmax, const int *__restrict__ imax, const float *__restrict__ B,
    int ldb, const float *__restrict__ nm, int T, int nt, float th2, float th2_lo,
    int *__restrict__ state, int *__restrict__ chunk_count, Slot *__restrict__ slots) {
    if (__builtin_nontemporal_load(&state[ST_DONE])) return;
    __shared__ float win[kChunk + 2 * kMaxNt];
    __shared__ int s_wc[4];
    const int t0 = blockIdx.x * kChunk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < kChunk + 2 * nt; e += 256) {
        const int x = t0 - nt + e;
        win[e] = (x >= 0 && x < T) ? cfmax[x] : -INFINITY;
    }
    __syncthreads();
    int count = 0;
    Slot *out = slots + (size_t)blockIdx.x * kSlotCap;
    for (int u = 0; u < 4; ++u) {
        const int x = t0 + u * 256 + tid;
        bool acc = false;
        float cmax = 0.f;
        if (x >= nt && x < T - nt) {
            const int e0 = x - t0 + nt;
            const float v = win[e0];
            if (v >= th2_lo) {
                cmax = v;
                for (int d = -nt; d <= nt; ++d) {
                    const float o = win[e0 + d];
                    if (o > cmax || o != o) cmax = o;
                    if (cmax != cmax) break;
                }
                acc = (cmax > th2) && (fabsf(cmax - v) < 1e-9f);
            }
        }
        const unsigned long long m = __ballot(acc);
        if (lane == 0) s_wc[wave] = __popcll(m);
        __syncthreads();
        int off = count;
        for (int w = 0; w < wave; ++w) off += s_wc[w];
        const int total = s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
        if (acc) {
            const int pos = off + __popcll(m & lanemask_lt(lane));
            if (pos < kSlotCap) {
                const int iy = imax[x];
                Slot sl;
                sl.x = x;
                sl.iy = iy;
                sl.amp = B[(size_t)iy * ldb + x] / nm[iy];
                sl.tha = sqrtf(cmax);
                out[pos] = sl;
            }
        }
        count += total;
        __syncthreads();
    }
    if (tid == 0) {
        if (count > kSlotCap) {
            atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
            atomicExch(&state[ST_DONE], 1);
            count = kSlotCap;
        }
        chunk_count[blockIdx.x] = count;
    }
}

__global__ __launch_bounds__(256) void peel_update_kernel(
    int it, float *__restrict__ B, int ldb, float *__restrict__ Xres, int ldx,
    const float *__restrict__ U, const float *__restrict__ W, const float *__restrict__ nm,
    const float *__restrict__ ctc, int ldl, const int32_t *__restrict__ rows,
    const int32_t *__restrict__ nrows, int K, int P, int C, int T, int nt, int nchunks,
    const int *__restrict__ chunk_count, const Slot *__restrict__ slots, int *__restrict__ state,
    int64_t *__restrict__ st, float *__restrict__ amps, float *__restrict__ th_amps, int capacity,
    float *__restrict__ cfmax, int *__restrict__ imax) {
    if (__builtin_nontemporal_load(&state[ST_DONE])) return;
    __shared__ Slot s_sp[kMaxTouch];
    __shared__ int s_rank[kMaxTouch];
    __shared__ int s_total, s_n;
    __shared__ float s_val[256];
    __shared__ int s_idx[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s0 = blockIdx.x * kSlab;
    const int s1 = min(s0 + kSlab, T);
    const int lo_x = s0 - nt, hi_x = s1 - 1 + nt;
    const int c_lo = max(lo_x, 0) / kChunk;
    const int c_hi = min(hi_x, T - 1) / kChunk;
    if (wave == 0) {
        int tot = 0, pre = 0;
        for (int c = lane; c < nchunks; c += 64) {
            const int v = chunk_count[c];
            tot += v;
            if (c < c_lo) pre += v;
        }
        tot = wave_sum_i(tot);
        pre = wave_sum_i(pre);
        int n = 0;
        for (int c = c_lo; c <= c_hi; ++c) {
            const int cnt = chunk_count[c];
            for (int b = 0; b < cnt; b += 64) {
                const int q = b + lane;
                Slot sl;
                bool in = false;
                if (q < cnt) {
                    sl = slots[(size_t)c * kSlotCap + q];
                    in = sl.x >= lo_x && sl.x <= hi_x;
                }
                const unsigned long long m = __ballot(in);
                const int pos = n + __popcll(m & lanemask_lt(lane));
                if (in && pos < kMaxTouch) {
                    s_sp[pos] = sl;
                    s_rank[pos] = pre + q;
                }
                n += __popcll(m);
            }
            pre += cnt;
        }
        if (lane == 0) {
            s_total = tot;
            s_n = n;
        }
    }
    __syncthreads();
    const int total = s_total;
    if (total == 0) {
        if (blockIdx.x == 0 && tid == 0) atomicExch(&state[ST_DONE], 1);
        return;
    }
    const int kb = state[ST_KBASE + it];
    if (kb + total > capacity) {
        if (blockIdx.x == 0 && tid == 0) {
            atomicExch(&state[ST_STATUS], KSM_ERR_CAPACITY);
            atomicExch(&state[ST_DONE], 1);
        }
        return;
    }
    if (blockIdx.x == 0 && tid == 0) {
        state[ST_KBASE + it + 1] = kb + total;
        state[ST_ITERS] = it + 1;
    }
    const int n = s_n;
    if (n == 0) return;
    if (n > kMaxTouch) {
        if (tid == 0) {
            atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
            atomicExch(&state[ST_DONE], 1);
        }
        return;
    }
    // record the spikes whose column lies in this slab
    if (tid < n) {
        const Slot sl = s_sp[tid];
        if (sl.x >= s0 && sl.x < s1) {
            const size_t pos = (size_t)kb + s_rank[tid];
            st[2 * pos] = sl.x;
            st[2 * pos + 1] = sl.iy;
            amps[pos] = sl.amp;
            th_amps[pos] = sl.tha;
        }
    }
    // B[:, x-nt .. x+nt] -= amp * ctc[:, id, :]: phase 0 (even rank), phase 1
    const int col = s0 + lane;
    for (int ph = 0; ph < 2; ++ph) {
        for (int q = 0; q < n; ++q) {
            if ((s_rank[q] & 1) != ph) continue;
            const int x = s_sp[q].x, j = s_sp[q].iy;
            const float a = s_sp[q].amp;
            const bool inwin = col < s1 && col >= x - nt && col <= x + nt;
            const int nr = nrows[j];
            const int32_t *rl = rows + (size_t)j * K;
            const float *cj = ctc + (size_t)j * K * ldl + (col - x + nt);
            if (inwin) {
                for (int r = wave; r < nr; r += 4) {
                    const int i = rl[r];
                    float *bp = B + (size_t)i * ldb + col;
                    const float prod = a * cj[(size_t)i * ldl];
                    *bp = *bp - prod;
                }
            }
        }
        __syncthreads();
    }
    slab_colmax(B, ldb, nm, K, s0, T, nt, cfmax, imax, s_val, s_idx);
    // Xres[:, x-nt/2 .. x+nt/2] -= amp * (U[id]^T W) for owned spikes
    const int h = nt / 2;
    for (int q = 0; q < n; ++q) {
        const Slot sl = s_sp[q];
        if (sl.x < s0 || sl.x >= s1) continue;
        const float *uj = U + (size_t)sl.iy * P * C;
        for (int c = wave; c < C; c += 4) {
            float *xr = Xres + (size_t)c * ldx + sl.x - h;
            for (int k = lane; k < nt; k += 64) {
                float tm = 0.f;
                for (int p = 0; p < P; ++p)
                    tm = __builtin_fmaf(uj[(size_t)p * C + c], W[p * nt + k], tm);
                const float prod = sl.amp * tm;
                xr[k] = xr[k] - prod;
            }
        }
    }
}

__global__ void peel_finish_kernel(const int *__restrict__ state, int32_t *__restrict__ info) {
    const int iters = state[ST_ITERS];
    info[0] = state[ST_KBASE + iters];
    info[1] = state[ST_STATUS];
    info[2] = iters;
    info[3] = 0;
}

// tF[s][n][p] = sum_k Xres[iCC[n, iU[id]], x + k - nt/2] W[p,k] + amps[s] Ucc[n,id,p]
__global__ __launch_bounds__(256) void features_kernel(
    const float *__restrict__ Xres, int ldx, const float *__restrict__ W,
    const int64_t *__restrict__ st, const float *__restrict__ amps, int nsp,
    const int64_t *__restrict__ iCC, const int64_t *__restrict__ iU, const float *__restrict__ Ucc,
    int nC, int K, int P, int C, int nt, float *__restrict__ tF) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nsp * nC) return;
    const int s = g / nC, n = g % nC;
    const int x = (int)st[2 * s];
    const int id = (int)st[2 * s + 1];
    const int ch = (int)iCC[(size_t)n * C + iU[id]];
    const float *xr = Xres + (size_t)ch * ldx + x - nt / 2;
    const float a = amps[s];
    for (int p = 0; p < P; ++p) {
        float f = 0.f;
        for (int k = 0; k < nt; ++k) f = __builtin_fmaf(xr[k], W[p * nt + k], f);
        const float own = a * Ucc[((size_t)n * K + id) * P + p];
        tF[((size_t)s * nC + n) * P + p] = f + own;
    }
}

struct PeelLayout {
    int nchunks, nslabs, n_state;
};

PeelLayout peel_layout(int T, int max_peels) {
    PeelLayout L;
    L.nchunks = (T + kChunk - 1) / kChunk;
    L.nslabs = (T + kSlab - 1) / kSlab;
    L.n_state = ST_KBASE + max_peels + 1;
    return L;
}

}  // namespace

extern "C" {

size_t ksm_peel_workspace(int T, int max_peels) {
    if (T <= 0 || max_peels < 0) return 0;
    const PeelLayout L = peel_layout(T, max_peels);
    const size_t Tp = ksm::align_up(T, kSlab);
    return ksm::align_up(Tp * 4, 256) * 2 + ksm::align_up(L.nchunks * 4, 256) +
           ksm::align_up((size_t)L.nchunks * kSlotCap * sizeof(Slot), 256) +
           ksm::align_up((size_t)L.n_state * 4, 256) + 256;
}

int ksm_peel(float *B, int ldb, float *Xres, int ldx, const float *U, const float *W,
             const float *nm, const float *ctc, int ldl, const int32_t *rows,
             const int32_t *nrows, int K, int P, int C, int T, int nt, float Th, int max_peels,
             int64_t *st, float *amps, float *th_amps, int capacity, int32_t *info,
             void *workspace, size_t ws_bytes, void *stream) {
    if (!B || !Xres || !U || !W || !nm || !ctc || !rows || !nrows || !st || !amps ||
        !th_amps || !info || K <= 0 || P <= 0 || C <= 0 || nt <= 0 || nt > kMaxNt ||
        (nt % 2) == 0 || T < 2 * nt + 1 || ldb < T || ldx < T || ldl < 2 * nt + 1 ||
        max_peels < 0 || capacity < 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_peel: bad arguments");
    if (!workspace || ws_bytes < ksm_peel_workspace(T, max_peels))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_peel: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const PeelLayout L = peel_layout(T, max_peels);
    const size_t Tp = ksm::align_up(T, kSlab);
    ksm::Carver cv(workspace);
    float *cfmax = cv.take<float>(Tp);
    int *imax = cv.take<int>(Tp);
    int *chunk_count = cv.take<int>(L.nchunks);
    Slot *slots = cv.take<Slot>((size_t)L.nchunks * kSlotCap);
    int *state = cv.take<int>(L.n_state);

    // Th^2 as the reference forms it (Python float, compared in float32)
    const double th2d = (double)Th * (double)Th;
    const float th2 = (float)th2d;
    // columns below th2_lo can never satisfy the acceptance test
    const float th2_lo = (float)(th2d - fmax(4e-9, fabs(th2d) * 1e-6));

    peel_init_kernel<<<1, 256, 0, s>>>(state, L.n_state);
    KSM_LAUNCH_CHECK("peel_init_kernel");
    {
        ksm::ProfileScope prof("colmax_all", s);
        colmax_all_kernel<<<L.nslabs, 256, 0, s>>>(B, ldb, nm, K, T, nt, cfmax, imax);
    }
    KSM_LAUNCH_CHECK("colmax_all_kernel");
    for (int it = 0; it < max_peels; ++it) {
        {
            ksm::ProfileScope prof("peel_select", s);
            peel_select_kernel<<<L.nchunks, 256, 0, s>>>(cfmax, imax, B, ldb, nm, T, nt, th2,
                                                         th2_lo, state, chunk_count, slots);
        }
        {
            ksm::ProfileScope prof("peel_update", s);
            peel_update_kernel<<<L.nslabs, 256, 0, s>>>(
                it, B, ldb, Xres, ldx, U, W, nm, ctc, ldl, rows, nrows, K, P, C, T, nt,
                L.nchunks, chunk_count, slots, state, st, amps, th_amps, capacity, cfmax, imax);
        }
    }
    KSM_LAUNCH_CHECK("peel loop");
    peel_finish_kernel<<<1, 1, 0, s>>>(state, info);
    KSM_LAUNCH_CHECK("peel_finish_kernel");
    return KSM_OK;
}

int ksm_features(const float *Xres, int ldx, const float *W, const int64_t *st,
                 const float *amps, int nsp, const int64_t *iCC, const int64_t *iU,
                 const float *Ucc, int nC, int K, int P, int C, int nt, float *tF, void *stream) {
    if (nsp < 0 || nC <= 0 || K <= 0 || P <= 0 || C <= 0 || nt <= 0 || nC > C)
        return ksm::fail(KSM_ERR_INVALID, "ksm_features: bad arguments");
    if (nsp == 0) return KSM_OK;
    if (!Xres || !W || !st || !amps || !iCC || !iU || !Ucc || !tF)
        return ksm::fail(KSM_ERR_INVALID, "ksm_features: null pointer");
    const int n = nsp * nC;
    features_kernel<<<(n + 255) / 256, 256, 0, ksm::as_stream(stream)>>>(
        Xres, ldx, W, st, amps, nsp, iCC, iU, Ucc, nC, K, P, C, nt, tF);
    KSM_LAUNCH_CHECK("features_kernel");
    return KSM_OK;
}

}  // extern "C"
