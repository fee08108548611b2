// The matching-pursuit peel and the per-spike PC features.
//
// Reference: kilosort/template_matching.py:183-244 (peel loop of
// run_matching) and :85-86 (features in extract).
//
// Layout.  Scores are column-major, Bt[t][r] (ksm_project), over templates in
// peel order r (perm[r] = template id, rank[id] = r; templates sorted by peak
// channel), so one time column is one contiguous row of Bt and the rows a
// spike's Gram update touches form one contiguous range [lo, hi).  The Gram
// is the compact lag-major array of ksm_gram_compact.
//
// Device state per batch (carved from the caller's workspace):
//   cfmax[T], imax[T]   best score and template id of every column
//                        (Cf = relu(B)^2/nm; columns < nt and >= T-nt are 0)
//   slots               accepted spikes of the current iteration, per chunk
//                        of kChunk columns, in increasing time
//   state               done flag, status, iteration count, and kbase[it]
//                        (spikes recorded before iteration it)
//
// One iteration is two launches on the caller's stream:
//   peel_select  window max, threshold and ordered compaction into slots;
//   peel_update  one block per slab of kSlab columns.  A wave takes one time
//                column at a time (lanes over templates): it subtracts
//                amp * ctc of the spikes whose +-nt window covers the column
//                -- the even-ranked spike first, as the reference's two
//                interleaved phases do -- writes the changed scores back, and
//                reduces the column's best score across lanes.  The block
//                also records the spikes whose column lies in the slab into
//                st/amps/th_amps and subtracts their templates from Xres.
// Once an iteration accepts nothing the done flag makes every later launch
// return at once, so the host never waits inside the loop.
#include "ksm_common.h"
#include "ksm_profile.h"

namespace {

constexpr int kChunk = 256;         // columns per select block (one per thread)
constexpr int kSlotCap = 16;        // accepted spikes per chunk (>= 256/62 + ties)
constexpr int kSlab = 64;           // columns per update block
constexpr int kUpdWaves = 8;        // waves per update block
constexpr int kUpdThreads = kUpdWaves * 64;
constexpr int kSlots4 = 3;          // 256-position slots of a column held in registers
constexpr int kMaxTouch = 16;       // spikes touching one slab (<= 3 without ties)
constexpr int kMaxNt = 255;
constexpr int kMaxK = 8192;

enum { ST_DONE = 0, ST_STATUS = 1, ST_ITERS = 2, ST_KBASE = 4 };

struct Slot {
    int x;       // column in the padded batch
    int iy;      // template id
    float amp;   // B[iy, x] / nm[iy]
    float tha;   // sqrt(Cmax[x])
};

// A touching spike as the column pass needs it.
struct SpikeRef {
    int x, lo, hi, w;       // column, Gram row range [lo, hi), w = hi - lo
    float amp;
    long long off;          // start of the template's block in the compact Gram
};

__device__ __forceinline__ int wave_sum_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// torch.max semantics for one reduction step: NaN wins, then larger value,
// then the smaller template id on equal values.
__device__ __forceinline__ bool beats(float v, int i, float bv, int bi) {
    const bool vn = v != v, bn = bv != bv;
    if (vn || bn) return vn && (!bn || i < bi);
    return v > bv || (v == bv && i < bi);
}

// Cf = relu(b)^2 / nm, rounded as torch rounds relu(B)**2 / nm.
__device__ __forceinline__ float score(float b, float n) {
    const float r = b > 0.f ? b : 0.f;
    return (r * r) / n;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ long long uni(long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// One covering spike, wave-uniform, in scalar registers: the Gram row of
// this column for positions [lo, hi) is cc[base + lo .. base + hi).
struct ColSpike {
    long long base;
    int lo, hi;
    float amp;
};

__device__ __forceinline__ ColSpike col_spike(const SpikeRef &rf, int t, int nt) {
    ColSpike cs;
    cs.lo = uni(rf.lo);
    cs.hi = uni(rf.hi);
    cs.amp = uni(rf.amp);
    cs.base = uni((long long)(rf.off + (long long)(t - rf.x + nt) * rf.w - rf.lo));
    return cs;
}

// One time column t (wave-uniform): apply the covering spikes' Gram updates
// (products rounded before the subtract; spike a before spike b; an absent
// spike has an empty range), store the changed scores, and return the
// column's best (score, template id) in every lane.  s_nm / s_pid:
// per-position norms and template ids in LDS.
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Column pass.  The scores are first ranked with a multiply by the rounded
// reciprocal of nm (within 3 ulp of the IEEE quotient); only positions whose
// approximate score is within 2^-20 of the column's approximate maximum can
// hold the exact maximum, and only those get the exact division and the
// (score, template id) comparison.  A NaN anywhere in the column, or a column
// maximum below 2^-100, sends the whole column through the exact comparison
// (NaN wins, as in torch.max).
__device__ __forceinline__ void column_pass(float *__restrict__ Bt, int ldbt, int t, int K,
                                            const float *__restrict__ cc, const float *s_nm,
                                            const float *s_inv, const int *s_pid, ColSpike a,
                                            ColSpike b, float &best, int &bid) {
    const int lane = threadIdx.x & 63;
    float *col = Bt + (size_t)t * ldbt;
    best = -INFINITY;
    bid = 0x7fffffff;
    // Each lane holds 4 consecutive positions (16-byte loads; ldbt, the Gram
    // ranges and their offsets are multiples of 4).  Loads are branch-free
    // (indices clamped into valid ranges, values discarded where not needed)
    // and all stores come after all loads, so no wait on the load counter has
    // to cover a store of the same column.
    const bool a_any = a.hi > a.lo, b_any = b.hi > b.lo;
    for (int r0 = 0; r0 < K; r0 += 256 * kSlots4) {
        float4 v[kSlots4], va[kSlots4], vb[kSlots4];
#pragma unroll
        for (int k = 0; k < kSlots4; ++k) {
            const int r = r0 + 256 * k + 4 * lane;
            const int ra = a_any ? min(max(r, a.lo), a.hi - 4) : 0;
            const int rb = b_any ? min(max(r, b.lo), b.hi - 4) : 0;
            v[k] = *reinterpret_cast<const float4 *>(col + min(r, ldbt - 4));
            va[k] = *reinterpret_cast<const float4 *>(cc + a.base + ra);
            vb[k] = *reinterpret_cast<const float4 *>(cc + b.base + rb);
        }
        float x[kSlots4][4];
        bool dirty[kSlots4];
        float amax = 0.f;
        bool nan_any = false;
#pragma unroll
        for (int k = 0; k < kSlots4; ++k) {
            const int r = r0 + 256 * k + 4 * lane;
            const bool ha = r >= a.lo && r < a.hi;
            const bool hb = r >= b.lo && r < b.hi;
            const float bv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            const float ca[4] = {va[k].x, va[k].y, va[k].z, va[k].w};
            const float cb[4] = {vb[k].x, vb[k].y, vb[k].z, vb[k].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float pa = a.amp * ca[e];
                const float pb = b.amp * cb[e];
                float y = bv[e];
                y = ha ? y - pa : y;
                y = hb ? y - pb : y;
                x[k][e] = y;
                if (r + e < K) {
                    const float rr = y > 0.f ? y : 0.f;
                    const float ap = (rr * rr) * s_inv[r + e];
                    nan_any |= ap != ap;
                    amax = fmaxf(amax, ap);
                }
            }
            dirty[k] = ha || hb;
        }
#pragma unroll
        for (int k = 0; k < kSlots4; ++k)
            if (dirty[k])
                *reinterpret_cast<float4 *>(col + r0 + 256 * k + 4 * lane) =
                    make_float4(x[k][0], x[k][1], x[k][2], x[k][3]);
        const float amax_w = wave_max(amax);
        // below 2^-100 the relative error bound of the approximation no
        // longer holds (subnormal products); compare exactly there too
        const bool exact_all = __ballot(nan_any) != 0ull || amax_w < 0x1p-100f;
        const float thr = amax_w * (1.0f - 0x1p-20f);
#pragma unroll
        for (int k = 0; k < kSlots4; ++k) {
            const int r = r0 + 256 * k + 4 * lane;
            bool cand[4];
            bool any = false;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float rr = x[k][e] > 0.f ? x[k][e] : 0.f;
                const float ap = (rr * rr) * s_inv[r + e < K ? r + e : 0];
                cand[e] = (r + e < K) && (exact_all || ap >= thr);
                any |= cand[e];
            }
            if (__ballot(any) == 0ull) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (!cand[e]) continue;
                const float cf = score(x[k][e], s_nm[r + e]);
                const int id = s_pid[r + e];
                if (beats(cf, id, best, bid)) {
                    best = cf;
                    bid = id;
                }
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bid, o, 64);
        if (beats(ov, oi, best, bid)) {
            best = ov;
            bid = oi;
        }
    }
}

// Per-position norms, their rounded reciprocals and template ids into LDS.
__device__ __forceinline__ void load_positions(const float *__restrict__ nm,
                                               const int32_t *__restrict__ perm, int K,
                                               float *s_nm, float *s_inv, int *s_pid, int first,
                                               int stride) {
    for (int r = first; r < K; r += stride) {
        const int id = perm[r];
        const float n = nm[id];
        s_pid[r] = id;
        s_nm[r] = n;
        s_inv[r] = 1.0f / n;
    }
}

__global__ void peel_init_kernel(int *state, int n_state) {
    for (int e = threadIdx.x; e < n_state; e += blockDim.x) state[e] = 0;
}

// Best score of every column (first pass, no updates).
__global__ __launch_bounds__(kUpdThreads) void colmax_all_kernel(
    float *__restrict__ Bt, int ldbt, const float *__restrict__ nm,
    const int32_t *__restrict__ perm, int K, int T, int nt, float *__restrict__ cfmax,
    int *__restrict__ imax) {
    extern __shared__ unsigned char smem[];
    float *s_nm = reinterpret_cast<float *>(smem);
    float *s_inv = s_nm + K;
    int *s_pid = reinterpret_cast<int *>(s_inv + K);
    load_positions(nm, perm, K, s_nm, s_inv, s_pid, threadIdx.x, blockDim.x);
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int s0 = blockIdx.x * kSlab;
    for (int c = wv; c < kSlab; c += kUpdWaves) {
        const int t = s0 + c;
        if (t >= T) break;
        float best;
        int bid;
        if (t >= nt && t < T - nt) {
            const ColSpike none = {0, 0, 0, 0.f};
            column_pass(Bt, ldbt, t, K, Bt, s_nm, s_inv, s_pid, none, none, best, bid);
        } else {
            best = 0.f;
            bid = 0;
        }
        if (lane == 0) {
            cfmax[t] = best;
            imax[t] = bid;
        }
    }
}

// Accepts columns whose score is the maximum of its +-nt window and exceeds
// Th^2: (Cmax > Th^2) & (|Cmax - Cfmax| < 1e-9) with Cmax the window max
// (template_matching.py:205-212), and records them in time order.
__global__ __launch_bounds__(kChunk) void peel_select_kernel(
    const float *__restrict__ cfmax, const int *__restrict__ imax, const float *__restrict__ Bt,
    int ldbt, const float *__restrict__ nm, const int32_t *__restrict__ rank, int T, int nt,
    float th2, float th2_lo, int early_exit, int *__restrict__ state,
    int *__restrict__ chunk_count, Slot *__restrict__ slots) {
    if (state[ST_DONE]) return;
    __shared__ float win[kChunk + 2 * kMaxNt];
    __shared__ int s_wc[kChunk / 64];
    const int t0 = blockIdx.x * kChunk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < kChunk + 2 * nt; e += kChunk) {
        const int x = t0 - nt + e;
        win[e] = (x >= 0 && x < T) ? cfmax[x] : -INFINITY;
    }
    __syncthreads();
    const int x = t0 + tid;
    bool acc = false;
    float cmax = 0.f;
    if (x >= nt && x < T - nt) {
        const int e0 = tid + nt;
        const float v = win[e0];
        if (v >= th2_lo) {
            if (early_exit) {
                // With v >= 2^-6 any larger neighbour differs from v by at
                // least one ulp of v (> 1e-9), so the column is rejected as
                // soon as one is found; scanning outwards finds it within a
                // few samples for columns that are not a peak.
                bool ok = true;
                for (int d = 1; d <= nt; ++d) {
                    const float a = win[e0 - d], b = win[e0 + d];
                    if (a > v || b > v || a != a || b != b) {
                        ok = false;
                        break;
                    }
                }
                cmax = v;
                acc = ok && (v > th2);
            } else {
                cmax = v;
                for (int d = -nt; d <= nt; ++d) {
                    const float o = win[e0 + d];
                    if (o > cmax || o != o) cmax = o;
                    if (cmax != cmax) break;
                }
                acc = (cmax > th2) && (fabsf(cmax - v) < 1e-9f);
            }
        }
    }
    const unsigned long long m = __ballot(acc);
    if (lane == 0) s_wc[wave] = __popcll(m);
    __syncthreads();
    int off = 0, count = 0;
#pragma unroll
    for (int w = 0; w < kChunk / 64; ++w) {
        if (w < wave) off += s_wc[w];
        count += s_wc[w];
    }
    if (acc) {
        const int pos = off + __popcll(m & lanemask_lt(lane));
        if (pos < kSlotCap) {
            const int iy = imax[x];
            Slot sl;
            sl.x = x;
            sl.iy = iy;
            sl.amp = Bt[(size_t)x * ldbt + rank[iy]] / nm[iy];
            sl.tha = sqrtf(cmax);
            slots[(size_t)blockIdx.x * kSlotCap + pos] = sl;
        }
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

__global__ __launch_bounds__(kUpdThreads) void peel_update_kernel(
    int it, float *__restrict__ Bt, int ldbt, float *__restrict__ Xres, int ldx,
    const float *__restrict__ U, const float *__restrict__ W, const int *__restrict__ crange,
    int P, int C, const float *__restrict__ nm, const int32_t *__restrict__ perm,
    const int32_t *__restrict__ rank, const float *__restrict__ cc,
    const int32_t *__restrict__ lohi, const int64_t *__restrict__ off, int K, int T, int nt,
    int nchunks, const int *__restrict__ chunk_count, const Slot *__restrict__ slots,
    int *__restrict__ state, int64_t *__restrict__ st, float *__restrict__ amps,
    float *__restrict__ th_amps, int capacity, float *__restrict__ cfmax,
    int *__restrict__ imax) {
    if (state[ST_DONE]) return;
    extern __shared__ unsigned char smem[];
    float *s_nm = reinterpret_cast<float *>(smem);
    float *s_inv = s_nm + K;
    int *s_pid = reinterpret_cast<int *>(s_inv + K);
    __shared__ Slot s_sp[kMaxTouch];
    __shared__ SpikeRef s_ref[kMaxTouch];
    __shared__ int s_rank[kMaxTouch];
    __shared__ int s_total, s_n, s_bad;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int s0 = blockIdx.x * kSlab;
    const int s1 = min(s0 + kSlab, T);
    const int lo_x = s0 - nt, hi_x = s1 - 1 + nt;
    const int c_lo = max(lo_x, 0) / kChunk;
    const int c_hi = min(hi_x, T - 1) / kChunk;
    if (wv == 0) {
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
            Slot sl = {0, 0, 0.f, 0.f};
            bool in = false;
            if (lane < cnt) {
                sl = slots[(size_t)c * kSlotCap + lane];
                in = sl.x >= lo_x && sl.x <= hi_x;
            }
            const unsigned long long m = __ballot(in);
            const int pos = n + __popcll(m & lanemask_lt(lane));
            if (in && pos < kMaxTouch) {
                s_sp[pos] = sl;
                s_rank[pos] = pre + lane;
            }
            n += __popcll(m);
            pre += cnt;
        }
        if (lane == 0) {
            s_total = tot;
            s_n = n;
            s_bad = 0;
        }
    } else {
        // the other waves stage the per-position tables meanwhile
        load_positions(nm, perm, K, s_nm, s_inv, s_pid, tid - 64, kUpdThreads - 64);
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
    if (tid < n) {
        const Slot sl = s_sp[tid];
        if (sl.x >= s0 && sl.x < s1) {
            // record the spikes whose column lies in this slab
            const size_t pos = (size_t)kb + s_rank[tid];
            st[2 * pos] = sl.x;
            st[2 * pos + 1] = sl.iy;
            amps[pos] = sl.amp;
            th_amps[pos] = sl.tha;
        }
        const int jr = rank[sl.iy];
        SpikeRef rf;
        rf.x = sl.x;
        rf.lo = lohi[2 * jr];
        rf.hi = lohi[2 * jr + 1];
        rf.w = rf.hi - rf.lo;
        rf.amp = sl.amp;
        rf.off = off[jr];
        s_ref[tid] = rf;
    }
    __syncthreads();
    // time columns of the slab, one per wave at a time
    for (int c = wv; c < kSlab; c += kUpdWaves) {
        const int t = s0 + c;
        if (t >= s1) break;
        int qa = -1, qb = -1;
        bool bad = false;
        for (int ph = 0; ph < 2; ++ph)
            for (int q = 0; q < n; ++q) {
                if ((s_rank[q] & 1) != ph) continue;
                const int x = s_ref[q].x;
                if (t < x - nt || t > x + nt) continue;
                if (qa < 0) {
                    qa = q;
                } else if (qb < 0 && (s_rank[qa] & 1) != ph) {
                    qb = q;
                } else {
                    bad = true;     // two spikes of one phase cover the column: exact ties
                }
            }
        if (bad) {
            if (lane == 0) s_bad = 1;
            continue;
        }
        if (qa < 0) continue;       // no spike window covers this column
        float best;
        int bid;
        const ColSpike none = {0, 0, 0, 0.f};
        const ColSpike ca = col_spike(s_ref[qa], t, nt);
        const ColSpike cb = qb >= 0 ? col_spike(s_ref[qb], t, nt) : none;
        column_pass(Bt, ldbt, t, K, cc, s_nm, s_inv, s_pid, ca, cb, best, bid);
        if (lane == 0) {
            const bool live = t >= nt && t < T - nt;
            cfmax[t] = live ? best : 0.f;
            imax[t] = live ? bid : 0;
        }
    }

    // Xres[:, x-nt/2 .. x+nt/2] -= amp * (U[id]^T W) for the spikes owned by
    // this slab (their windows are disjoint: accepted spikes are > nt apart).
    // Lane = lag; the waves take channels of the template's nonzero range
    // (outside it the template is exactly zero and subtracting is a no-op).
    const int h = nt / 2;
    for (int q = 0; q < n; ++q) {
        const Slot sl = s_sp[q];
        if (sl.x < s0 || sl.x >= s1) continue;
        const int clo = crange[2 * sl.iy], chi = crange[2 * sl.iy + 1];
        const float *uj = U + (size_t)sl.iy * P * C;
        for (int kb0 = 0; kb0 < nt; kb0 += 64) {
            const int k = kb0 + lane;
            const bool kin = k < nt;
            float wk[8];
#pragma unroll
            for (int p = 0; p < 8; ++p) wk[p] = (p < P && kin) ? W[p * nt + k] : 0.f;
            float *xb = Xres + sl.x - h + k;
            for (int c0 = clo + wv; c0 < chi; c0 += kUpdWaves * 4) {
                float tm[4], xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = c0 + kUpdWaves * u;
                    tm[u] = 0.f;
                    xv[u] = 0.f;
                    if (c < chi) {
#pragma unroll
                        for (int p = 0; p < 8; ++p)
                            if (p < P) tm[u] = __builtin_fmaf(uj[(size_t)p * C + c], wk[p], tm[u]);
                        if (kin) xv[u] = xb[(size_t)c * ldx];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = c0 + kUpdWaves * u;
                    if (c < chi && kin) {
                        const float prod = sl.amp * tm[u];
                        xb[(size_t)c * ldx] = xv[u] - prod;
                    }
                }
            }
        }
    }
    __syncthreads();
    if (s_bad && tid == 0) {
        atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
        atomicExch(&state[ST_DONE], 1);
    }
}

// FP32 re-scoring of the accepted spikes (projection forms that compute B in
// reduced precision).  For the spike in slot q of chunk c: the score of its
// template at its column, recomputed from the current residual,
//   B = sum_c sum_k Xres[c, x + k - nt/2] * (sum_p U[id,p,c] W[p,k]),
// replaces the slot's amplitude (B/nm) and th_amp (sqrt(relu(B)^2/nm)) and
// the score entry itself, so that the subtraction and everything recorded
// for the spike use FP32 values.  Block per slot: lanes = lags, waves split
// the template's nonzero channel range.
__global__ __launch_bounds__(256) void peel_rescore_kernel(
    const float *__restrict__ Xres, int ldx, const float *__restrict__ U,
    const float *__restrict__ W, const int *__restrict__ crange, int P, int C, int nt,
    const float *__restrict__ nm, const int32_t *__restrict__ rank, float *__restrict__ Bt,
    int ldbt, const int *__restrict__ state, const int *__restrict__ chunk_count,
    Slot *__restrict__ slots) {
    if (state[ST_DONE]) return;
    const int c = blockIdx.x / kSlotCap, q = blockIdx.x % kSlotCap;
    if (q >= chunk_count[c]) return;
    __shared__ float s_part[4];
    Slot *sl = slots + (size_t)c * kSlotCap + q;
    const int x = sl->x, id = sl->iy;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int h = nt / 2;
    const int clo = crange[2 * id], chi = crange[2 * id + 1];
    const float *uj = U + (size_t)id * P * C;
    float part = 0.f;
    for (int kb0 = 0; kb0 < nt; kb0 += 64) {
        const int k = kb0 + lane;
        const bool kin = k < nt;
        float wk[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) wk[p] = (p < P && kin) ? W[p * nt + k] : 0.f;
        const float *xb = Xres + x - h + (kin ? k : 0);
        for (int c0 = clo + wv; c0 < chi; c0 += 4 * 8) {
            float xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int ch = c0 + 4 * u;
                xv[u] = (ch < chi) ? xb[(size_t)ch * ldx] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int ch = c0 + 4 * u;
                if (ch < chi) {
                    float tm = 0.f;
#pragma unroll
                    for (int p = 0; p < 8; ++p)
                        if (p < P) tm = __builtin_fmaf(uj[(size_t)p * C + ch], wk[p], tm);
                    part = __builtin_fmaf(kin ? xv[u] : 0.f, tm, part);
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if (lane == 0) s_part[wv] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = (s_part[0] + s_part[1]) + (s_part[2] + s_part[3]);
        const float n = nm[id];
        const float cf = score(b, n);
        sl->amp = b / n;
        sl->tha = sqrtf(cf);
        Bt[(size_t)x * ldbt + rank[id]] = b;
    }
}

// crange[2j], crange[2j+1] = first and one-past-last channel where template
// j has a nonzero component (an empty range is [0, 0)).  Wave per template.
__global__ __launch_bounds__(256) void channel_range_kernel(const float *__restrict__ U, int K,
                                                            int P, int C,
                                                            int *__restrict__ crange) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= K) return;
    const float *u = U + (size_t)j * P * C;
    int lo = C, hi = 0;
    for (int c0 = 0; c0 < C; c0 += 64) {
        const int c = c0 + lane;
        bool nz = false;
        if (c < C)
            for (int p = 0; p < P; ++p) nz |= u[(size_t)p * C + c] != 0.f;
        const unsigned long long m = __ballot(nz);
        if (m) {
            lo = min(lo, c0 + (int)__ffsll((long long)m) - 1);
            hi = c0 + 64 - __clzll((long long)m);
        }
    }
    if (lane == 0) {
        crange[2 * j] = lo < hi ? lo : 0;
        crange[2 * j + 1] = hi;
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
           ksm::align_up((size_t)L.n_state * 4, 256) + ksm::align_up(2 * (size_t)kMaxK * 4, 256) +
           256;
}

int ksm_peel(float *Bt, int ldbt, float *Xres, int ldx, const float *U, const float *W,
             const float *nm, const int32_t *perm, const int32_t *rank, const float *cc,
             const int32_t *lohi, const int64_t *off, int K, int P, int C, int T, int nt,
             float Th, int max_peels, int flags, int64_t *st, float *amps, float *th_amps,
             int capacity, int32_t *info, void *workspace, size_t ws_bytes, void *stream) {
    if (!Bt || !Xres || !U || !W || !nm || !perm || !rank || !cc || !lohi || !off || !st ||
        !amps || !th_amps || !info || K <= 0 || K > kMaxK || ldbt < K || (ldbt % 4) != 0 ||
        P <= 0 || P > 8 ||
        C <= 0 || nt <= 0 || nt > kMaxNt || (nt % 2) == 0 || T < 2 * nt + 1 || ldx < T ||
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
    int *crange = cv.take<int>(2 * (size_t)K);

    // Th^2 as the reference forms it (Python float, compared in float32)
    const double th2d = (double)Th * (double)Th;
    const float th2 = (float)th2d;
    // columns below th2_lo can never satisfy the acceptance test
    const float th2_lo = (float)(th2d - fmax(4e-9, fabs(th2d) * 1e-6));
    // the early-exit window scan is exact when every candidate is >= 2^-6
    const int early_exit = th2_lo >= 0.015625f ? 1 : 0;
    const size_t pos_shm = (size_t)K * (2 * sizeof(float) + sizeof(int));

    peel_init_kernel<<<1, 256, 0, s>>>(state, L.n_state);
    KSM_LAUNCH_CHECK("peel_init_kernel");
    channel_range_kernel<<<(K + 3) / 4, 256, 0, s>>>(U, K, P, C, crange);
    KSM_LAUNCH_CHECK("channel_range_kernel");
    {
        ksm::ProfileScope prof("colmax_all", ksm::kProfPeel, s);
        colmax_all_kernel<<<L.nslabs, kUpdThreads, pos_shm, s>>>(Bt, ldbt, nm, perm, K, T, nt,
                                                                 cfmax, imax);
    }
    KSM_LAUNCH_CHECK("colmax_all_kernel");
    for (int it = 0; it < max_peels; ++it) {
        {
            ksm::ProfileScope prof("peel_select", ksm::kProfPeel, s);
            peel_select_kernel<<<L.nchunks, kChunk, 0, s>>>(cfmax, imax, Bt, ldbt, nm, rank, T,
                                                            nt, th2, th2_lo, early_exit, state,
                                                            chunk_count, slots);
        }
        if (flags & KSM_PEEL_RESCORE) {
            ksm::ProfileScope prof("peel_rescore", ksm::kProfPeel, s);
            peel_rescore_kernel<<<L.nchunks * kSlotCap, 256, 0, s>>>(
                Xres, ldx, U, W, crange, P, C, nt, nm, rank, Bt, ldbt, state, chunk_count, slots);
        }
        {
            ksm::ProfileScope prof("peel_update", ksm::kProfPeel, s);
            peel_update_kernel<<<L.nslabs, kUpdThreads, pos_shm, s>>>(
                it, Bt, ldbt, Xres, ldx, U, W, crange, P, C, nm, perm, rank, cc, lohi, off, K, T,
                nt, L.nchunks, chunk_count, slots, state, st, amps, th_amps, capacity, cfmax,
                imax);
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
