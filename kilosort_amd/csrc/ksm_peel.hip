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
//   gkey[T][G]          best (score, template id) of each group of kGroup
//                        consecutive positions of each column, packed as one
//                        64-bit key (score bits high, ~id low: the larger key
//                        is the larger score, then the smaller id; NaN wins)
//   cfmax[T], imax[T]   best score and template id of every column
//                        (Cf = relu(B)^2/nm; columns < nt and >= T-nt are 0)
//   pnm, pinv, ppid     norm, rounded reciprocal and template id per position
//   slots               accepted spikes of the current iteration, per chunk
//                        of kChunk columns, in increasing time
//   state               done flag, status, iteration count, and kbase[it]
//                        (spikes recorded before iteration it)
//
// One iteration is two launches on the caller's stream:
//   peel_select  window max, threshold and ordered compaction into slots
//                (plus the FP32 re-scoring of reduced-precision scores);
//   peel_update  a work list over the accepted spikes, spread over every
//                wave of a grid that fills the chip: per spike, one item per
//                column of its +-nt window and kXresItems items for its
//                residual window.  A column item subtracts amp * ctc of the
//                spikes covering the column -- the even-ranked spike first,
//                as the reference's two interleaved phases do -- re-reduces
//                only the position groups those spikes touch, and folds the
//                group keys into the column's best score.
// Once an iteration accepts nothing the done flag makes every later launch
// return at once, so the host never waits inside the loop.
#include "ksm_common.h"
#include "ksm_profile.h"

namespace {

constexpr int kChunk = 256;         // columns per select block (one per thread)
constexpr int kSlotCap = 16;        // accepted spikes per chunk (>= 256/62 + ties)
constexpr int kGroup = 64;          // positions per group (16 lanes x 4)
constexpr int kMaxK = 8192;
constexpr int kMaxG = kMaxK / kGroup;
constexpr int kMaxNt = 255;
constexpr int kMaxChunks = 4096;    // T <= kChunk * kMaxChunks
constexpr int kUpdThreads = 256;    // update / first-pass block (4 waves)
constexpr int kXresItems = 4;       // residual work items per spike
constexpr int kXresCh = 8;          // channels per residual step

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
// then the smaller template id on equal values.
__device__ __forceinline__ bool beats(float v, int i, float bv, int bi) {
    const bool vn = v != v, bn = bv != bv;
    if (vn || bn) return vn && (!bn || i < bi);
    return v > bv || (v == bv && i < bi);
}

// relu as torch applies it: NaN passes through.
__device__ __forceinline__ float relu(float b) { return b < 0.f ? 0.f : b; }

// Cf = relu(b)^2 / nm, rounded as torch rounds relu(B)**2 / nm.
__device__ __forceinline__ float score(float b, float n) {
    const float r = relu(b);
    return (r * r) / n;
}

typedef unsigned long long u64;

// (score, template id) as one ordered key; NaN scores are made the canonical
// quiet NaN, which orders above +inf.  Scores are >= +0 or NaN.
__device__ __forceinline__ u64 make_key(float cf, int id) {
    const unsigned hi = cf != cf ? 0x7fc00000u : __float_as_uint(cf);
    return ((u64)hi << 32) | (unsigned)(~id);
}
__device__ __forceinline__ float key_score(u64 k) { return __uint_as_float((unsigned)(k >> 32)); }
__device__ __forceinline__ int key_id(u64 k) { return (int)(~(unsigned)k); }

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int o) {
    const unsigned lo = __shfl_xor((unsigned)v, o, 64);
    const unsigned hi = __shfl_xor((unsigned)(v >> 32), o, 64);
    return ((u64)hi << 32) | lo;
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
// the current column for positions [lo, hi) is cc[base + lo .. base + hi).
struct ColSpike {
    long long base;
    int lo, hi;
    float amp;
};

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Bits [glo, ghi] of 64-bit word w of a group mask (empty if glo > ghi).
__device__ __forceinline__ u64 range_bits(int glo, int ghi, int w) {
    const int lo = max(glo, 64 * w), hi = min(ghi, 64 * w + 63);
    if (lo > hi) return 0ull;
    return (~0ull >> (63 - (hi - lo))) << (lo - 64 * w);
}

// Column pass over position groups.  Applies the covering spikes' Gram
// updates to column t (products rounded before the subtract; spike a before
// spike b; an absent spike has an empty range), stores the changed scores,
// re-reduces every group that a spike touches (all groups when `all`), stores
// those group keys, and returns the column's best key (wave-uniform).
//
// 16 lanes take one group (4 consecutive positions each: 16-byte loads; ldbt,
// the Gram ranges and their offsets are multiples of 4), so a wave reduces
// four groups per round.  Within a group the scores are first ranked with a
// multiply by the rounded reciprocal of nm (within 3 ulp of the IEEE
// quotient); only positions whose approximate score is within 2^-20 of the
// group's approximate maximum can hold the exact maximum, and only those get
// the exact division.  A NaN in the group, or a group maximum below 2^-100,
// sends the whole group through the exact division.
__device__ __forceinline__ u64 column_groups(float *__restrict__ Bt, int ldbt, int t, int K,
                                             int G, const float *__restrict__ cc,
                                             const float *__restrict__ pnm,
                                             const float *__restrict__ pinv,
                                             const int *__restrict__ ppid, ColSpike a,
                                             ColSpike b, u64 *__restrict__ gk, bool all,
                                             u64 *s_key) {
    const int lane = threadIdx.x & 63;
    const int sg = lane >> 4, sl = lane & 15;
    float *col = Bt + (size_t)t * ldbt;
    const bool a_any = a.hi > a.lo, b_any = b.hi > b.lo;
    // groups to reduce (uniform masks over up to 128 groups)
    u64 m0, m1;
    if (all) {
        m0 = range_bits(0, G - 1, 0);
        m1 = range_bits(0, G - 1, 1);
    } else {
        const int ga0 = a_any ? a.lo / kGroup : 1, ga1 = a_any ? (a.hi - 1) / kGroup : 0;
        const int gb0 = b_any ? b.lo / kGroup : 1, gb1 = b_any ? (b.hi - 1) / kGroup : 0;
        m0 = range_bits(ga0, ga1, 0) | range_bits(gb0, gb1, 0);
        m1 = range_bits(ga0, ga1, 1) | range_bits(gb0, gb1, 1);
    }
    const u64 t0 = m0, t1 = m1;
    // keys of the groups left as they are (read before any store of this column)
    u64 old0 = 0ull, old1 = 0ull;
    if (!all) {
        if (lane < G && !((t0 >> lane) & 1ull)) old0 = gk[lane];
        if (lane + 64 < G && !((t1 >> lane) & 1ull)) old1 = gk[lane + 64];
    }
    while (m0 | m1) {
        int gs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (m0) {
                gs[j] = __builtin_ctzll(m0);
                m0 &= m0 - 1;
            } else if (m1) {
                gs[j] = 64 + __builtin_ctzll(m1);
                m1 &= m1 - 1;
            } else {
                gs[j] = -1;
            }
        }
        const int g = sg == 0 ? gs[0] : sg == 1 ? gs[1] : sg == 2 ? gs[2] : gs[3];
        if (g < 0) continue;
        const int r = g * kGroup + 4 * sl;
        const int ra = a_any ? min(max(r, a.lo), a.hi - 4) : 0;
        const int rb = b_any ? min(max(r, b.lo), b.hi - 4) : 0;
        const float4 v = *reinterpret_cast<const float4 *>(col + min(r, ldbt - 4));
        const float4 va = *reinterpret_cast<const float4 *>(cc + a.base + ra);
        const float4 vb = *reinterpret_cast<const float4 *>(cc + b.base + rb);
        const float4 n4 = *reinterpret_cast<const float4 *>(pnm + r);
        const float4 i4 = *reinterpret_cast<const float4 *>(pinv + r);
        const int4 p4 = *reinterpret_cast<const int4 *>(ppid + r);
        const bool ha = r >= a.lo && r < a.hi;
        const bool hb = r >= b.lo && r < b.hi;
        const float bv[4] = {v.x, v.y, v.z, v.w};
        const float ca[4] = {va.x, va.y, va.z, va.w};
        const float cb[4] = {vb.x, vb.y, vb.z, vb.w};
        const float nv[4] = {n4.x, n4.y, n4.z, n4.w};
        const float iv[4] = {i4.x, i4.y, i4.z, i4.w};
        const int pv[4] = {p4.x, p4.y, p4.z, p4.w};
        float y[4], ap[4];
        float amax = 0.f;
        bool nan_any = false;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float pa = a.amp * ca[e];
            const float pb = b.amp * cb[e];
            float x = bv[e];
            x = ha ? x - pa : x;
            x = hb ? x - pb : x;
            y[e] = x;
            const float rr = relu(x);
            ap[e] = (rr * rr) * iv[e];
            if (r + e < K) {
                nan_any |= ap[e] != ap[e];
                amax = fmaxf(amax, ap[e]);
            }
        }
        if (ha || hb) *reinterpret_cast<float4 *>(col + r) = make_float4(y[0], y[1], y[2], y[3]);
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
        const unsigned long long nb = __ballot(nan_any);
        const bool exact_all = ((nb >> (16 * sg)) & 0xffffull) != 0ull || amax < 0x1p-100f;
        const float thr = amax * (1.0f - 0x1p-20f);
        u64 key = 0ull;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (r + e < K && (exact_all || ap[e] >= thr)) {
                const u64 k = make_key(score(y[e], nv[e]), pv[e]);
                key = k > key ? k : key;
            }
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            const u64 k = shfl_xor_u64(key, o);
            key = k > key ? k : key;
        }
        if (sl == 0) {
            gk[g] = key;
            s_key[g] = key;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    u64 best = 0ull;
    if (lane < G) best = ((t0 >> lane) & 1ull) ? s_key[lane] : old0;
    if (lane + 64 < G) {
        const u64 k = ((t1 >> lane) & 1ull) ? s_key[lane + 64] : old1;
        best = k > best ? k : best;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const u64 k = shfl_xor_u64(best, o);
        best = k > best ? k : best;
    }
    return best;
}

// Per-position norms, rounded reciprocals and template ids, padded with
// (1, 0, 0) to G * kGroup positions.
__global__ void position_tables_kernel(const float *__restrict__ nm,
                                       const int32_t *__restrict__ perm, int K, int Kg,
                                       float *__restrict__ pnm, float *__restrict__ pinv,
                                       int *__restrict__ ppid) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Kg) return;
    if (r < K) {
        const int id = perm[r];
        const float n = nm[id];
        pnm[r] = n;
        pinv[r] = 1.0f / n;
        ppid[r] = id;
    } else {
        pnm[r] = 1.f;
        pinv[r] = 0.f;
        ppid[r] = 0;
    }
}

__global__ void peel_init_kernel(int *state, int n_state) {
    for (int e = threadIdx.x; e < n_state; e += blockDim.x) state[e] = 0;
}

// Group keys and best score of every live column (first pass, no updates);
// one wave per column.
__global__ __launch_bounds__(kUpdThreads) void colmax_all_kernel(
    float *__restrict__ Bt, int ldbt, const float *__restrict__ pnm,
    const float *__restrict__ pinv, const int *__restrict__ ppid, int K, int G, int T, int nt,
    float *__restrict__ cfmax, int *__restrict__ imax, u64 *__restrict__ gkey) {
    __shared__ u64 s_key[kUpdThreads / 64][kMaxG];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (kUpdThreads / 64) + wv;
    if (t >= T) return;
    if (t >= nt && t < T - nt) {
        const ColSpike none = {0, 0, 0, 0.f};
        const u64 best = column_groups(Bt, ldbt, t, K, G, Bt, pnm, pinv, ppid, none, none,
                                       gkey + (size_t)t * G, true, s_key[wv]);
        if (lane == 0) {
            cfmax[t] = key_score(best);
            imax[t] = key_id(best);
        }
    } else if (lane == 0) {
        cfmax[t] = 0.f;
        imax[t] = 0;
    }
}

// Accepts columns whose score is the maximum of its +-nt window and exceeds
// Th^2: (Cmax > Th^2) & (|Cmax - Cfmax| < 1e-9) with Cmax the window max
// (template_matching.py:205-212), and records them in time order.
//
// With `rescore` set (scores from a reduced-precision projection), each
// accepted spike is then re-scored in FP32 from the residual,
//   b = sum_c sum_k Xres[c, x + k - nt/2] * (sum_p U[id,p,c] W[p,k]),
// and b replaces the score entry, the amplitude (b/nm) and th_amp
// (sqrt(relu(b)^2/nm)), so that what is recorded and subtracted is FP32.
// Work split: lanes = lags, wave w takes channels clo+w, clo+w+4, ... of the
// template's nonzero channel range; partial sums meet in LDS.
__global__ __launch_bounds__(kChunk) void peel_select_kernel(
    const float *__restrict__ cfmax, const int *__restrict__ imax, float *__restrict__ Bt,
    int ldbt, const float *__restrict__ nm, const int32_t *__restrict__ rank, int T, int nt,
    float th2, float th2_lo, int early_exit, int rescore, const float *__restrict__ Xres,
    int ldx, const float *__restrict__ U, const float *__restrict__ W,
    const int *__restrict__ crange, int P, int C, int *__restrict__ state,
    int *__restrict__ chunk_count, Slot *__restrict__ slots) {
    if (state[ST_DONE]) return;
    __shared__ float win[kChunk + 2 * kMaxNt];
    __shared__ int s_wc[kChunk / 64];
    __shared__ int s_sx[kSlotCap], s_sid[kSlotCap];
    __shared__ float s_part[kSlotCap][kChunk / 64];
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
            s_sx[pos] = x;
            s_sid[pos] = iy;
        }
    }
    if (tid == 0) {
        if (count > kSlotCap) {
            atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
            atomicExch(&state[ST_DONE], 1);
        }
        chunk_count[blockIdx.x] = count < kSlotCap ? count : kSlotCap;
    }
    if (!rescore || count == 0) return;      // block-uniform
    const int n = count < kSlotCap ? count : kSlotCap;
    __syncthreads();
    const int h = nt / 2;
    for (int q = 0; q < n; ++q) {
        const int sx = uni(s_sx[q]), id = uni(s_sid[q]);
        const int clo = uni(crange[2 * id]), chi = uni(crange[2 * id + 1]);
        const float *uj = U + (size_t)id * P * C;
        float part = 0.f;
        for (int kb0 = 0; kb0 < nt; kb0 += 64) {
            const int k = kb0 + lane;
            const bool kin = k < nt;
            float wk[8];
#pragma unroll
            for (int p = 0; p < 8; ++p) wk[p] = (p < P && kin) ? W[p * nt + k] : 0.f;
            const float *xb = Xres + sx - h + (kin ? k : 0);
            // wave w: channels [cb, cb + 16) with cb = clo + 16 w + 64 i (uniform,
            // so the template coefficients are scalar loads); loads are clamped
            // to the last channel and masked, with no branch between them
            for (int cb = clo + 16 * uni(wave); cb < chi; cb += 64) {
                float xv[16], tm[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int ch = cb + u < chi ? cb + u : chi - 1;
                    xv[u] = xb[(size_t)ch * ldx];
                    tm[u] = 0.f;
                }
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    if (p < P) {
#pragma unroll
                        for (int u = 0; u < 16; ++u) {
                            const int ch = cb + u < chi ? cb + u : chi - 1;
                            tm[u] = __builtin_fmaf(uj[(size_t)p * C + ch], wk[p], tm[u]);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    part = __builtin_fmaf((kin && cb + u < chi) ? xv[u] : 0.f, tm[u], part);
            }
        }
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
        if (lane == 0) s_part[q][wave] = part;
    }
    __syncthreads();
    if (tid < n) {
        const int sx = s_sx[tid], id = s_sid[tid];
        const float b = (s_part[tid][0] + s_part[tid][1]) + (s_part[tid][2] + s_part[tid][3]);
        const float nmv = nm[id];
        Slot *sl = slots + (size_t)blockIdx.x * kSlotCap + tid;
        sl->amp = b / nmv;
        sl->tha = sqrtf(score(b, nmv));
        Bt[(size_t)sx * ldbt + rank[id]] = b;
    }
}

// Spike q of the current iteration (time order) from the per-chunk slots;
// pre[c] = spikes in chunks before c (block-shared).
__device__ __forceinline__ Slot spike_at(int q, const int *pre, int nchunks,
                                         const Slot *__restrict__ slots) {
    int lo = 0, hi = nchunks - 1;         // largest c with pre[c] <= q
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return slots[(size_t)lo * kSlotCap + (q - pre[lo])];
}

// A spike of the current iteration with its Gram block, wave-uniform.
struct SpikeG {
    int x, lo, hi;
    float amp;
    long long off;
};

__device__ __forceinline__ SpikeG spike_gram(const Slot &sl, const int32_t *__restrict__ rank,
                                             const int32_t *__restrict__ lohi,
                                             const int64_t *__restrict__ off) {
    const int jr = uni(rank[uni(sl.iy)]);
    SpikeG g;
    g.x = uni(sl.x);
    g.lo = uni(lohi[2 * jr]);
    g.hi = uni(lohi[2 * jr + 1]);
    g.amp = uni(sl.amp);
    g.off = uni((long long)off[jr]);
    return g;
}

// The spike's Gram row for column t (lag t - x + nt).
__device__ __forceinline__ ColSpike col_spike(const SpikeG &g, int t, int nt) {
    ColSpike cs;
    cs.lo = g.lo;
    cs.hi = g.hi;
    cs.amp = g.amp;
    cs.base = g.off + (long long)(t - g.x + nt) * (g.hi - g.lo) - g.lo;
    return cs;
}

__global__ __launch_bounds__(kUpdThreads) void peel_update_kernel(
    int it, float *__restrict__ Bt, int ldbt, float *__restrict__ Xres, int ldx,
    const float *__restrict__ U, const float *__restrict__ W, const int *__restrict__ crange,
    int P, int C, const float *__restrict__ pnm, const float *__restrict__ pinv,
    const int *__restrict__ ppid, const int32_t *__restrict__ rank,
    const float *__restrict__ cc, const int32_t *__restrict__ lohi,
    const int64_t *__restrict__ off, int K, int G, int T, int nt, int nchunks,
    const int *__restrict__ chunk_count, const Slot *__restrict__ slots,
    int *__restrict__ state, int64_t *__restrict__ st, float *__restrict__ amps,
    float *__restrict__ th_amps, int capacity, float *__restrict__ cfmax,
    int *__restrict__ imax, u64 *__restrict__ gkey) {
    if (state[ST_DONE]) return;
    extern __shared__ int s_pre[];          // nchunks + 1 entries
    __shared__ int s_wsum[kUpdThreads / 64];
    __shared__ u64 s_key[kUpdThreads / 64][kMaxG];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);

    // exclusive prefix of the per-chunk spike counts (thread: consecutive chunks)
    const int per_t = (nchunks + kUpdThreads - 1) / kUpdThreads;
    const int c0 = tid * per_t, c1 = min(c0 + per_t, nchunks);
    int mine = 0;
    for (int c = c0; c < c1; ++c) mine += chunk_count[c];
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; ++w) base += s_wsum[w];
    int run = base + incl - mine;
    for (int c = c0; c < c1; ++c) {
        s_pre[c] = run;
        run += chunk_count[c];
    }
    const int S = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    if (tid == 0) s_pre[nchunks] = S;
    __syncthreads();

    if (S == 0) {
        if (blockIdx.x == 0 && tid == 0) atomicExch(&state[ST_DONE], 1);
        return;
    }
    const int kb = state[ST_KBASE + it];
    if (kb + S > capacity) {
        if (blockIdx.x == 0 && tid == 0) {
            atomicExch(&state[ST_STATUS], KSM_ERR_CAPACITY);
            atomicExch(&state[ST_DONE], 1);
        }
        return;
    }
    if (blockIdx.x == 0 && tid == 0) {
        state[ST_KBASE + it + 1] = kb + S;
        state[ST_ITERS] = it + 1;
    }

    // work items: per spike, 2nt+1 columns then kXresItems residual items;
    // each wave takes one contiguous range of items
    const int ncol = 2 * nt + 1;
    const int per_spike = ncol + kXresItems;
    const long long n_items = (long long)S * per_spike;
    const long long nwaves = (long long)gridDim.x * (kUpdThreads / 64);
    const long long gw = (long long)blockIdx.x * (kUpdThreads / 64) + wv;
    const long long per = (n_items + nwaves - 1) / nwaves;
    const long long i0 = gw * per, i1 = min(n_items, i0 + per);
    int cq = -1;
    Slot sq = {0, 0, 0.f, 0.f};
    SpikeG gq = {0, 0, 0, 0.f, 0}, gn = gq;
    bool has_p = false, has_n = false, has_nn = false;
    int x_p = 0, x_nn = 0;
    const int h = nt / 2;
    for (long long i = i0; i < i1; ++i) {
        const int q = (int)(i / per_spike), d = (int)(i % per_spike);
        if (q != cq) {
            cq = q;
            sq = spike_at(q, s_pre, nchunks, slots);
            has_p = q > 0;
            if (has_p) x_p = uni(spike_at(q - 1, s_pre, nchunks, slots).x);
            has_n = q + 1 < S;
            has_nn = q + 2 < S;
            if (has_nn) x_nn = uni(spike_at(q + 2, s_pre, nchunks, slots).x);
            gq = spike_gram(sq, rank, lohi, off);
            if (has_n) gn = spike_gram(spike_at(q + 1, s_pre, nchunks, slots), rank, lohi, off);
        }
        if (d < ncol) {
            // column t, owned by the earliest spike whose window covers it
            const int t = gq.x - nt + d;
            if (t < nt || t >= T - nt) continue;        // never scored again
            if (has_p && t <= x_p + nt) continue;       // spike q-1 owns it
            const bool with_n = has_n && t >= gn.x - nt;
            if (has_nn && t >= x_nn - nt) {
                // three spikes within 2nt+1 columns: only exact score ties
                // place accepted spikes this close
                if (lane == 0) {
                    atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
                    atomicExch(&state[ST_DONE], 1);
                }
                continue;
            }
            const ColSpike none = {0, 0, 0, 0.f};
            const ColSpike cq_ = col_spike(gq, t, nt);
            const ColSpike cn_ = with_n ? col_spike(gn, t, nt) : none;
            // the even-ranked spike of the pair is subtracted first
            const bool q_even = (q & 1) == 0;
            const u64 best = column_groups(Bt, ldbt, t, K, G, cc, pnm, pinv, ppid,
                                           q_even ? cq_ : cn_, q_even ? cn_ : cq_,
                                           gkey + (size_t)t * G, false, s_key[wv]);
            if (lane == 0) {
                cfmax[t] = key_score(best);
                imax[t] = key_id(best);
            }
        } else {
            // residual: Xres[:, x-nt/2 .. x+nt/2] -= amp * (U[id]^T W) over the
            // template's nonzero channel range (the windows of one iteration's
            // spikes are disjoint); lanes = lags, kXresCh channels per step
            const int k = d - ncol;
            const int id = uni(sq.iy), x = uni(sq.x);
            const float amp = uni(sq.amp);
            if (k == 0 && lane == 0) {
                const size_t pos = (size_t)kb + q;
                st[2 * pos] = x;
                st[2 * pos + 1] = id;
                amps[pos] = amp;
                th_amps[pos] = sq.tha;
            }
            const int clo = uni(crange[2 * id]), chi = uni(crange[2 * id + 1]);
            const float *uj = U + (size_t)id * P * C;
            for (int kb0 = 0; kb0 < nt; kb0 += 64) {
                const int kk = kb0 + lane;
                const bool kin = kk < nt;
                float wk[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) wk[p] = (p < P && kin) ? W[p * nt + kk] : 0.f;
                float *xb = Xres + x - h + (kin ? kk : 0);
                for (int cb = clo + kXresCh * k; cb < chi; cb += kXresCh * kXresItems) {
                    float tm[kXresCh], xv[kXresCh];
#pragma unroll
                    for (int u = 0; u < kXresCh; ++u) {
                        const int c = cb + u < chi ? cb + u : chi - 1;
                        xv[u] = xb[(size_t)c * ldx];
                        tm[u] = 0.f;
                    }
#pragma unroll
                    for (int p = 0; p < 8; ++p) {
                        if (p < P) {
#pragma unroll
                            for (int u = 0; u < kXresCh; ++u) {
                                const int c = cb + u < chi ? cb + u : chi - 1;
                                tm[u] = __builtin_fmaf(uj[(size_t)p * C + c], wk[p], tm[u]);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kXresCh; ++u) {
                        if (cb + u < chi && kin) {
                            const float prod = amp * tm[u];
                            xb[(size_t)(cb + u) * ldx] = xv[u] - prod;
                        }
                    }
                }
            }
        }
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
    int nchunks, n_state, G, Kg;
};

PeelLayout peel_layout(int K, int T, int max_peels) {
    PeelLayout L;
    L.nchunks = (T + kChunk - 1) / kChunk;
    L.n_state = ST_KBASE + max_peels + 1;
    L.G = (K + kGroup - 1) / kGroup;
    L.Kg = L.G * kGroup;
    return L;
}

// Blocks of the update kernel that fit on the device at once, for `dyn`
// bytes of dynamic LDS (cached per device and size).
int update_grid(size_t dyn) {
    static thread_local int c_dev = -1, c_grid = 0;
    static thread_local size_t c_dyn = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev == c_dev && dyn == c_dyn && c_grid > 0) return c_grid;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_update_kernel, kUpdThreads,
                                                     dyn) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    c_dev = dev;
    c_dyn = dyn;
    c_grid = cus * per_cu;
    return c_grid;
}

}  // namespace

extern "C" {

size_t ksm_peel_workspace(int K, int T, int max_peels) {
    if (K <= 0 || T <= 0 || max_peels < 0) return 0;
    const PeelLayout L = peel_layout(K, T, max_peels);
    return ksm::align_up((size_t)T * L.G * 8, 256) + ksm::align_up((size_t)T * 4, 256) * 2 +
           ksm::align_up((size_t)L.Kg * 4, 256) * 3 + ksm::align_up(L.nchunks * 4, 256) +
           ksm::align_up((size_t)L.nchunks * kSlotCap * sizeof(Slot), 256) +
           ksm::align_up((size_t)L.n_state * 4, 256) + ksm::align_up(2 * (size_t)K * 4, 256) +
           256;
}

int ksm_peel(float *Bt, int ldbt, float *Xres, int ldx, const float *U, const float *W,
             const float *nm, const int32_t *perm, const int32_t *rank, const float *cc,
             const int32_t *lohi, const int64_t *off, int K, int P, int C, int T, int nt,
             float Th, int max_peels, int flags, int64_t *st, float *amps, float *th_amps,
             int capacity, int32_t *info, void *workspace, size_t ws_bytes, void *stream) {
    if (!Bt || !Xres || !U || !W || !nm || !perm || !rank || !cc || !lohi || !off || !st ||
        !amps || !th_amps || !info || K <= 0 || K > kMaxK || ldbt < K || (ldbt % 4) != 0 ||
        P <= 0 || P > 8 || C <= 0 || nt <= 0 || nt > kMaxNt || (nt % 2) == 0 ||
        T < 2 * nt + 1 || T > kChunk * kMaxChunks || ldx < T || max_peels < 0 ||
        capacity < 0 || (flags & ~KSM_PEEL_RESCORE) != 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_peel: bad arguments");
    if (!workspace || ws_bytes < ksm_peel_workspace(K, T, max_peels))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_peel: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const PeelLayout L = peel_layout(K, T, max_peels);
    ksm::Carver cv(workspace);
    u64 *gkey = cv.take<u64>((size_t)T * L.G);
    float *cfmax = cv.take<float>(T);
    int *imax = cv.take<int>(T);
    float *pnm = cv.take<float>(L.Kg);
    float *pinv = cv.take<float>(L.Kg);
    int *ppid = cv.take<int>(L.Kg);
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
    const size_t upd_dyn = (L.nchunks + 1) * sizeof(int);
    const int grid = update_grid(upd_dyn);

    peel_init_kernel<<<1, 256, 0, s>>>(state, L.n_state);
    KSM_LAUNCH_CHECK("peel_init_kernel");
    channel_range_kernel<<<(K + 3) / 4, 256, 0, s>>>(U, K, P, C, crange);
    KSM_LAUNCH_CHECK("channel_range_kernel");
    position_tables_kernel<<<(L.Kg + 255) / 256, 256, 0, s>>>(nm, perm, K, L.Kg, pnm, pinv,
                                                               ppid);
    KSM_LAUNCH_CHECK("position_tables_kernel");
    {
        ksm::ProfileScope prof("colmax_all", ksm::kProfPeel, s);
        colmax_all_kernel<<<(T + 3) / 4, kUpdThreads, 0, s>>>(Bt, ldbt, pnm, pinv, ppid, K,
                                                              L.G, T, nt, cfmax, imax, gkey);
    }
    KSM_LAUNCH_CHECK("colmax_all_kernel");
    for (int it = 0; it < max_peels; ++it) {
        {
            ksm::ProfileScope prof("peel_select", ksm::kProfPeel, s);
            peel_select_kernel<<<L.nchunks, kChunk, 0, s>>>(
                cfmax, imax, Bt, ldbt, nm, rank, T, nt, th2, th2_lo, early_exit,
                (flags & KSM_PEEL_RESCORE) ? 1 : 0, Xres, ldx, U, W, crange, P, C, state,
                chunk_count, slots);
        }
        {
            ksm::ProfileScope prof("peel_update", ksm::kProfPeel, s);
            peel_update_kernel<<<grid, kUpdThreads, upd_dyn, s>>>(
                it, Bt, ldbt, Xres, ldx, U, W, crange, P, C, pnm, pinv, ppid, rank, cc, lohi,
                off, K, L.G, T, nt, L.nchunks, chunk_count, slots, state, st, amps, th_amps,
                capacity, cfmax, imax, gkey);
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
