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
//   pnm, ppid           norm and template id per position
//   slots               accepted spikes of the current iteration, per chunk
//                        of kChunk columns, in increasing time
//   state               done flag, status, iteration count, and kbase[it]
//                        (spikes recorded before iteration it)
//
// One iteration is two launches on the caller's stream:
//   peel_select  window max, threshold and ordered compaction into slots
//                (for reduced-precision scores peel_select_cert: decisions
//                on their error bound, open ones on FP32 re-scores);
//   peel_update  a work list over the accepted spikes, spread over every
//                wave of a grid that fills the chip: per spike, one item per
//                column of its +-nt window and kXresItems items for its
//                residual window.  A column item subtracts amp * ctc of the
//                spikes covering the column -- the even-ranked spike first,
//                as the reference's two interleaved phases do -- re-reduces
//                only the position groups those spikes touch, and folds the
//                group keys into the column's best score.
// Once an iteration accepts nothing the done flag makes every later launch
// return at once, so results never depend on the host: it follows the
// device's progress words only to stay at most kLag iterations ahead and to
// stop enqueuing once the loop has ended.
#include <chrono>
#include <mutex>
#include <thread>

#include "ksm_common.h"
#include "ksm_bound.h"
#include "ksm_profile.h"
#include "ksm_graph.h"

namespace {

constexpr int kChunk = 256;         // columns per select block (one per thread)
constexpr int kSlotCap = 256;       // accepted spikes per chunk: every column of it
static_assert(kSlotCap >= kChunk, "a chunk can accept each of its columns (exact ties)");
constexpr int kGroup = 32;          // positions per group (8 lanes x 4)
constexpr int kPassGroups = 8;      // groups a wave reduces per pass
constexpr int kMaxK = 8192;
constexpr int kMaxNt = 255;
constexpr int kMaxC = 65535;        // channel ranges are packed in 16 bits
constexpr int kMaxChunks = 4096;    // T <= kChunk * kMaxChunks
constexpr int kUpdThreads = 256;    // update / first-pass block (4 waves)
constexpr int kXresItems = 8;       // residual work items per spike
constexpr int kXresCh = 16;         // channels per residual step (loads first)

enum { ST_DONE = 0, ST_STATUS = 1, ST_ITERS = 2, ST_RESCORED = 3, ST_KBASE = 4 };

struct Slot {
    int x;          // column in the padded batch
    int iy;         // template id
    float amp;      // B[iy, x] / nm[iy]
    float tha;      // sqrt(Cmax[x])
    unsigned lohi;  // the template's Gram block rows lo | hi << 16 (see GramRef)
    unsigned crg;   // its nonzero channels clo | chi << 16
    long long off;
};

// Per template id: its position in the peel order, its norm and its Gram
// block: rows [lo, hi) of the peel order at cc[off + lag * (hi - lo) + (r - lo)].
struct GramRef {
    int rank;
    float nm;
    int lo, hi;
    long long off;
    int clo, chi;   // nonzero channels [clo, chi) of the template
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

// A progress word for the host, in pinned host memory (system scope).
__device__ __forceinline__ void publish(u64 *p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

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

// Cross-lane steps on DPP (no LDS round trip).  quad_perm [1,0,3,2],
// quad_perm [2,3,0,1] and row_half_mirror applied with a commutative
// reduction leave every lane of an 8-lane group holding the group's result;
// row_mirror then joins the two groups of a 16-lane row.
template <int CTRL> __device__ __forceinline__ unsigned dpp_u32(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ u64 dpp_max_step(u64 k) {
    const unsigned lo = dpp_u32<CTRL>((unsigned)k);
    const unsigned hi = dpp_u32<CTRL>((unsigned)(k >> 32));
    const u64 o = ((u64)hi << 32) | lo;
    return o > k ? o : k;
}
__device__ __forceinline__ u64 group_max_u64(u64 k) {
    k = dpp_max_step<0xb1>(k);
    k = dpp_max_step<0x4e>(k);
    k = dpp_max_step<0x141>(k);
    return k;
}
// Maximum over the wave (uniform result): 16-lane rows on DPP, then the four
// row results through scalar registers.
__device__ __forceinline__ u64 wave_max_u64(u64 k) {
    k = group_max_u64(k);
    k = dpp_max_step<0x140>(k);
    u64 best = 0ull;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)k, 16 * r);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(k >> 32), 16 * r);
        const u64 o = ((u64)hi << 32) | lo;
        best = o > best ? o : best;
    }
    return best;
}

// The groups a column pass reduces, as up to three intervals: the first
// spike's groups, then the second spike's groups below and above them.
struct GroupList {
    int lo0, n0, lo1, n1, lo2, n2;
    __device__ __forceinline__ int count() const { return n0 + n1 + n2; }
    __device__ __forceinline__ int at(int j) const {
        return j < n0 ? lo0 + j : j < n0 + n1 ? lo1 + (j - n0) : lo2 + (j - n0 - n1);
    }
    __device__ __forceinline__ bool has(int g) const {
        return (g >= lo0 && g < lo0 + n0) || (g >= lo1 && g < lo1 + n1) ||
               (g >= lo2 && g < lo2 + n2);
    }
};

__device__ __forceinline__ GroupList group_list(const ColSpike &a, const ColSpike &b, int G,
                                                bool all) {
    GroupList L = {0, 0, 0, 0, 0, 0};
    if (all) {
        L.n0 = G;
        return L;
    }
    int a0 = 0, a1 = -1;
    if (a.hi > a.lo) {
        a0 = a.lo / kGroup;
        a1 = (a.hi - 1) / kGroup;
        L.lo0 = a0;
        L.n0 = a1 - a0 + 1;
    }
    if (b.hi > b.lo) {
        const int b0 = b.lo / kGroup, b1 = (b.hi - 1) / kGroup;
        if (a1 < a0) {
            L.lo1 = b0;
            L.n1 = b1 - b0 + 1;
        } else {
            const int h1 = min(b1, a0 - 1);
            if (h1 >= b0) {
                L.lo1 = b0;
                L.n1 = h1 - b0 + 1;
            }
            const int l2 = max(b0, a1 + 1);
            if (b1 >= l2) {
                L.lo2 = l2;
                L.n2 = b1 - l2 + 1;
            }
        }
    }
    return L;
}

// A column to update: its index and the spikes covering it, the one
// subtracted first in `a` (an absent spike has an empty range).
struct ColJob {
    int t;
    ColSpike a, b;
};

// Loads of one pass (kPassGroups groups) over one column: lanes 8 g .. 8 g + 7
// take group g of the pass, 4 consecutive positions each (16-byte loads;
// ldbt, the Gram ranges and their offsets are multiples of 4).  An absent
// group reads group 0 and is discarded.
struct ColPass {
    float4 v, va, vb;
    int gg;
};

__device__ __forceinline__ void pass_load(ColPass &ps, const ColJob &J, const GroupList &L,
                                          int j0, const float *__restrict__ Bt, int ldbt,
                                          const float *__restrict__ cc) {
    const int lane = threadIdx.x & 63;
    const int j = j0 + (lane >> 3);
    ps.gg = j < L.count() ? L.at(j) : -1;
    const int r = (ps.gg < 0 ? 0 : ps.gg) * kGroup + 4 * (lane & 7);
    const bool a_any = J.a.hi > J.a.lo, b_any = J.b.hi > J.b.lo;
    const int ra = a_any ? min(max(r, J.a.lo), J.a.hi - 4) : 0;
    const int rb = b_any ? min(max(r, J.b.lo), J.b.hi - 4) : 0;
    ps.v = *reinterpret_cast<const float4 *>(Bt + (size_t)J.t * ldbt + min(r, ldbt - 4));
    ps.va = *reinterpret_cast<const float4 *>(cc + J.a.base + ra);
    ps.vb = *reinterpret_cast<const float4 *>(cc + J.b.base + rb);
}

// Applies the Gram updates (products rounded before the subtract; a before
// b) in place in ps.v and reduces the lane's group to its key (exact score
// relu(b)^2/nm; larger score, then smaller template id; NaN wins).  s_nm,
// s_pid: per-position norms and template ids.
__device__ __forceinline__ u64 pass_compute(ColPass &ps, const ColJob &J, int K, const float *s_nm,
                                            const int *s_pid) {
    const int r = (ps.gg < 0 ? 0 : ps.gg) * kGroup + 4 * (threadIdx.x & 7);
    const bool ha = r >= J.a.lo && r < J.a.hi;
    const bool hb = r >= J.b.lo && r < J.b.hi;
    const float4 n4 = *reinterpret_cast<const float4 *>(s_nm + r);
    const int4 p4 = *reinterpret_cast<const int4 *>(s_pid + r);
    float bv[4] = {ps.v.x, ps.v.y, ps.v.z, ps.v.w};
    const float ca[4] = {ps.va.x, ps.va.y, ps.va.z, ps.va.w};
    const float cb[4] = {ps.vb.x, ps.vb.y, ps.vb.z, ps.vb.w};
    const float nv[4] = {n4.x, n4.y, n4.z, n4.w};
    const int pv[4] = {p4.x, p4.y, p4.z, p4.w};
    u64 kk = 0ull;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float pa = J.a.amp * ca[e];
        const float pb = J.b.amp * cb[e];
        float x = bv[e];
        x = ha ? x - pa : x;
        x = hb ? x - pb : x;
        bv[e] = x;
        const u64 k = r + e < K ? make_key(score(x, nv[e]), pv[e]) : 0ull;
        kk = k > kk ? k : kk;
    }
    ps.v = make_float4(bv[0], bv[1], bv[2], bv[3]);
    return group_max_u64(ps.gg < 0 ? 0ull : kk);
}

__device__ __forceinline__ void pass_store(const ColPass &ps, const ColJob &J,
                                           float *__restrict__ Bt, int ldbt,
                                           u64 *__restrict__ gk, u64 key, u64 &acc) {
    const int sl = threadIdx.x & 7;
    const int r = (ps.gg < 0 ? 0 : ps.gg) * kGroup + 4 * sl;
    const bool dirty = ps.gg >= 0 && ((r >= J.a.lo && r < J.a.hi) || (r >= J.b.lo && r < J.b.hi));
    if (dirty) *reinterpret_cast<float4 *>(Bt + (size_t)J.t * ldbt + r) = ps.v;
    if (ps.gg >= 0 && sl == 0) gk[ps.gg] = key;
    acc = key > acc ? key : acc;
}

// Keys of the groups a column pass leaves as they are (read before any store).
__device__ __forceinline__ u64 kept_keys(const GroupList &L, int G, const u64 *__restrict__ gk) {
    const int lane = threadIdx.x & 63;
    u64 acc = 0ull;
    for (int g = lane; g < G; g += 64)
        if (!L.has(g)) {
            const u64 k = gk[g];
            acc = k > acc ? k : acc;
        }
    return acc;
}

// A pass in flight: its column job, the column's group list, the pass's
// first group j0, the kept keys (first pass of a column) and its loads.
struct PassWork {
    ColJob J;
    GroupList L;
    int j0;
    u64 kept;
    ColPass ps;
};

__device__ __forceinline__ void pass_begin(PassWork &w, const ColJob &J, const GroupList &L,
                                           int j0, bool all, int G,
                                           const float *__restrict__ Bt, int ldbt,
                                           const float *__restrict__ cc,
                                           const u64 *__restrict__ gkey) {
    w.J = J;
    w.L = L;
    w.j0 = j0;
    w.kept = (all || j0 > 0) ? 0ull : kept_keys(L, G, gkey + (size_t)J.t * G);
    pass_load(w.ps, J, L, j0, Bt, ldbt, cc);
}

// Column passes as a pipeline over the columns `src` yields: each pass's
// loads are issued before the previous pass's stores (a store issued ahead of
// a load would hold that load's wait on gfx9, whose counter covers both), a
// column with more than kPassGroups groups simply contributes several
// consecutive passes, and after a column's last pass its best (score,
// template id) -- over its re-reduced and kept group keys -- goes to
// cfmax / imax.
template <class Source>
__device__ __forceinline__ void column_pipeline(Source &src, bool all, float *__restrict__ Bt,
                                                int ldbt, int K, int G,
                                                const float *__restrict__ cc, const float *s_nm,
                                                const int *s_pid, u64 *__restrict__ gkey,
                                                float *__restrict__ cfmax,
                                                int *__restrict__ imax) {
    ColJob J;
    if (!src.next(J)) return;
    GroupList L = group_list(J.a, J.b, G, all);
    int j0 = 0;
    PassWork a, b;
    pass_begin(a, J, L, 0, all, G, Bt, ldbt, cc, gkey);
    u64 acc = 0ull;
    for (;;) {
        const u64 key = pass_compute(a.ps, a.J, K, s_nm, s_pid);
        // the next pass: the same column's next groups, or the next column
        bool more = true;
        if (j0 + kPassGroups < L.count()) {
            j0 += kPassGroups;
        } else if (src.next(J)) {
            L = group_list(J.a, J.b, G, all);
            j0 = 0;
        } else {
            more = false;
        }
        if (more) pass_begin(b, J, L, j0, all, G, Bt, ldbt, cc, gkey);
        if (a.j0 == 0) acc = a.kept;
        pass_store(a.ps, a.J, Bt, ldbt, gkey + (size_t)a.J.t * G, key, acc);
        if (a.j0 + kPassGroups >= a.L.count()) {
            const u64 best = wave_max_u64(acc);
            if ((threadIdx.x & 63) == 0) {
                cfmax[a.J.t] = key_score(best);
                imax[a.J.t] = key_id(best);
            }
        }
        if (!more) break;
        a = b;
    }
}

// Per-position norms and template ids (padded to Kg) into LDS.
__device__ __forceinline__ void stage_tables(const float *__restrict__ pnm,
                                             const int *__restrict__ ppid, int Kg, float *s_nm,
                                             int *s_pid) {
    for (int r = threadIdx.x; r < Kg; r += blockDim.x) {
        s_nm[r] = pnm[r];
        s_pid[r] = ppid[r];
    }
}

// Per-position norms and template ids, padded with (1, 0) to G * kGroup
// positions.
__global__ void position_tables_kernel(const float *__restrict__ nm,
                                       const int32_t *__restrict__ perm, int K, int Kg,
                                       float *__restrict__ pnm, int *__restrict__ ppid) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Kg) return;
    if (r < K) {
        const int id = perm[r];
        pnm[r] = nm[id];
        ppid[r] = id;
    } else {
        pnm[r] = 1.f;
        ppid[r] = 0;
    }
}

__global__ void peel_init_kernel(int *state, int n_state) {
    for (int e = threadIdx.x; e < n_state; e += blockDim.x) state[e] = 0;
}

// Group keys and best score of every live column (first pass, no updates);
// a block takes kFirstCols columns, a wave every fourth one, pipelined.
constexpr int kFirstCols = 64;
struct FirstPassColumns {
    int t, hi, step;
    __device__ __forceinline__ bool next(ColJob &J) {
        if (t >= hi) return false;
        const ColSpike none = {0, 0, 0, 0.f};
        J.t = t;
        J.a = none;
        J.b = none;
        t += step;
        return true;
    }
};
__global__ __launch_bounds__(kUpdThreads) void colmax_all_kernel(
    float *__restrict__ Bt, int ldbt, const float *__restrict__ pnm,
    const int *__restrict__ ppid, int K, int G, int Kg, int T, int nt,
    float *__restrict__ cfmax, int *__restrict__ imax, u64 *__restrict__ gkey) {
    extern __shared__ float s_tab[];        // s_nm[Kg], s_pid[Kg]
    float *s_nm = s_tab;
    int *s_pid = reinterpret_cast<int *>(s_tab + Kg);
    stage_tables(pnm, ppid, Kg, s_nm, s_pid);
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const ColSpike none = {0, 0, 0, 0.f};
    const int t0 = blockIdx.x * kFirstCols;
    const int tend = min(t0 + kFirstCols, T);
    for (int t = t0 + wv; t < tend; t += kUpdThreads / 64)
        if (lane == 0 && (t < nt || t >= T - nt)) {
            cfmax[t] = 0.f;
            imax[t] = 0;
        }
    // live columns of the block handled by this wave
    FirstPassColumns src;
    src.t = max(t0, nt) + wv;
    src.hi = min(tend, T - nt);
    src.step = kUpdThreads / 64;
    column_pipeline(src, true, Bt, ldbt, K, G, Bt, s_nm, s_pid, gkey, cfmax, imax);
}

// Accepts columns whose score is the maximum of its +-nt window and exceeds
// Th^2: (Cmax > Th^2) & (|Cmax - Cfmax| < 1e-9) with Cmax the window max
// (template_matching.py:205-212), and records them in time order.  The window
// max of the chunk's columns comes from maxima of 2^j consecutive values built
// by doubling in LDS (max_pool1d semantics: -inf padding, NaN propagates).
// The scores are FP32 (the reduced-precision forms use
// peel_select_cert_kernel below).
// NaN-propagating maximum (max_pool1d semantics).
__device__ __forceinline__ float nan_max(float a, float b) {
    return a != a ? a : (b != b ? b : fmaxf(a, b));
}

template <int NTH>
__global__ __launch_bounds__(NTH) void peel_select_kernel(
    const float *__restrict__ cfmax, const int *__restrict__ imax, float *__restrict__ Bt,
    int ldbt, const GramRef *__restrict__ gref, int T, int nt, float th2,
    const int32_t *__restrict__ rank, int K, int *__restrict__ state,
    int *__restrict__ chunk_count, Slot *__restrict__ slots) {
    // every load of the prologue is issued together: the done flag, the
    // window, the column's template and the rank table
    const int done = state[ST_DONE];
    extern __shared__ int s_rank[];          // position in the peel order, by id
    __shared__ float win[kChunk + 2 * kMaxNt];
    __shared__ float win2[kChunk + 2 * kMaxNt];
    __shared__ int s_wc[kChunk / 64];
    const int t0 = blockIdx.x * kChunk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x = t0 + tid;
    const int my_id = (tid < kChunk && x < T) ? imax[x] : 0;     // issued with the window
    const int WL = kChunk + 2 * nt;
    for (int e = tid; e < WL; e += NTH) {
        const int xe = t0 - nt + e;
        win[e] = (xe >= 0 && xe < T) ? cfmax[xe] : -INFINITY;
    }
    for (int e = tid; e < K; e += NTH) s_rank[e] = rank[e];
    __syncthreads();
    if (done) return;                        // block-uniform
    const float v = tid < kChunk ? win[tid + nt] : 0.f;     // Cfmax of column x
    // window max of width L = 2nt+1 (max_pool1d, -inf padding): maxima of
    // 2^k consecutive values by doubling, then two overlapping 2^k windows
    const int L = 2 * nt + 1;
    const int k = 31 - __clz(L);
    float *cur = win, *nxt = win2;
    for (int j = 0; j < k; ++j) {
        const int half = 1 << j;
        for (int e = tid; e < WL; e += NTH)
            nxt[e] = e + half < WL ? nan_max(cur[e], cur[e + half]) : cur[e];
        __syncthreads();
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    bool acc = false;
    float cmax = 0.f;
    if (tid < kChunk && x >= nt && x < T - nt) {
        // the reference's test: (Cmax > Th^2) & (|Cmax - Cfmax| < 1e-9)
        cmax = nan_max(cur[tid], cur[tid + L - (1 << k)]);
        acc = (cmax > th2) && (fabsf(cmax - v) < 1e-9f);
    }
    const unsigned long long m = __ballot(acc);
    if (lane == 0 && wave < kChunk / 64) s_wc[wave] = __popcll(m);
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
            const int iy = my_id;
            const float b = Bt[(size_t)x * ldbt + s_rank[iy]];
            const GramRef g = gref[iy];
            Slot sl;
            sl.x = x;
            sl.iy = iy;
            sl.amp = b / g.nm;
            sl.tha = sqrtf(cmax);
            sl.lohi = (unsigned)g.lo | ((unsigned)g.hi << 16);
            sl.crg = (unsigned)g.clo | ((unsigned)g.chi << 16);
            sl.off = g.off;
            slots[(size_t)blockIdx.x * kSlotCap + pos] = sl;
        }
    }
    if (tid == 0) {
        if (count > kSlotCap) {
            atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
            atomicExch(&state[ST_DONE], 1);
        }
        chunk_count[blockIdx.x] = count < kSlotCap ? count : kSlotCap;
    }
}

// ---------------------------------------------------------------------------
// Certified selection for reduced-precision scores (BF16, split BF16, FP8).
//
// The scores B~ of a reduced-precision projection come with a bound
// (ksm_project_bounded): per unit template norm, |B~[r,t] - B[r,t]| /
// sqrt(nm_r) <= D_r(t) + colA[t] <= colD[t], so sqrt(Cf~) = relu(B~)/sqrt(nm)
// lies within that of sqrt(Cf) for every template (relu is 1-Lipschitz), and
// a column's best score within colD[t] of its approximate best.  The
// reference's decision for column x (template_matching.py:197-212: Cmax >
// Th^2 and Cmax == Cfmax, Cmax the +-nt window maximum) is taken on these
// intervals where they settle it and on exact values where they do not: a
// column whose decision is open is re-scored in FP32 from the residual, for
// every template that can still be its best,
//   b = sum_c sum_k Xres[c, x + k - nt/2] * T_r[c, k]
// (T_r the template's waveforms on its significant channels), which makes
// the column exact (its best score and template known to FP32 accuracy).
// Rounds repeat until every column of the chunk is decided:
//   * certainly rejected: its upper bound is <= Th^2, or some column of its
//     window certainly scores more (a lower bound above its upper bound);
//   * open and not exact: re-scored;
//   * exact, with non-exact window columns that may reach its score:
//     those are re-scored;
//   * exact, every other window column exact or certainly below: the
//     reference's test on the exact values.
// Each round makes at least one window column exact, so the rounds end.
// Re-scored values are written back into Bt (the scores the update kernel
// subtracts from) for the chunk's own columns, and a column's exactness is
// kept in the sign bit of its cfmax entry (cfmax = -Cf): the update kernel
// writes every column it changes with a non-negative value, which clears
// it.  Halo columns re-scored by a neighbouring chunk stay in that block's
// LDS.  Accepted spikes take their FP32 b for amps (b/nm) and th_amps.
constexpr int kCertThreads = 512;
constexpr int kCertMaxWin = kChunk + 2 * kMaxNt;
constexpr float kSqrtSlack = 0x1p-20f;   // sqrt and score rounding, relative

__device__ __forceinline__ bool cf_is_exact(float v) {
    return (__float_as_uint(v) >> 31) != 0u || v != v;
}

// The bound D_r(y) of template position r at column y (without colA),
// rounded up, from the column's sub-block statistics sc[b] = S[b][y] in LDS.
__device__ __forceinline__ float bound_lds(const ksm::BoundView &bv, const float *sc, int K,
                                           int r) {
    float a = 0.f;
    // 16 weights per step, loads issued together
    for (int b0 = 0; b0 < bv.nsb; b0 += 16) {
        float w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) w[u] = b0 + u < bv.nsb ? bv.Wt[(size_t)(b0 + u) * K + r] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (b0 + u < bv.nsb) a = __builtin_fmaf(w[u], sc[b0 + u], a);
    }
    a *= 1.f + 0x1p-16f;
    if (bv.mode == ksm::kBoundSqrt) a = sqrtf(a) * (1.f + kSqrtSlack);
    return a;
}

// Partial FP32 score of template position r at column y over its significant
// channels [c0, c1) (at most kTaskCh): lanes take lags, each lane sums its
// channels in order, then a fixed butterfly over the lanes (wave-uniform).
// A column's score b is the sum of its chunks' partials in chunk order, the
// same whichever wave computes which chunk.
constexpr int kTaskCh = 16;
__device__ float rescore_chunk(int r, int y, int c0, int c1, const float *__restrict__ Xres,
                               int ldx, const ksm::BoundView &bv, int C, int nt) {
    const int lane = threadIdx.x & 63;
    const int h = nt / 2;
    const int lo = uni(bv.srng[2 * r]);
    const float *tr = bv.T + (size_t)r * C * nt;
    float part = 0.f;
    for (int kb0 = 0; kb0 < nt; kb0 += 64) {
        const int k = kb0 + lane;
        const bool kin = k < nt;
        const int kk = kin ? k : 0;
        const float *xb = Xres + (size_t)(y - h + kk);
        const float *tb = tr + kk;
        float xv[kTaskCh], tv[kTaskCh];
#pragma unroll
        for (int u = 0; u < kTaskCh; ++u) {
            const int c = c0 + u < c1 ? c0 + u : c1 - 1;
            xv[u] = xb[(size_t)c * ldx];
            tv[u] = tb[(size_t)(c - lo) * nt];
        }
#pragma unroll
        for (int u = 0; u < kTaskCh; ++u)
            part = __builtin_fmaf((kin && c0 + u < c1) ? xv[u] : 0.f, tv[u], part);
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    return uni(part);
}

// Partials of two consecutive chunks [c0, c1) and [c1, c2) (each at most
// kTaskCh channels) with their loads issued together; each equals
// rescore_chunk of its chunk.
__device__ void rescore_chunk2(int r, int y, int c0, int c1, int c2,
                               const float *__restrict__ Xres, int ldx, const ksm::BoundView &bv,
                               int C, int nt, float &p0, float &p1) {
    const int lane = threadIdx.x & 63;
    const int h = nt / 2;
    const int lo = uni(bv.srng[2 * r]);
    const float *tr = bv.T + (size_t)r * C * nt;
    float a0 = 0.f, a1 = 0.f;
    for (int kb0 = 0; kb0 < nt; kb0 += 64) {
        const int k = kb0 + lane;
        const bool kin = k < nt;
        const int kk = kin ? k : 0;
        const float *xb = Xres + (size_t)(y - h + kk);
        const float *tb = tr + kk;
        float xv[2 * kTaskCh], tv[2 * kTaskCh];
#pragma unroll
        for (int u = 0; u < 2 * kTaskCh; ++u) {
            const int cs = u < kTaskCh ? c0 + u : c1 + (u - kTaskCh);
            const int ce = u < kTaskCh ? c1 : c2;
            const int c = cs < ce ? cs : ce - 1;
            xv[u] = xb[(size_t)c * ldx];
            tv[u] = tb[(size_t)(c - lo) * nt];
        }
#pragma unroll
        for (int u = 0; u < kTaskCh; ++u)
            a0 = __builtin_fmaf((kin && c0 + u < c1) ? xv[u] : 0.f, tv[u], a0);
#pragma unroll
        for (int u = 0; u < kTaskCh; ++u)
            a1 = __builtin_fmaf((kin && c1 + u < c2) ? xv[kTaskCh + u] : 0.f, tv[kTaskCh + u], a1);
    }
    for (int o = 32; o > 0; o >>= 1) a0 += __shfl_xor(a0, o, 64);
    for (int o = 32; o > 0; o >>= 1) a1 += __shfl_xor(a1, o, 64);
    p0 = uni(a0);
    p1 = uni(a1);
}

// The full FP32 score of position r at column y: its chunks' partials in
// order (two chunks' loads at a time).
__device__ float rescore_full(int r, int y, const float *__restrict__ Xres, int ldx,
                              const ksm::BoundView &bv, int C, int nt) {
    const int lo = uni(bv.srng[2 * r]), hi = uni(bv.srng[2 * r + 1]);
    float b = 0.f;
    int c0 = lo;
    for (; c0 + kTaskCh < hi; c0 += 2 * kTaskCh) {
        float p0, p1;
        rescore_chunk2(r, y, c0, c0 + kTaskCh, min(hi, c0 + 2 * kTaskCh), Xres, ldx, bv, C, nt, p0,
                       p1);
        b += p0;
        b += p1;
    }
    if (c0 < hi) b += rescore_chunk(r, y, c0, hi, Xres, ldx, bv, C, nt);
    return b;
}

constexpr int kMaxPairs = 256;           // (column, template) re-scores per round, shared
constexpr int kMaxTasks = 1024;          // their channel chunks
constexpr int kWavePl = 128;             // plausible positions a wave collects per column
constexpr int kDirectPl = 2;             // at most this many: the scanning wave scores them

__global__ __launch_bounds__(kCertThreads) void peel_select_cert_kernel(
    float *__restrict__ cfmax, int *__restrict__ imax, ksm::BoundView bv,
    const int *__restrict__ bhdr, const u64 *__restrict__ gkey, int G, float *__restrict__ Bt,
    int ldbt, const GramRef *__restrict__ gref, const float *__restrict__ pnm,
    const int *__restrict__ ppid, int T, int nt, float th2, const float *__restrict__ Xres,
    int ldx, int C, int K, int *__restrict__ state, int *__restrict__ chunk_count,
    Slot *__restrict__ slots) {
    const int done = state[ST_DONE];
    bv.mode = bhdr[0];
    __shared__ float s_v[kCertMaxWin];       // Cf of the column's best (approximate or exact)
    __shared__ float s_d[kCertMaxWin];       // its bound in sqrt(Cf) units; 0: exact
    __shared__ int s_id[kCertMaxWin];        // its template
    __shared__ float s_b[kCertMaxWin];       // FP32 score re-scored here (NaN: not re-scored)
    __shared__ float s_lo0[kCertMaxWin];     // lower bounds
    __shared__ float s_lo[2][kCertMaxWin];   // their window maxima by doubling
    __shared__ float s_hn[kCertMaxWin];      // upper bounds of non-exact columns
    __shared__ float s_ex[kCertMaxWin];      // exact values
    __shared__ unsigned char s_req[kCertMaxWin];
    __shared__ unsigned char s_dec[kChunk];  // 0 open, 1 rejected, 2 accepted
    __shared__ int s_list[kCertMaxWin];      // requested columns (window index)
    __shared__ int s_cbase[kCertMaxWin], s_ccnt[kCertMaxWin];   // their pairs (-1: done)
    __shared__ int s_pe[kMaxPairs], s_pr[kMaxPairs], s_pt[kMaxPairs + 1];
    __shared__ float s_part[kMaxTasks];
    __shared__ float s_scol[kCertThreads / 64][ksm::kBoundMaxC / ksm::kBoundSub];
    __shared__ int s_pl[kCertThreads / 64][kWavePl];
    __shared__ int s_n, s_np, s_nu, s_wc[kChunk / 64];
    __shared__ int s_ulist[kChunk];          // open columns a wave decides
    const int t0 = blockIdx.x * kChunk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = kCertThreads / 64;
    const int WL = kChunk + 2 * nt;
    for (int e = tid; e < WL; e += kCertThreads) {
        const int xe = t0 - nt + e;
        float v = -INFINITY, d = 0.f;
        int id = 0;
        if (xe >= nt && xe < T - nt) {
            const float c = cfmax[xe], dc = bv.colD[xe];     // loads issued together
            id = imax[xe];
            if (cf_is_exact(c)) {
                v = c != c ? c : -c;
            } else {
                v = c;
                d = dc;
            }
        } else if (xe >= 0 && xe < T) {
            v = 0.f;                          // the zeroed edge columns
        }
        s_v[e] = v;
        s_d[e] = d;
        s_id[e] = id;
        s_b[e] = __builtin_nanf("");
        s_req[e] = 0;
    }
    if (tid < kChunk) s_dec[tid] = 0;
    __syncthreads();
    if (done) return;                        // block-uniform
    const float sth = sqrtf(th2) * (1.f - kSqrtSlack);
    const int L = 2 * nt + 1;
    const int kdbl = 31 - __clz(L);
    int rescored = 0;
    for (int round = 0; round <= 2 * WL; ++round) {
        // interval ends (sqrt units) of every window column: its lower
        // bound, the upper bound of a non-exact column (-inf: exact) and the
        // exact value (-inf: not exact)
        for (int e = tid; e < WL; e += kCertThreads) {
            const float v = s_v[e], d = s_d[e];
            const float sv = v == -INFINITY ? -INFINITY : sqrtf(v);
            const bool ex = d == 0.f;
            const float lo = ex ? sv * (1.f - kSqrtSlack) : fmaxf(0.f, sv * (1.f - kSqrtSlack) - d);
            s_lo0[e] = lo;
            s_lo[0][e] = lo;
            s_hn[e] = ex ? -INFINITY : sv * (1.f + kSqrtSlack) + d;
            s_ex[e] = ex ? v : -INFINITY;
        }
        if (tid == 0) {
            s_n = 0;
            s_np = 0;
            s_nu = 0;
        }
        __syncthreads();
        // first round: window maxima of the lower bounds of every column by
        // doubling, which settles most columns
        int cur = 0;
        if (round == 0) {
            for (int j = 0; j < kdbl; ++j) {
                const int half = 1 << j;
                for (int e = tid; e < WL; e += kCertThreads) {
                    const bool in = e + half < WL;
                    s_lo[cur ^ 1][e] = in ? nan_max(s_lo[cur][e], s_lo[cur][e + half])
                                          : s_lo[cur][e];
                }
                __syncthreads();
                cur ^= 1;
            }
        }
        const int x = t0 + tid;
        if (tid < kChunk && s_dec[tid] == 0) {
            const int e = tid + nt;
            if (x < nt || x >= T - nt) {
                s_dec[tid] = 1;
            } else if (round == 0 && s_d[e] != 0.f) {
                const int o = L - (1 << kdbl);
                const float wlo = nan_max(s_lo[cur][tid], s_lo[cur][tid + o]);
                const float sv = sqrtf(s_v[e]);
                const float hi = sv * (1.f + kSqrtSlack) + s_d[e];
                if (!(hi >= sth) || wlo > hi) s_dec[tid] = 1;   // certainly not accepted
                else s_req[e] = 1;                              // open: re-score it
            } else {
                s_ulist[atomicAdd(&s_nu, 1)] = tid;             // decided by a wave below
            }
        }
        __syncthreads();
        // the other open columns, one wave each: the window's maxima by a
        // lane-parallel read and a butterfly
        const int nu = s_nu;
        for (int jj = wave; jj < nu; jj += NW) {
            const int tx = uni(s_ulist[jj]);
            const int e = tx + nt;
            const float v = s_v[e], d = s_d[e];
            const float sv = sqrtf(v);
            const float hi = sv * (1.f + kSqrtSlack) + d;
            float wlo = -INFINITY, whn = -INFINITY, wex = -INFINITY;
            for (int q = tx + lane; q < tx + L; q += 64) {
                wlo = nan_max(wlo, s_lo0[q]);
                whn = fmaxf(whn, s_hn[q]);
                wex = nan_max(wex, s_ex[q]);
            }
            for (int o = 32; o > 0; o >>= 1) {
                wlo = nan_max(wlo, __shfl_xor(wlo, o, 64));
                whn = fmaxf(whn, __shfl_xor(whn, o, 64));
                wex = nan_max(wex, __shfl_xor(wex, o, 64));
            }
            if (!(hi >= sth) || wlo > hi) {
                if (lane == 0) s_dec[tx] = 1;                   // certainly not accepted
            } else if (d != 0.f) {
                if (lane == 0) s_req[e] = 1;                    // open: re-score it
            } else {
                // exact: re-score the window's non-exact columns that may
                // reach it; with none, the reference's test on the exact
                // values (the others are certainly below)
                const float lim = sv * (1.f - kSqrtSlack);
                if (whn >= lim) {
                    for (int q = tx + lane; q < tx + L; q += 64)
                        if (s_hn[q] >= lim) s_req[q] = 1;
                } else if (lane == 0) {
                    const bool acc = (wex > th2) && (fabsf(wex - v) < 1e-9f);
                    s_dec[tx] = acc ? 2 : 1;
                }
            }
        }
        __syncthreads();
        for (int e = tid; e < WL; e += kCertThreads)
            if (s_req[e]) {
                s_req[e] = 0;
                s_list[atomicAdd(&s_n, 1)] = e;
            }
        __syncthreads();
        const int nreq = s_n;
        if (nreq == 0) break;                        // block-uniform: all decided
        // scan: one wave per requested column finds the templates that can
        // still be its best (groups of 32 positions by their best key first);
        // their (column, template) pairs go to the shared list, or, where it
        // is full, the wave re-scores them itself
        for (int j = wave; j < nreq; j += NW) {
            const int e = uni(s_list[j]);
            const int y = t0 - nt + e;
            // the column's loads issued together: sub-block statistics, its
            // bounds, the group keys (the first 64 groups)
            float *sc = s_scol[wave];
            const u64 *gk = gkey + (size_t)y * G;
            const float scv = lane < bv.nsb ? bv.S[(size_t)lane * T + y] : 0.f;
            const u64 gk0 = lane < G ? gk[lane] : 0ull;
            const float ay = bv.colA[y], dcol = bv.colD[y];
            if (lane < bv.nsb) sc[lane] = scv;
            for (int b = 64 + lane; b < bv.nsb; b += 64) sc[b] = bv.S[(size_t)b * T + y];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the true best is at least the column's approximate best less colD
            const float lo_col = fmaxf(0.f, sqrtf(s_v[e]) * (1.f - kSqrtSlack) - dcol);
            // a group's key is within colD of the current scores of its
            // positions (its values are the approximate ones of the last
            // reduction, and re-scored values lie within the bound of them)
            int *pl = s_pl[wave];
            int npl = 0;
            bool over = false;
            for (int g0 = 0; g0 < G; g0 += 64) {
                const int g = g0 + lane;
                bool gp = false;
                if (g < G) {
                    const float kf = key_score(g0 == 0 ? gk0 : gk[g]);
                    gp = kf != kf || sqrtf(kf) * (1.f + kSqrtSlack) + 2.f * dcol >= lo_col;
                }
                unsigned long long gm = __ballot(gp);
                while (gm) {
                    const int gg = g0 + __ffsll((long long)gm) - 1;
                    gm &= gm - 1;
                    // lanes 0-31 take the group's positions; the template
                    // weights are loaded with the scores
                    const int r = gg * kGroup + (lane & 31);
                    bool p = false;
                    if (lane < 32 && r < K) {
                        const float bval = Bt[(size_t)y * ldbt + r], nmr = pnm[r];
                        const float dr = bound_lds(bv, sc, K, r);
                        const float cf = score(bval, nmr);
                        const float up = sqrtf(cf) * (1.f + kSqrtSlack);
                        p = cf != cf || up + dr + ay >= lo_col;
                    }
                    const unsigned long long m = __ballot(p);
                    const int cnt = __popcll(m);
                    if (npl + cnt > kWavePl) {
                        over = true;
                    } else if (p) {
                        pl[npl + __popcll(m & lanemask_lt(lane))] = r;
                    }
                    npl = over ? npl : npl + cnt;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            int base = -1;
            if (!over && npl <= kDirectPl) {
                // few templates: the scanning wave scores them itself
                const bool own = e >= nt && e < nt + kChunk;
                u64 best = 0ull;
                float bbest = 0.f;
                for (int q = 0; q < npl; ++q) {
                    const int rr = uni(pl[q]);
                    const float b = rescore_full(rr, y, Xres, ldx, bv, C, nt);
                    const u64 key = make_key(score(b, pnm[rr]), uni(ppid[rr]));
                    if (key > best) {
                        best = key;
                        bbest = b;
                    }
                    if (own && lane == 0) Bt[(size_t)y * ldbt + rr] = b;
                }
                rescored += npl;
                if (best == 0ull) {
                    best = make_key(s_v[e], s_id[e]);
                    bbest = __builtin_nanf("");
                }
                if (lane == 0) {
                    const float cf = key_score(best);
                    s_v[e] = cf;
                    s_id[e] = key_id(best);
                    s_b[e] = bbest;
                    s_d[e] = 0.f;
                    s_ccnt[j] = -1;
                    if (own) {
                        cfmax[y] = cf != cf ? cf : -cf;  // sign bit: exact
                        imax[y] = key_id(best);
                    }
                }
                continue;
            }
            if (!over && lane == 0) {
                // reserve npl entries only where they fit, so that the list
                // [0, s_np) holds no unwritten entries
                int old = s_np;
                for (;;) {
                    if (old + npl > kMaxPairs) break;
                    const int prev = atomicCAS(&s_np, old, old + npl);
                    if (prev == old) {
                        base = old;
                        break;
                    }
                    old = prev;
                }
            }
            base = uni(__shfl(base, 0, 64));
            const bool own = e >= nt && e < nt + kChunk;
            if (base >= 0) {
                for (int q = lane; q < npl; q += 64) {
                    s_pe[base + q] = e;
                    s_pr[base + q] = pl[q];
                }
                if (lane == 0) {
                    s_cbase[j] = base;
                    s_ccnt[j] = npl;
                }
                continue;
            }
            // the wave re-scores the column itself: every position of every
            // group that can hold its best (the full scan when the collected
            // list overflowed)
            u64 best = 0ull;
            float bbest = 0.f;
            for (int g = 0; g < G; ++g) {
                const float kf = key_score(gk[g]);
                if (!(kf != kf || sqrtf(kf) * (1.f + kSqrtSlack) + 2.f * dcol >= lo_col)) continue;
                const int r = g * kGroup + (lane & 31);
                bool p = false;
                if (lane < 32 && r < K) {
                    const float cf = score(Bt[(size_t)y * ldbt + r], pnm[r]);
                    const float up = sqrtf(cf) * (1.f + kSqrtSlack);
                    p = cf != cf || (up + dcol >= lo_col &&
                                     up + bound_lds(bv, sc, K, r) + ay >= lo_col);
                }
                unsigned long long m = __ballot(p);
                while (m) {
                    const int rr = g * kGroup + __ffsll((long long)m) - 1;
                    m &= m - 1;
                    const float b = rescore_full(rr, y, Xres, ldx, bv, C, nt);
                    const u64 key = make_key(score(b, pnm[rr]), uni(ppid[rr]));
                    if (key > best) {
                        best = key;
                        bbest = b;
                    }
                    ++rescored;
                    if (own && lane == 0) Bt[(size_t)y * ldbt + rr] = b;
                }
            }
            if (best == 0ull) {                // no template scored (cannot occur for
                best = make_key(s_v[e], s_id[e]);  // finite bounds): keep the column's own
                bbest = __builtin_nanf("");
            }
            if (lane == 0) {
                const float cf = key_score(best);
                s_v[e] = cf;
                s_id[e] = key_id(best);
                s_b[e] = bbest;
                s_d[e] = 0.f;
                s_ccnt[j] = -1;
                if (own) {
                    cfmax[y] = cf != cf ? cf : -cf;  // sign bit: exact
                    imax[y] = key_id(best);
                }
            }
        }
        __syncthreads();
        // channel chunks of the listed pairs, then one chunk per wave
        const int np = min(s_np, kMaxPairs);
        if (wave == 0) {
            int run = 0;
            for (int q0 = 0; q0 < np; q0 += 64) {
                const int q = q0 + lane;
                int nch = 0;
                if (q < np) {
                    const int r = s_pr[q];
                    nch = (bv.srng[2 * r + 1] - bv.srng[2 * r] + kTaskCh - 1) / kTaskCh;
                }
                int incl = nch;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += t;
                }
                if (q < np) s_pt[q] = run + incl - nch;
                run += __shfl(incl, 63, 64);
            }
            if (lane == 0) s_pt[np] = run;
        }
        __syncthreads();
        const int ntask = min(s_pt[np], kMaxTasks);
        for (int j = wave; j < ntask; j += NW) {
            int lo_q = 0, hi_q = np - 1;       // the pair q with s_pt[q] <= j < s_pt[q+1]
            while (lo_q < hi_q) {
                const int mid = (lo_q + hi_q + 1) >> 1;
                if (s_pt[mid] <= j) lo_q = mid;
                else hi_q = mid - 1;
            }
            const int q = uni(lo_q);
            const int r = uni(s_pr[q]), y = t0 - nt + uni(s_pe[q]);
            const int c0 = uni(bv.srng[2 * r]) + (j - s_pt[q]) * kTaskCh;
            const int c1 = min(uni(bv.srng[2 * r + 1]), c0 + kTaskCh);
            const float part = rescore_chunk(r, y, c0, c1, Xres, ldx, bv, C, nt);
            if (lane == 0) s_part[j] = part;
        }
        __syncthreads();
        // per listed column: its pairs' scores (chunks in order) and the best
        // of them; pairs whose chunks did not fit are scored here in full
        for (int j = wave; j < nreq; j += NW) {
            const int cnt = s_ccnt[j];
            if (cnt < 0) continue;             // done by its scanning wave
            const int e = uni(s_list[j]);
            const int y = t0 - nt + e;
            const int base = uni(s_cbase[j]);
            const bool own = e >= nt && e < nt + kChunk;
            u64 best = 0ull;
            float bbest = 0.f;
            for (int q = base; q < base + cnt; ++q) {
                const int r = uni(s_pr[q]);
                float b = 0.f;
                if (s_pt[q + 1] <= kMaxTasks) {
                    for (int t = s_pt[q]; t < s_pt[q + 1]; ++t) b += s_part[t];
                } else {
                    b = rescore_full(r, y, Xres, ldx, bv, C, nt);
                }
                b = uni(b);
                const u64 key = make_key(score(b, pnm[r]), uni(ppid[r]));
                if (key > best) {
                    best = key;
                    bbest = b;
                }
                if (own && lane == 0) Bt[(size_t)y * ldbt + r] = b;
            }
            rescored += cnt;
            if (best == 0ull) {                // no template scored (cannot occur for
                best = make_key(s_v[e], s_id[e]);  // finite bounds): keep the column's own
                bbest = __builtin_nanf("");
            }
            if (lane == 0) {
                const float cf = key_score(best);
                s_v[e] = cf;
                s_id[e] = key_id(best);
                s_b[e] = bbest;
                s_d[e] = 0.f;
                if (own) {
                    cfmax[y] = cf != cf ? cf : -cf;  // sign bit: exact
                    imax[y] = key_id(best);
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0 && rescored) atomicAdd(&state[ST_RESCORED], rescored);
    // ordered compaction of the accepted columns (time order)
    const bool acc = tid < kChunk && s_dec[tid] == 2;
    const unsigned long long m = __ballot(acc);
    if (lane == 0 && wave < kChunk / 64) s_wc[wave] = __popcll(m);
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
            const int x = t0 + tid, e = tid + nt;
            const int iy = s_id[e];
            const GramRef g = gref[iy];
            float b = s_b[e];
            if (b != b) b = Bt[(size_t)x * ldbt + g.rank];   // exact since an earlier launch
            Slot sl;
            sl.x = x;
            sl.iy = iy;
            sl.amp = b / g.nm;
            sl.tha = sqrtf(s_v[e]);
            sl.lohi = (unsigned)g.lo | ((unsigned)g.hi << 16);
            sl.crg = (unsigned)g.clo | ((unsigned)g.chi << 16);
            sl.off = g.off;
            slots[(size_t)blockIdx.x * kSlotCap + pos] = sl;
        }
    }
    if (tid == 0) {
        if (count > kSlotCap) {
            atomicExch(&state[ST_STATUS], KSM_ERR_TIES);
            atomicExch(&state[ST_DONE], 1);
        }
        chunk_count[blockIdx.x] = count < kSlotCap ? count : kSlotCap;
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

// Lane l < n of the wave loads spike q0 + l (lanes outside [0, S): x = -1),
// so that a spike's neighbours arrive with one load.
__device__ __forceinline__ Slot spike_lanes(int q0, int n, int S, const int *pre, int nchunks,
                                            const Slot *__restrict__ slots) {
    const int lane = threadIdx.x & 63;
    const int q = q0 + lane;
    Slot sl = {-1, 0, 0.f, 0.f, 0u, 0u, 0ll};
    if (lane < n && q >= 0 && q < S) sl = spike_at(q, pre, nchunks, slots);
    return sl;
}

__device__ __forceinline__ int rl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ unsigned rl(unsigned v, int l) {
    return (unsigned)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float rl(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ long long rl(long long v, int l) {
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)v, l);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)v >> 32), l);
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// The spike held by lane l.
__device__ __forceinline__ SpikeG spike_gram(const Slot &w, int l) {
    SpikeG s;
    s.x = rl(w.x, l);
    s.amp = rl(w.amp, l);
    const unsigned lh = rl(w.lohi, l);
    s.lo = (int)(lh & 0xffffu);
    s.hi = (int)(lh >> 16);
    s.off = rl(w.off, l);
    return s;
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

// Subtractions of a column.  The reference subtracts each phase
// (even-ranked spikes, then odd-ranked) with one indexed write, so where
// windows of one phase overlap -- exact score ties place accepted spikes
// within nt columns of each other, and any two within 2nt+1 share columns of
// their +-nt windows -- the last spike's write is the one that remains
// (torch's non-accumulating index_put on the CPU).  The spikes covering a
// column are consecutive in time order, so of alternating rank parity: the
// column takes exactly two subtractions, the last even-ranked and the last
// odd-ranked of them -- the last covering spike and the one before it -- even
// first, however many spikes cover it.  A column is therefore owned by the
// last spike whose window covers it.

// Spike q with its neighbours: q owns the columns of its window that the
// window of spike q+1 does not reach, and spike q-1 covers those it reaches.
struct SpikeCtx {
    SpikeG g0;              // spike q
    SpikeG gp;              // spike q-1 (has_p)
    bool has_p, has_n;
    int x_n;                // column of spike q+1 (has_n)
};

__device__ __forceinline__ void spike_ctx(SpikeCtx &c, int q, int S, const int *pre, int nchunks,
                                          const Slot *__restrict__ slots) {
    // lanes 0, 1, 2: spikes q - 1, q, q + 1
    const Slot w = spike_lanes(q - 1, 3, S, pre, nchunks, slots);
    c.g0 = spike_gram(w, 1);
    c.has_p = q > 0;
    c.gp = spike_gram(w, 0);
    c.has_n = q + 1 < S;
    c.x_n = rl(w.x, 2);
}

// Per-lane constants of one pass (kPassGroups groups) over the columns of a
// segment: lanes 8 g .. 8 g + 7 take group g of the pass, 4 positions each.
struct SegPass {
    int r, rb_, ra, rb;     // first position; B offset (clamped); Gram offsets (clamped)
    bool in, ha, hb;        // lane has a group; its positions take a's / b's subtraction
};

__device__ __forceinline__ void seg_pass(SegPass &sp, const GroupList &L, int j0,
                                         const ColSpike &a, const ColSpike &b, int ldbt) {
    const int lane = threadIdx.x & 63;
    const int j = j0 + (lane >> 3);
    const int g = j < L.count() ? L.at(j) : -1;
    sp.in = g >= 0;
    sp.r = (g < 0 ? 0 : g) * kGroup + 4 * (lane & 7);
    sp.rb_ = min(sp.r, ldbt - 4);
    sp.ra = a.hi > a.lo ? min(max(sp.r, a.lo), a.hi - 4) : 0;
    sp.rb = b.hi > b.lo ? min(max(sp.r, b.lo), b.hi - 4) : 0;
    sp.ha = sp.in && sp.r >= a.lo && sp.r < a.hi;
    sp.hb = sp.in && sp.r >= b.lo && sp.r < b.hi;
}

struct ColLoads {
    float4 v0, va0, vb0, v1, va1, vb1;
    u64 kept;
};

// 16- and 8-byte accesses at base + i elements.
__device__ __forceinline__ float4 ld4(const float *base, int i) {
    return *reinterpret_cast<const float4 *>(base + i);
}
__device__ __forceinline__ void st4(float *base, int i, const float4 &v) {
    *reinterpret_cast<float4 *>(base + i) = v;
}
__device__ __forceinline__ u64 ld8(const u64 *base, int i) { return base[i]; }
__device__ __forceinline__ void st8(u64 *base, int i, u64 v) { base[i] = v; }

// Key of group `lane` if it is outside L (read before any store); the
// segment paths run only where G <= 64.  keep: lane < G and lane not in L.
__device__ __forceinline__ u64 seg_kept(bool keep, const GroupList &L, int G,
                                        const u64 *__restrict__ gk) {
    return keep ? ld8(gk, threadIdx.x & 63) : 0ull;
}

__device__ __forceinline__ void seg_load(ColLoads &l, int t, long long ba, long long bb,
                                         bool two, const SegPass &s0, const SegPass &s1,
                                         bool keep, const GroupList &L,
                                         const float *__restrict__ Bt, int ldbt,
                                         const float *__restrict__ cc, int G,
                                         const u64 *__restrict__ gkey) {
    const float *col = Bt + (size_t)t * ldbt;
    l.v0 = ld4(col, s0.rb_);
    l.va0 = ld4(cc + ba, s0.ra);
    l.vb0 = ld4(cc + bb, s0.rb);
    if (two) {
        l.v1 = ld4(col, s1.rb_);
        l.va1 = ld4(cc + ba, s1.ra);
        l.vb1 = ld4(cc + bb, s1.rb);
    }
    l.kept = seg_kept(keep, L, G, gkey + (size_t)t * G);
}

// Applies a then b to the lane's 4 positions (products rounded before the
// subtract) in place and returns the lane group's key.
__device__ __forceinline__ u64 seg_compute(float4 &v, const float4 &va, const float4 &vb,
                                           const SegPass &sp, float amp_a, float amp_b, int K,
                                           const float *s_nm, const int *s_pid) {
    const float4 n4 = *reinterpret_cast<const float4 *>(s_nm + sp.r);
    const int4 p4 = *reinterpret_cast<const int4 *>(s_pid + sp.r);
    float bv[4] = {v.x, v.y, v.z, v.w};
    const float ca[4] = {va.x, va.y, va.z, va.w};
    const float cb[4] = {vb.x, vb.y, vb.z, vb.w};
    const float nv[4] = {n4.x, n4.y, n4.z, n4.w};
    const int pv[4] = {p4.x, p4.y, p4.z, p4.w};
    u64 kk = 0ull;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float pa = amp_a * ca[e];
        const float pb = amp_b * cb[e];
        float x = bv[e];
        x = sp.ha ? x - pa : x;
        x = sp.hb ? x - pb : x;
        bv[e] = x;
        const u64 k = sp.r + e < K ? make_key(score(x, nv[e]), pv[e]) : 0ull;
        kk = k > kk ? k : kk;
    }
    v = make_float4(bv[0], bv[1], bv[2], bv[3]);
    return group_max_u64(sp.in ? kk : 0ull);
}

__device__ __forceinline__ void seg_store(const float4 &v, const SegPass &sp, u64 key, int t,
                                          float *__restrict__ Bt, int ldbt, int G,
                                          u64 *__restrict__ gkey, u64 &acc) {
    if (sp.ha || sp.hb) st4(Bt + (size_t)t * ldbt, sp.r, v);
    if (sp.in && (threadIdx.x & 7) == 0) st8(gkey + (size_t)t * G, sp.r / kGroup, key);
    acc = key > acc ? key : acc;
}

// Column t of a segment, whose loads are in `cur`: compute, issue the next
// column's loads into `nxt`, store, and write the column's best key.
__device__ __forceinline__ void seg_column(ColLoads &cur, ColLoads &nxt, int t, int t1,
                                           long long &ba, long long &bb, int wa, int wb,
                                           bool two, const SegPass &s0, const SegPass &s1,
                                           bool keep, const GroupList &L, float amp_a,
                                           float amp_b, float *__restrict__ Bt, int ldbt, int K,
                                           int G,
                                           const float *__restrict__ cc, const float *s_nm,
                                           const int *s_pid, u64 *__restrict__ gkey,
                                           float *__restrict__ cfmax, int *__restrict__ imax) {
    const u64 k0 = seg_compute(cur.v0, cur.va0, cur.vb0, s0, amp_a, amp_b, K, s_nm, s_pid);
    u64 k1 = 0ull;
    if (two) k1 = seg_compute(cur.v1, cur.va1, cur.vb1, s1, amp_a, amp_b, K, s_nm, s_pid);
    ba += wa;
    bb += wb;
    if (t < t1) seg_load(nxt, t + 1, ba, bb, two, s0, s1, keep, L, Bt, ldbt, cc, G, gkey);
    u64 acc = cur.kept;
    seg_store(cur.v0, s0, k0, t, Bt, ldbt, G, gkey, acc);
    if (two) seg_store(cur.v1, s1, k1, t, Bt, ldbt, G, gkey, acc);
    const u64 best = wave_max_u64(acc);
    if ((threadIdx.x & 63) == 0) {
        cfmax[t] = key_score(best);
        imax[t] = key_id(best);
    }
}

// Columns t0 .. t1 of one segment: fixed covering spikes a (first) and b
// (either may be absent: empty range), whose Gram rows for column t start at
// a.base + (t - t0) wa and b.base + (t - t0) wb.  At most 2 passes.  The
// columns run as a pipeline: the next column's loads are issued before the
// current column's stores (a store issued ahead of a load would hold that
// load's wait on gfx9, whose counter covers both).
__device__ __forceinline__ void seg_columns(int t0, int t1, const ColSpike &a, const ColSpike &b,
                                            int wa, int wb, const GroupList &L,
                                            float *__restrict__ Bt, int ldbt, int K, int G,
                                            const float *__restrict__ cc, const float *s_nm,
                                            const int *s_pid, u64 *__restrict__ gkey,
                                            float *__restrict__ cfmax, int *__restrict__ imax) {
    const int lane = threadIdx.x & 63;
    const bool two = L.count() > kPassGroups;     // uniform
    SegPass s0, s1;
    seg_pass(s0, L, 0, a, b, ldbt);
    if (two) seg_pass(s1, L, kPassGroups, a, b, ldbt);
    else s1 = s0;
    const bool keep = lane < G && !L.has(lane);
    long long ba = a.base, bb = b.base;
    ColLoads cur, nxt;
    seg_load(cur, t0, ba, bb, two, s0, s1, keep, L, Bt, ldbt, cc, G, gkey);
    for (int t = t0;; ++t) {
        seg_column(cur, nxt, t, t1, ba, bb, wa, wb, two, s0, s1, keep, L, a.amp, b.amp, Bt, ldbt, K,
                   G, cc, s_nm, s_pid, gkey, cfmax, imax);
        if (t >= t1) break;
        cur = nxt;
    }
}

// Columns of one segment with 3 passes (two covering spikes with far-apart
// Gram ranges), as two jobs per column -- passes 0 and 1, then pass 2 -- in
// the same pipeline as above, so that a column costs two memory round trips.
__device__ __forceinline__ void seg_columns3(int t0, int t1, const ColSpike &a, const ColSpike &b,
                                             int wa, int wb, const GroupList &L,
                                             float *__restrict__ Bt, int ldbt, int K, int G,
                                             const float *__restrict__ cc, const float *s_nm,
                                             const int *s_pid, u64 *__restrict__ gkey,
                                             float *__restrict__ cfmax,
                                             int *__restrict__ imax) {
    const int lane = threadIdx.x & 63;
    const bool keep = lane < G && !L.has(lane);
    long long ba = a.base, bb = b.base;
    SegPass s0, s1;
    seg_pass(s0, L, 0, a, b, ldbt);
    seg_pass(s1, L, kPassGroups, a, b, ldbt);
    ColLoads cur, nxt;
    seg_load(cur, t0, ba, bb, true, s0, s1, keep, L, Bt, ldbt, cc, G, gkey);
    u64 acc = 0ull;
    int t = t0, h = 0;
    for (;;) {
        // job h of column t: passes (0, 1) or pass 2
        SegPass p, q;
        if (h == 0) {
            p = s0;
            q = s1;
        } else {
            seg_pass(p, L, 2 * kPassGroups, a, b, ldbt);
            q = p;
        }
        const u64 k0 = seg_compute(cur.v0, cur.va0, cur.vb0, p, a.amp, b.amp, K, s_nm, s_pid);
        u64 k1 = 0ull;
        if (h == 0) k1 = seg_compute(cur.v1, cur.va1, cur.vb1, q, a.amp, b.amp, K, s_nm, s_pid);
        const int tn = h ? t + 1 : t, hn = h ^ 1;
        if (h) {
            ba += wa;
            bb += wb;
        }
        const bool more = tn <= t1;
        if (more) {
            if (hn == 0) {
                seg_load(nxt, tn, ba, bb, true, s0, s1, keep, L, Bt, ldbt, cc, G, gkey);
            } else {
                SegPass r2;
                seg_pass(r2, L, 2 * kPassGroups, a, b, ldbt);
                seg_load(nxt, tn, ba, bb, false, r2, r2, false, L, Bt, ldbt, cc, G, gkey);
            }
        }
        if (h == 0) acc = cur.kept;
        seg_store(cur.v0, p, k0, t, Bt, ldbt, G, gkey, acc);
        if (h == 0) seg_store(cur.v1, q, k1, t, Bt, ldbt, G, gkey, acc);
        if (h) {
            const u64 best = wave_max_u64(acc);
            if (lane == 0) {
                cfmax[t] = key_score(best);
                imax[t] = key_id(best);
            }
        }
        if (!more) break;
        cur = nxt;
        t = tn;
        h = hn;
    }
}

// Columns of one segment with more than 3 passes (K > 768 only), or of any
// segment where G > 64 (K > 2048), through the general pass pipeline.
struct SegColumnsGeneral {
    int t, t1, wa, wb;
    ColSpike a, b;
    __device__ __forceinline__ bool next(ColJob &J) {
        if (t > t1) return false;
        J.t = t;
        J.a = a;
        J.b = b;
        ++t;
        a.base += wa;
        b.base += wb;
        return true;
    }
};

// The column items [d_lo, d_hi] of spike q (time order): the columns it owns
// (live, not covered by spike q-1), cut into segments of fixed covering
// spikes.
__device__ __forceinline__ void spike_columns(int q, int d_lo, int d_hi, const SpikeCtx &c,
                                              int nt, int T, float *__restrict__ Bt, int ldbt,
                                              int K, int G, const float *__restrict__ cc,
                                              const float *s_nm, const int *s_pid,
                                              u64 *__restrict__ gkey, float *__restrict__ cfmax,
                                              int *__restrict__ imax) {
    int tl = c.g0.x - nt + d_lo, th = c.g0.x - nt + d_hi;
    tl = max(tl, nt);
    th = min(th, T - nt - 1);
    if (c.has_n) th = min(th, c.x_n - nt - 1);       // spike q+1 owns the later ones
    const int pe = c.has_p ? c.gp.x + nt : tl - 1;  // spike q-1 covers columns <= pe
    const bool q_even = (q & 1) == 0;
    // segment 0: covered by q-1 and q; segment 1: by q alone
    for (int m = 0; m < 2; ++m) {
        const int s0 = m == 0 ? tl : max(tl, pe + 1);
        const int s1 = m == 0 ? min(th, pe) : th;
        if (s0 > s1) continue;
        const ColSpike none = {0, 0, 0, 0.f};
        const ColSpike sq = col_spike(c.g0, s0, nt);
        ColSpike sp = none;
        int wp = 0;
        if (m == 0) {
            sp = col_spike(c.gp, s0, nt);
            wp = c.gp.hi - c.gp.lo;
        }
        const int wq = c.g0.hi - c.g0.lo;
        // the even-ranked spike is subtracted first
        const ColSpike a = q_even ? sq : sp, b = q_even ? sp : sq;
        const int wa = q_even ? wq : wp, wb = q_even ? wp : wq;
        const GroupList L = group_list(a, b, G, false);
        if (G <= 64 && L.count() <= 2 * kPassGroups) {
            seg_columns(s0, s1, a, b, wa, wb, L, Bt, ldbt, K, G, cc, s_nm, s_pid, gkey, cfmax,
                        imax);
        } else if (G <= 64 && L.count() <= 3 * kPassGroups) {
            seg_columns3(s0, s1, a, b, wa, wb, L, Bt, ldbt, K, G, cc, s_nm, s_pid, gkey, cfmax,
                         imax);
        } else {
            SegColumnsGeneral src{s0, s1, wa, wb, a, b};
            column_pipeline(src, false, Bt, ldbt, K, G, cc, s_nm, s_pid, gkey, cfmax, imax);
        }
    }
}

// Residual subtraction of spike q's item k where its residual window overlaps
// a neighbour's (spikes closer than nt samples: exact score ties).  As for B,
// a sample takes the subtractions of its last covering spike and of the one
// before it, even-ranked first; spike q takes the samples of its window that
// spike q+1's window does not reach, on its share of the union of the two
// templates' channel ranges.  Lanes = samples; the template coefficients are
// read per channel (wave-uniform).
__device__ void xres_pair(int q, int k, const Slot &rw, bool has_p, bool has_n,
                          float *__restrict__ Xres, int ldx, const float *__restrict__ U,
                          const float *__restrict__ W, const int *__restrict__ crange, int P,
                          int C, int nt) {
    const int lane = threadIdx.x & 63;
    const int h = nt / 2;
    const int x = rl(rw.x, 1), id = rl(rw.iy, 1);
    const float amp = rl(rw.amp, 1);
    const int xp = rl(rw.x, 0), idp = rl(rw.iy, 0);
    const float ampp = rl(rw.amp, 0);
    const int p_lo = x - h;
    const int p_hi = has_n ? min(x + h, rl(rw.x, 2) - h - 1) : x + h;
    const int pe = has_p ? xp + h : p_lo - 1;       // spike q-1 covers samples <= pe
    int clo = uni(crange[2 * id]), chi = uni(crange[2 * id + 1]);
    if (has_p && pe >= p_lo) {
        const int plo = uni(crange[2 * idp]), phi = uni(crange[2 * idp + 1]);
        if (phi > plo) {
            clo = chi > clo ? min(clo, plo) : plo;
            chi = max(chi, phi);
        }
    }
    if (chi <= clo || p_hi < p_lo) return;
    const int cw = (chi - clo + kXresItems - 1) / kXresItems;
    const int cl0 = clo + k * cw, cl1 = min(chi, cl0 + cw);
    const bool q_even = (q & 1) == 0;
    for (int s0 = p_lo; s0 <= p_hi; s0 += 64) {
        const int sm = s0 + lane;
        const bool in = sm <= p_hi;
        const bool cov_p = in && has_p && sm <= pe;
        const int kq = in ? sm - x + h : 0;
        const int kp = cov_p ? sm - xp + h : 0;
        for (int c = cl0; c < cl1; ++c) {
            if (!in) continue;
            float xv = Xres[(size_t)c * ldx + sm];
            float tq = 0.f, tp = 0.f;
            for (int p = 0; p < P; ++p) {
                tq = __builtin_fmaf(U[((size_t)id * P + p) * C + c], W[p * nt + kq], tq);
                if (cov_p)
                    tp = __builtin_fmaf(U[((size_t)idp * P + p) * C + c], W[p * nt + kp], tp);
            }
            const float prq = amp * tq;
            const float prp = ampp * tp;
            if (q_even) {
                xv = xv - prq;
                if (cov_p) xv = xv - prp;
            } else {
                if (cov_p) xv = xv - prp;
                xv = xv - prq;
            }
            Xres[(size_t)c * ldx + sm] = xv;
        }
    }
}

// One iteration's subtraction.  Work items per spike q (time order): 2nt+1
// columns of its window, which the waves take in contiguous ranges, and
// kXresItems residual items, dealt round-robin over all waves so that no
// wave gets more than about one.  A column is owned by the last spike whose
// window covers it; its owner applies its own subtraction and that of the
// spike before it where that one covers the column too, even-ranked first
// (the reference's two interleaved phases).  A wave runs its columns as a
// pipeline (the next column's loads are issued before the current column's
// stores), then its residual items: each subtracts amp * (U[id]^T W) from
// its share of the template's nonzero channels in Xres[:, x-nt/2 .. x+nt/2]
// (windows of one iteration's spikes are disjoint unless exact ties put
// spikes closer than nt: those go through xres_pair).  The block also
// records the spikes into st / amps / th_amps.
__global__ __launch_bounds__(kUpdThreads) void peel_update_kernel(
    int it, float *__restrict__ Bt, int ldbt, float *__restrict__ Xres, int ldx,
    const float *__restrict__ U, const float *__restrict__ W, const int *__restrict__ crange,
    int P, int C, const float *__restrict__ pnm, const int *__restrict__ ppid,
    const float *__restrict__ cc, int K, int G, int Kg, int T,
    int nt, int nchunks, const int *__restrict__ chunk_count, const Slot *__restrict__ slots,
    int *__restrict__ state, int64_t *__restrict__ st, float *__restrict__ amps,
    float *__restrict__ th_amps, int capacity, float *__restrict__ cfmax,
    int *__restrict__ imax, u64 *__restrict__ gkey, u64 *hflags, unsigned seq) {
    extern __shared__ float s_tab[];        // s_nm[Kg], s_pid[Kg], s_pre[nchunks + 1]
    float *s_nm = s_tab;
    int *s_pid = reinterpret_cast<int *>(s_tab + Kg);
    int *s_pre = s_pid + Kg;
    __shared__ int s_wsum[kUpdThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);

    // the loads of the prologue are issued together: the done flag, the
    // iteration's output base, the per-chunk spike counts and the per-position
    // tables
    const int done = state[ST_DONE];
    const int kb = state[ST_KBASE + it];
    const int per_t = (nchunks + kUpdThreads - 1) / kUpdThreads;
    const int c0 = tid * per_t, c1 = min(c0 + per_t, nchunks);
    int mine = 0;
    for (int c = c0; c < c1; ++c) mine += chunk_count[c];
    stage_tables(pnm, ppid, Kg, s_nm, s_pid);
    // exclusive prefix of the per-chunk spike counts (thread: consecutive chunks)
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    if (done) {                             // block-uniform
        if (hflags && blockIdx.x == 0 && tid == 0) publish(hflags + 1, seq);
        return;
    }
    if (hflags && blockIdx.x == 0 && tid == 0) publish(hflags, ((u64)seq << 32) | (unsigned)it);
    int base = 0;
    for (int w = 0; w < wv; ++w) base += s_wsum[w];
    int run = base + incl - mine;
    for (int c = c0; c < c1; ++c) {
        s_pre[c] = run;
        run += chunk_count[c];
    }
    const int S = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    if (tid == 0) s_pre[nchunks] = S;

    if (S == 0) {
        if (blockIdx.x == 0 && tid == 0) {
            atomicExch(&state[ST_DONE], 1);
            if (hflags) publish(hflags + 1, seq);
        }
        return;
    }
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
    const int ncol = 2 * nt + 1;
    // column items (spike, lag): contiguous ranges per wave; residual items
    // (spike, channel share): dealt round-robin over all waves
    const int n_items = S * ncol;               // < 2^31: S <= nchunks * kSlotCap
    const int n_xres = S * kXresItems;
    const int nwaves = gridDim.x * (kUpdThreads / 64);
    const int per = (n_items + nwaves - 1) / nwaves;
    const int gw0 = blockIdx.x * (kUpdThreads / 64);
    if ((long long)gw0 * per >= n_items && gw0 >= n_xres) return;   // no work here
    __syncthreads();

    // record this iteration's spikes: block b takes spikes b, b + grid, ...
    for (int q = blockIdx.x * kUpdThreads + tid; q < S; q += gridDim.x * kUpdThreads) {
        const Slot sl = spike_at(q, s_pre, nchunks, slots);
        const size_t pos = (size_t)kb + q;
        st[2 * pos] = sl.x;
        st[2 * pos + 1] = sl.iy;
        amps[pos] = sl.amp;
        th_amps[pos] = sl.tha;
    }

    const int gw = gw0 + wv;
    const int i0 = (int)min((long long)gw * per, (long long)n_items);
    const int i1 = min(n_items, i0 + per);

    // columns: the wave's items, spike by spike
    if (i0 < i1) {
        const int q0 = i0 / ncol, q1 = (i1 - 1) / ncol;
        for (int q = q0; q <= q1; ++q) {
            const int d_lo = q == q0 ? i0 % ncol : 0;
            const int d_hi = q == q1 ? (i1 - 1) % ncol : ncol - 1;
            SpikeCtx c;
            spike_ctx(c, q, S, s_pre, nchunks, slots);
            spike_columns(q, d_lo, d_hi, c, nt, T, Bt, ldbt, K, G, cc, s_nm, s_pid, gkey, cfmax,
                          imax);
        }
    }

    // residual items j = gw, gw + nwaves, ...: item k of spike q takes
    // channels [cl0, cl1) of the template's nonzero range (at most 64 at a
    // time: lane l holds U[id, :, cl0 + l], broadcast per channel with readlane)
    const int h = nt / 2;
    for (int j = gw; j < n_xres; j += nwaves) {
        const int q = j / kXresItems;
        const int k0 = j % kXresItems, k1 = k0 + 1;
        // lanes 0, 1, 2: spikes q - 1, q, q + 1
        const Slot rw = spike_lanes(q - 1, 3, S, s_pre, nchunks, slots);
        const int id = rl(rw.iy, 1), x = rl(rw.x, 1);
        // windows closer than nt samples (exact ties) overlap a neighbour's
        const bool has_p = q > 0, has_n = q + 1 < S;
        if ((has_p && x - rl(rw.x, 0) < nt) || (has_n && rl(rw.x, 2) - x < nt)) {
            for (int k = k0; k < k1; ++k)
                xres_pair(q, k, rw, has_p, has_n && rl(rw.x, 2) - x < nt, Xres, ldx, U, W,
                          crange, P, C, nt);
            continue;
        }
        const float amp = rl(rw.amp, 1);
        const unsigned crg = rl(rw.crg, 1);
        const int clo = (int)(crg & 0xffffu), chi = (int)(crg >> 16);
        const int cw = (chi - clo + kXresItems - 1) / kXresItems;
        const float *uj = U + (size_t)id * P * C;
        for (int k = k0; k < k1; ++k) {
            const int cl0 = clo + k * cw, cl1 = min(chi, cl0 + cw);
            for (int cs = cl0; cs < cl1; cs += 64) {
                const int nch = min(cl1 - cs, 64);
                float ul[8];
#pragma unroll
                for (int p = 0; p < 8; ++p)
                    ul[p] = (p < P && lane < nch) ? uj[(size_t)p * C + cs + lane] : 0.f;
                for (int kb0 = 0; kb0 < nt; kb0 += 64) {
                    const int kk = kb0 + lane;
                    const bool kin = kk < nt;
                    float wk[8];
#pragma unroll
                    for (int p = 0; p < 8; ++p) wk[p] = (p < P && kin) ? W[p * nt + kk] : 0.f;
                    float *xb = Xres + x - h + (kin ? kk : 0) + (size_t)cs * ldx;
                    for (int cb = 0; cb < nch; cb += kXresCh) {
                        // all loads of the step first (clamped), then compute and
                        // masked stores
                        float xv[kXresCh];
#pragma unroll
                        for (int u = 0; u < kXresCh; ++u) {
                            const int c = cb + u < nch ? cb + u : nch - 1;
                            xv[u] = xb[(size_t)c * ldx];
                        }
#pragma unroll
                        for (int u = 0; u < kXresCh; ++u) {
                            const int c = cb + u < nch ? cb + u : nch - 1;
                            float tm = 0.f;
#pragma unroll
                            for (int p = 0; p < 8; ++p)
                                if (p < P)
                                    tm = __builtin_fmaf(
                                        __int_as_float(__builtin_amdgcn_readlane(
                                            __float_as_int(ul[p]), c)),
                                        wk[p], tm);
                            if (cb + u < nch && kin) {
                                const float prod = amp * tm;
                                xb[(size_t)c * ldx] = xv[u] - prod;
                            }
                        }
                    }
                }
            }
        }
    }
}

// Gram block of every template, by template id.
__global__ void gram_refs_kernel(const int32_t *__restrict__ rank,
                                 const int32_t *__restrict__ lohi,
                                 const int64_t *__restrict__ off, const float *__restrict__ nm,
                                 const int *__restrict__ crange, int K,
                                 GramRef *__restrict__ gref) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= K) return;
    const int jr = rank[j];
    GramRef g;
    g.rank = jr;
    g.nm = nm[j];
    g.lo = lohi[2 * jr];
    g.hi = lohi[2 * jr + 1];
    g.off = off[jr];
    g.clo = crange[2 * j];
    g.chi = crange[2 * j + 1];
    gref[j] = g;
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
    info[3] = state[ST_RESCORED];
}

// tF[s][n][p] = sum_k Xres[iCC[n, iU[id]], x + k - nt/2] W[p,k] + amps[s] Ucc[n,id,p]
// One wave per spike: lanes = lags load the spike's nC residual rows
// (coalesced) into LDS, then lane j < nC*P computes feature (n, p) = (j / P,
// j % P) from LDS.  W sits in LDS for the whole block.
constexpr int kFeatWaves = 4;
constexpr int kFeatMaxNC = 32;
__global__ __launch_bounds__(kFeatWaves * 64) void features_kernel(
    const float *__restrict__ Xres, int ldx, const float *__restrict__ W,
    const int64_t *__restrict__ st, const float *__restrict__ amps, int nsp,
    const int64_t *__restrict__ iCC, const int64_t *__restrict__ iU, const float *__restrict__ Ucc,
    int nC, int K, int P, int C, int nt, float *__restrict__ tF, const int32_t *__restrict__ count) {
    extern __shared__ float f_smem[];
    float *s_w = f_smem;                                  // [P][nt]
    float *s_x = f_smem + P * nt + (threadIdx.x >> 6) * nC * nt;   // this wave: [nC][nt]
    for (int e = threadIdx.x; e < P * nt; e += blockDim.x) s_w[e] = W[e];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // nsp spikes, or the device count (a peel's info[0]) when given, at most nsp
    const int n = count ? min(uni((int)count[0]), nsp) : nsp;
    for (int s = blockIdx.x * kFeatWaves + (threadIdx.x >> 6); s < n;
         s += gridDim.x * kFeatWaves) {
        const int x = (int)st[2 * s];
        const int id = (int)st[2 * s + 1];
        const int iu = (int)iU[id];
        for (int r = 0; r < nC; ++r) {
            const int ch = (int)iCC[(size_t)r * C + iu];
            const float *xr = Xres + (size_t)ch * ldx + x - nt / 2;
            for (int k = lane; k < nt; k += 64) s_x[r * nt + k] = xr[k];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const float a = amps[s];
        for (int j = lane; j < nC * P; j += 64) {
            const int r = j / P, p = j % P;
            const float *xs = s_x + r * nt;
            const float *ws = s_w + p * nt;
            float f = 0.f;
            for (int k = 0; k < nt; ++k) f = __builtin_fmaf(xs[k], ws[k], f);
            const float own = a * Ucc[((size_t)r * K + id) * P + p];
            tF[((size_t)s * nC + r) * P + p] = f + own;
        }
        __builtin_amdgcn_wave_barrier();             // s_x is rewritten for the next spike
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

// Host-side end of the peel loop.  The update kernel of iteration it writes
// (seq, it) to flags[0] when it starts and seq to flags[1] once an iteration
// has accepted nothing; the host keeps at most kLag iterations enqueued ahead
// of the device and stops enqueuing once the done word carries its call's
// sequence number (words from earlier calls carry older numbers).  The flags
// live in pinned host memory, one pair per host thread.
constexpr int kLag = 6;

struct HostFlags {
    u64 *host = nullptr;    // host view
    u64 *dev = nullptr;     // device view
    unsigned seq = 0;
};

HostFlags &host_flags() {
    static thread_local HostFlags f;
    static thread_local bool tried = false;
    if (!tried) {
        tried = true;
        void *p = nullptr;
        if (hipHostMalloc(&p, 256,
                          hipHostMallocCoherent | hipHostMallocMapped | hipHostMallocPortable) ==
            hipSuccess) {
            void *d = nullptr;
            if (hipHostGetDevicePointer(&d, p, 0) == hipSuccess) {
                f.host = static_cast<u64 *>(p);
                f.dev = static_cast<u64 *>(d);
                f.host[0] = 0;
                f.host[1] = 0;
            } else {
                (void)hipHostFree(p);
            }
        }
        (void)hipGetLastError();
    }
    return f;
}

// Waits until the device has started iteration >= want or finished the loop
// (true: finished).  The host runs kLag iterations ahead, tens of
// microseconds each, so the wait spins with a pause hint for its first
// 100 us and then yields the core between polls.  Gives up after a bounded
// wait (sets *gave_up; the caller then enqueues the remaining iterations,
// which return at once past the end).
bool wait_progress(const HostFlags &f, unsigned seq, int want, bool *gave_up) {
    const auto t0 = std::chrono::steady_clock::now();
    bool yielding = false;
    for (unsigned spin = 0;; ++spin) {
        const u64 done = __atomic_load_n(f.host + 1, __ATOMIC_ACQUIRE);
        if ((unsigned)done == seq) return true;
        const u64 pr = __atomic_load_n(f.host, __ATOMIC_ACQUIRE);
        if ((unsigned)(pr >> 32) == seq && (int)(unsigned)pr >= want) return false;
        if (yielding) std::this_thread::yield();
        else __builtin_ia32_pause();
        if ((spin & 255) == 255) {
            const auto waited = std::chrono::steady_clock::now() - t0;
            if (waited > std::chrono::seconds(2)) {
                *gave_up = true;
                return false;
            }
            yielding = waited > std::chrono::microseconds(100);
        }
    }
}

// Scratch of ksm_peel_static: the peel state a workspace holds for ksm_peel,
// in one device array per GPU (held until the process ends).  Host code never
// passes the array's name to a kernel: on the host it is a stand-in symbol,
// not a device address.  The device address is looked up once per device
// (with that device current) and kept.
constexpr size_t kPeelScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kPeelScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int peel_scratch_address(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return ksm::hip_fail(e, "ksm_peel_static: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
        return ksm::fail(KSM_ERR_INVALID, "ksm_peel_static: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return ksm::hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                 "ksm_peel_static: hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return KSM_OK;
}

}  // namespace

extern "C" {

size_t ksm_peel_workspace(int K, int T, int max_peels) {
    if (K <= 0 || T <= 0 || max_peels < 0) return 0;
    const PeelLayout L = peel_layout(K, T, max_peels);
    return ksm::align_up((size_t)T * L.G * 8, 256) + ksm::align_up((size_t)T * 4, 256) * 2 +
           ksm::align_up((size_t)L.Kg * 4, 256) * 2 + ksm::align_up(L.nchunks * 4, 256) +
           ksm::align_up((size_t)L.nchunks * kSlotCap * sizeof(Slot), 256) +
           ksm::align_up((size_t)K * sizeof(GramRef), 256) +
           ksm::align_up((size_t)L.n_state * 4, 256) + ksm::align_up(2 * (size_t)K * 4, 256) +
           256;
}

// Validation shared by ksm_peel and ksm_peel_static.
static bool peel_args_ok(const float *Bt, int ldbt, const float *Xres, int ldx, const float *U,
                         const float *W, const float *nm, const int32_t *perm,
                         const int32_t *rank, const float *cc, const int32_t *lohi,
                         const int64_t *off, int K, int P, int C, int T, int nt, int max_peels,
                         int flags, const void *bound, const int64_t *st, const float *amps,
                         const float *th_amps, int capacity, const int32_t *info) {
    // re-scoring (reduced-precision scores) needs their bound, and only then is one given
    if (((flags & KSM_PEEL_RESCORE) != 0) != (bound != nullptr)) return false;
    return Bt && Xres && U && W && nm && perm && rank && cc && lohi && off && st && amps &&
           th_amps && info && K > 0 && K <= kMaxK && ldbt >= K && (ldbt % 4) == 0 && P > 0 &&
           P <= 8 && C > 0 && C <= kMaxC && nt > 0 && nt <= kMaxNt && (nt % 2) == 1 &&
           T >= 2 * nt + 1 && T <= kChunk * kMaxChunks && ldx >= T && max_peels >= 0 &&
           capacity >= 0 && (flags & ~KSM_PEEL_RESCORE) == 0;
}

// The peel on a validated argument set with its scratch state at `workspace`
// (at least ksm_peel_workspace bytes).  may_poll: the host may follow the
// device's progress through its pinned progress words (allocated on a host
// thread's first call); otherwise every iteration up to max_peels is enqueued.
static int peel_run(float *Bt, int ldbt, float *Xres, int ldx, const float *U, const float *W,
                    const float *nm, const int32_t *perm, const int32_t *rank, const float *cc,
                    const int32_t *lohi, const int64_t *off, int K, int P, int C, int T, int nt,
                    float Th, int max_peels, int flags, const void *bound, int64_t *st,
                    float *amps, float *th_amps, int capacity, int32_t *info, void *workspace,
                    hipStream_t s, bool may_poll, int it_begin, int it_end) {
    // the kernels are timed only for the entry point that may allocate (events)
    const unsigned prof_group = may_poll ? ksm::kProfPeel : 0u;
    const PeelLayout L = peel_layout(K, T, max_peels);
    ksm::Carver cv(workspace);
    u64 *gkey = cv.take<u64>((size_t)T * L.G);
    float *cfmax = cv.take<float>(T);
    int *imax = cv.take<int>(T);
    float *pnm = cv.take<float>(L.Kg);
    int *ppid = cv.take<int>(L.Kg);
    int *chunk_count = cv.take<int>(L.nchunks);
    Slot *slots = cv.take<Slot>((size_t)L.nchunks * kSlotCap);
    GramRef *gref = cv.take<GramRef>(K);
    int *state = cv.take<int>(L.n_state);
    int *crange = cv.take<int>(2 * (size_t)K);

    // Th^2 as the reference forms it (Python float, compared in float32)
    const float th2 = (float)((double)Th * (double)Th);
    // the reduced-precision scores' bound (ksm_project_bounded)
    ksm::BoundView bview = {};
    const int *bhdr = nullptr;
    if (bound) {
        bview = ksm::bound_view(bound, C, T, K, nt, 0);
        bhdr = static_cast<const int *>(bound);
    }
    const size_t tab_dyn = (size_t)L.Kg * (sizeof(float) + sizeof(int));
    const size_t upd_dyn = tab_dyn + (size_t)(L.nchunks + 1) * sizeof(int);
    const int grid = update_grid(upd_dyn);
    const size_t sel_dyn = (size_t)K * sizeof(int);

    if (it_begin == 0) {
        // a new peel: state, tables and the first pass over the scores (a
        // later range continues from the state these leave in the workspace)
        ksm::launch(peel_init_kernel, 1, 256, 0, s, state, L.n_state);
        KSM_LAUNCH_CHECK("peel_init_kernel");
        ksm::launch(channel_range_kernel, (K + 3) / 4, 256, 0, s, U, K, P, C, crange);
        KSM_LAUNCH_CHECK("channel_range_kernel");
        ksm::launch(position_tables_kernel, (L.Kg + 255) / 256, 256, 0, s, nm, perm, K, L.Kg, pnm,
                    ppid);
        KSM_LAUNCH_CHECK("position_tables_kernel");
        ksm::launch(gram_refs_kernel, (K + 255) / 256, 256, 0, s, rank, lohi, off, nm, crange, K,
                    gref);
        KSM_LAUNCH_CHECK("gram_refs_kernel");
        ksm::ProfileScope prof("colmax_all", prof_group, s);
        prof.launch(colmax_all_kernel, dim3((T + kFirstCols - 1) / kFirstCols), dim3(kUpdThreads),
                    (unsigned)tab_dyn, Bt, ldbt, pnm, ppid, K, L.G, L.Kg, T, nt, cfmax, imax, gkey);
        KSM_LAUNCH_CHECK("colmax_all_kernel");
    }
    // stop enqueuing once the device reports the end (not under stream capture)
    HostFlags none_flags;
    HostFlags &hf = may_poll ? host_flags() : none_flags;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    const bool capturing = ksm::recording(s) != nullptr ||
        hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone;
    (void)hipGetLastError();
    const bool poll = hf.host != nullptr && !capturing;
    if (poll && ++hf.seq == 0u) ++hf.seq;      // 0 is never a live sequence number
    const unsigned seq = poll ? hf.seq : 0u;
    bool gave_up = false;
    for (int it = it_begin; it < it_end; ++it) {
        if (poll && !gave_up && it >= kLag && wait_progress(hf, seq, it - kLag, &gave_up)) break;
        {
            ksm::ProfileScope prof("peel_select", prof_group, s);
            if (flags & KSM_PEEL_RESCORE)
                prof.launch(peel_select_cert_kernel, dim3(L.nchunks), dim3(kCertThreads), 0u,
                            cfmax, imax, bview, bhdr, static_cast<const u64 *>(gkey), L.G, Bt,
                            ldbt, static_cast<const GramRef *>(gref),
                            static_cast<const float *>(pnm),
                            static_cast<const int *>(ppid), T, nt, th2, Xres, ldx, C, K, state,
                            chunk_count, slots);
            else
                prof.launch(peel_select_kernel<kChunk>, dim3(L.nchunks), dim3(kChunk),
                            (unsigned)sel_dyn, cfmax, imax, Bt, ldbt, gref, T, nt, th2, rank, K,
                            state, chunk_count, slots);
        }
        {
            ksm::ProfileScope prof("peel_update", prof_group, s);
            prof.launch(peel_update_kernel, dim3(grid), dim3(kUpdThreads), (unsigned)upd_dyn, it,
                        Bt, ldbt, Xres, ldx, U, W, crange, P, C, pnm, ppid, cc, K, L.G, L.Kg, T,
                        nt, L.nchunks, chunk_count, slots, state, st, amps, th_amps, capacity,
                        cfmax, imax, gkey, poll ? hf.dev : nullptr, seq);
        }
    }
    KSM_LAUNCH_CHECK("peel loop");
    ksm::launch(peel_finish_kernel, 1, 1, 0, s, state, info);
    KSM_LAUNCH_CHECK("peel_finish_kernel");
    return KSM_OK;
}

int ksm_peel(float *Bt, int ldbt, float *Xres, int ldx, const float *U, const float *W,
             const float *nm, const int32_t *perm, const int32_t *rank, const float *cc,
             const int32_t *lohi, const int64_t *off, int K, int P, int C, int T, int nt,
             float Th, int max_peels, int flags, const void *bound, int64_t *st,
             float *amps, float *th_amps, int capacity, int32_t *info, void *workspace,
             size_t ws_bytes, void *stream) {
    if (!peel_args_ok(Bt, ldbt, Xres, ldx, U, W, nm, perm, rank, cc, lohi, off, K, P, C, T, nt,
                      max_peels, flags, bound, st, amps, th_amps, capacity, info))
        return ksm::fail(KSM_ERR_INVALID, "ksm_peel: bad arguments");
    if (!workspace || ws_bytes < ksm_peel_workspace(K, T, max_peels))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_peel: workspace too small");
    if (ksm::recording(ksm::as_stream(stream)))
        return ksm::fail(KSM_ERR_INVALID,
                         "ksm_peel: the stream is being recorded (ksm_peel_range records)");
    return peel_run(Bt, ldbt, Xres, ldx, U, W, nm, perm, rank, cc, lohi, off, K, P, C, T, nt, Th,
                    max_peels, flags, bound, st, amps, th_amps, capacity, info, workspace,
                    ksm::as_stream(stream), true, 0, max_peels);
}

int ksm_peel_range(float *Bt, int ldbt, float *Xres, int ldx, const float *U, const float *W,
                   const float *nm, const int32_t *perm, const int32_t *rank, const float *cc,
                   const int32_t *lohi, const int64_t *off, int K, int P, int C, int T, int nt,
                   float Th, int max_peels, int flags, const void *bound, int64_t *st,
                   float *amps, float *th_amps, int capacity, int32_t *info, void *workspace,
                   size_t ws_bytes, int it_begin, int it_end, void *stream) {
    if (!peel_args_ok(Bt, ldbt, Xres, ldx, U, W, nm, perm, rank, cc, lohi, off, K, P, C, T, nt,
                      max_peels, flags, bound, st, amps, th_amps, capacity, info) ||
        it_begin < 0 || it_end < it_begin || it_end > max_peels)
        return ksm::fail(KSM_ERR_INVALID, "ksm_peel_range: bad arguments");
    if (!workspace || ws_bytes < ksm_peel_workspace(K, T, max_peels))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_peel_range: workspace too small");
    return peel_run(Bt, ldbt, Xres, ldx, U, W, nm, perm, rank, cc, lohi, off, K, P, C, T, nt, Th,
                    max_peels, flags, bound, st, amps, th_amps, capacity, info, workspace,
                    ksm::as_stream(stream), false, it_begin, it_end);
}

size_t ksm_peel_static_capacity(void) { return kPeelScratchBytes; }

int ksm_peel_static(float *Bt, int ldbt, float *Xres, int ldx, const float *U, const float *W,
                    const float *nm, const int32_t *perm, const int32_t *rank, const float *cc,
                    const int32_t *lohi, const int64_t *off, int K, int P, int C, int T, int nt,
                    float Th, int max_peels, int flags, const void *bound, int64_t *st,
                    float *amps, float *th_amps, int capacity, int32_t *info, void *stream) {
    if (!peel_args_ok(Bt, ldbt, Xres, ldx, U, W, nm, perm, rank, cc, lohi, off, K, P, C, T, nt,
                      max_peels, flags, bound, st, amps, th_amps, capacity, info))
        return ksm::fail(KSM_ERR_INVALID, "ksm_peel_static: bad arguments");
    const size_t need = ksm_peel_workspace(K, T, max_peels);
    if (need > kPeelScratchBytes)
        return ksm::fail(KSM_ERR_WORKSPACE,
                         "ksm_peel_static: K=%d, T=%d, max_peels=%d need %zu bytes of scratch, "
                         "more than the library's static %zu (use ksm_peel with a workspace)",
                         K, T, max_peels, need, kPeelScratchBytes);
    void *ws = nullptr;
    const int rc = peel_scratch_address(&ws);
    if (rc != KSM_OK) return rc;
    return peel_run(Bt, ldbt, Xres, ldx, U, W, nm, perm, rank, cc, lohi, off, K, P, C, T, nt, Th,
                    max_peels, flags, bound, st, amps, th_amps, capacity, info, ws,
                    ksm::as_stream(stream), false, 0, max_peels);
}

int ksm_features(const float *Xres, int ldx, const float *W, const int64_t *st,
                 const float *amps, int nsp, const int64_t *iCC, const int64_t *iU,
                 const float *Ucc, int nC, int K, int P, int C, int nt, float *tF, void *stream) {
    if (nsp < 0 || nC <= 0 || K <= 0 || P <= 0 || C <= 0 || nt <= 0 || nC > C ||
        nC > kFeatMaxNC || nt > kMaxNt)
        return ksm::fail(KSM_ERR_INVALID, "ksm_features: bad arguments");
    if (nsp == 0) return KSM_OK;
    if (!Xres || !W || !st || !amps || !iCC || !iU || !Ucc || !tF)
        return ksm::fail(KSM_ERR_INVALID, "ksm_features: null pointer");
    const size_t smem = (size_t)(P * nt + kFeatWaves * nC * nt) * sizeof(float);
    ksm::ProfileScope prof("features", ksm::kProfPeel, ksm::as_stream(stream));
    prof.launch(features_kernel, dim3((nsp + kFeatWaves - 1) / kFeatWaves), dim3(kFeatWaves * 64),
                (unsigned)smem, Xres, ldx, W, st, amps, nsp, iCC, iU, Ucc, nC, K, P, C, nt, tF,
                static_cast<const int32_t *>(nullptr));
    KSM_LAUNCH_CHECK("features_kernel");
    return KSM_OK;
}

int ksm_features_counted(const float *Xres, int ldx, const float *W, const int64_t *st,
                         const float *amps, const int32_t *info, int capacity,
                         const int64_t *iCC, const int64_t *iU, const float *Ucc, int nC, int K,
                         int P, int C, int nt, float *tF, void *stream) {
    if (capacity < 0 || nC <= 0 || K <= 0 || P <= 0 || C <= 0 || nt <= 0 || nC > C ||
        nC > kFeatMaxNC || nt > kMaxNt)
        return ksm::fail(KSM_ERR_INVALID, "ksm_features_counted: bad arguments");
    if (capacity == 0) return KSM_OK;
    if (!Xres || !W || !st || !amps || !info || !iCC || !iU || !Ucc || !tF)
        return ksm::fail(KSM_ERR_INVALID, "ksm_features_counted: null pointer");
    const size_t smem = (size_t)(P * nt + kFeatWaves * nC * nt) * sizeof(float);
    // a fixed grid whose waves step over the spikes the device reports
    const int blocks = (min(capacity, 8192) + kFeatWaves - 1) / kFeatWaves;
    ksm::ProfileScope prof("features", ksm::kProfPeel, ksm::as_stream(stream));
    prof.launch(features_kernel, dim3(blocks), dim3(kFeatWaves * 64), (unsigned)smem, Xres, ldx, W,
                st, amps, capacity, iCC, iU, Ucc, nC, K, P, C, nt, tF, info);
    KSM_LAUNCH_CHECK("features_kernel");
    return KSM_OK;
}

}  // extern "C"
