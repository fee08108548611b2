// Output-format kernels: the per-spike work of the Phy export.
//
// Reference: kilosort/postprocessing.py:9-30 (remove_duplicates),
// kilosort/template_matching.py:361-379 (roll_features),
// :65-130 (make_pc_features, with clustering_qr.py:560-600 get_data_cpu),
// kilosort/CCG.py:8-36 (compute_CCG) and :105-120 (refract), as called by
// kilosort/io.py:248-469 (save_to_phy).
//
// * ksm_group: a stable grouping of spike indices by label (cluster id):
//   order[offsets[g] .. offsets[g+1]) lists the spikes of label g in input
//   order.  Three passes: per-chunk label counts in LDS, a per-label scan
//   over chunks, and an in-order placement where each wave ranks equal labels
//   among its 64 lanes by bitwise ballots.
// * ksm_remove_duplicates: one wave per label walks its spikes in input
//   order, 64 at a time, and keeps a spike when it is at least dt samples
//   after the last kept spike of its label (the first spike of a label is
//   kept).
// * ksm_acg: autocorrelogram counts of every active label: for every spike j
//   of a label, the spikes k of the same label with t_j - dt < t_k <=
//   t_j + dt (dt = nbins * tbin) land in bin rint((t_j - t_k) / tbin) +
//   nbins (times in seconds, float64, sorted within each label), as the
//   reference's two-pointer loop.  Integer counts, so the result is exact.
// * ksm_ccg: the same counts for two sorted trains (t2 against t1).
// * ksm_pc_map: the PC features of chosen spikes times a P x P matrix (the
//   time shift of merging_function's roll_features, expressed in PC space).
// * ksm_template_sums: per detection template, the double sums of the PC
//   features of its spikes and their count, and which templates each label
//   holds (the inputs of the per-cluster channel ranking).
// * ksm_pc_channel_rank: the per-cluster channel ranking of make_pc_features
//   replayed in cluster order on those sums (one block).
// * ksm_gather_features: each spike's features re-indexed by its template's
//   channel map (sigma[t][k] = source channel row, or -1 for zeros).
#include "ksm_common.h"

namespace {

constexpr int kGrpThreads = 64;            // one wave per grouping block
constexpr int kMaxGroups = 16384;          // label counters in 64 KB of LDS
constexpr long long kMaxSpikes = 0x7fffffffLL;
constexpr int kMaxBins = 1023;             // 2 nbins + 1 <= 2047 ACG bins

long long group_chunk(long long n, int G) {
    // chunks of at least 4096 spikes, with at most 2^24 chunk x label counters
    long long chunk = 4096;
    while (((n + chunk - 1) / chunk) * (long long)G > (1LL << 24)) chunk *= 2;
    return chunk;
}

__global__ void zero_bytes_kernel(uint8_t *__restrict__ p, size_t nbytes) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes; i += stride)
        p[i] = 0;
}

int zero_async(void *p, size_t nbytes, hipStream_t s) {
    if (!nbytes) return KSM_OK;
    const size_t blocks = std::min<size_t>((nbytes + 255) / 256, 4096);
    zero_bytes_kernel<<<(unsigned)blocks, 256, 0, s>>>(static_cast<uint8_t *>(p), nbytes);
    KSM_LAUNCH_CHECK("zero_bytes_kernel");
    return KSM_OK;
}

// counts[b][g] = spikes of label g in chunk b.
__global__ __launch_bounds__(kGrpThreads) void group_count_kernel(
    const int32_t *__restrict__ lab, long long n, int G, long long chunk,
    uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t s_hist[];
    const int lane = threadIdx.x;
    for (int g = lane; g < G; g += kGrpThreads) s_hist[g] = 0;
    __syncthreads();
    const long long b0 = (long long)blockIdx.x * chunk;
    const long long b1 = b0 + chunk < n ? b0 + chunk : n;
    for (long long i = b0 + lane; i < b1; i += kGrpThreads) {
        const int c = lab[i];
        if ((unsigned)c < (unsigned)G) atomicAdd(&s_hist[c], 1u);
    }
    __syncthreads();
    uint32_t *row = counts + (size_t)blockIdx.x * G;
    for (int g = lane; g < G; g += kGrpThreads) row[g] = s_hist[g];
}

// counts[b][g] -> spikes of label g in chunks before b; total[g].
__global__ void group_scan_kernel(uint32_t *__restrict__ counts, long long nb, int G,
                                  int64_t *__restrict__ total) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    uint32_t run = 0;
    for (long long b = 0; b < nb; ++b) {
        uint32_t *c = counts + (size_t)b * G + g;
        const uint32_t v = *c;
        *c = run;
        run += v;
    }
    total[g] = run;
}

// offsets[0] = 0, offsets[g + 1] = total[0] + ... + total[g]; one block.
__global__ __launch_bounds__(1024) void group_offsets_kernel(const int64_t *__restrict__ total,
                                                             int G, int64_t *__restrict__ off) {
    __shared__ int64_t s[1024];
    const int t = threadIdx.x;
    const int per = (G + 1023) / 1024;
    const int g0 = t * per, g1 = g0 + per < G ? g0 + per : G;
    int64_t local = 0;
    for (int g = g0; g < g1; ++g) local += total[g];
    s[t] = local;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t v = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int64_t run = s[t] - local;
    if (t == 0) off[0] = 0;
    for (int g = g0; g < g1; ++g) {
        run += total[g];
        off[g + 1] = run;
    }
}

// order[offsets[g] + rank] = i, rank = spikes of label g before i.
__global__ __launch_bounds__(kGrpThreads) void group_place_kernel(
    const int32_t *__restrict__ lab, long long n, int G, int bits, long long chunk,
    const uint32_t *__restrict__ before, const int64_t *__restrict__ off,
    int64_t *__restrict__ order) {
    extern __shared__ uint32_t s_base[];
    const int lane = threadIdx.x;
    const uint32_t *row = before + (size_t)blockIdx.x * G;
    for (int g = lane; g < G; g += kGrpThreads) s_base[g] = row[g];
    __syncthreads();
    const long long b0 = (long long)blockIdx.x * chunk;
    const long long b1 = b0 + chunk < n ? b0 + chunk : n;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (long long i0 = b0; i0 < b1; i0 += kGrpThreads) {
        const long long i = i0 + lane;
        const int c = i < b1 ? lab[i] : -1;
        const bool ok = (unsigned)c < (unsigned)G;
        uint64_t eq = __ballot(ok);
        for (int bit = 0; bit < bits; ++bit) {
            const bool on = (c >> bit) & 1;
            const uint64_t m = __ballot(ok && on);
            eq &= on ? m : ~m;
        }
        if (ok) order[off[c] + s_base[c] + __popcll(eq & lt)] = i;
        __syncthreads();
        if (ok && (eq & lt) == 0) s_base[c] += (uint32_t)__popcll(eq);
        __syncthreads();
    }
}

// One wave per label.  Per chunk of 64 consecutive spikes of the label, the
// kept ones are found in order by ballots: the next kept spike is the first
// lane after the previous kept one with t >= t0 + dt, and t0 becomes its time.
__global__ __launch_bounds__(64) void dedup_kernel(const int64_t *__restrict__ times,
                                                   const int64_t *__restrict__ off,
                                                   const int64_t *__restrict__ order, int G,
                                                   long long dt, uint8_t *__restrict__ keep) {
    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const long long a = off[g], b = off[g + 1];
    if (a == b) return;
    long long t0 = times[order[a]] - dt;
    for (long long p0 = a; p0 < b; p0 += 64) {
        const long long p = p0 + lane;
        const bool valid = p < b;
        const long long i = valid ? order[p] : 0;
        const long long t = valid ? times[i] : 0;
        uint64_t rest = __ballot(valid);
        uint64_t kept = 0;
        while (true) {
            const uint64_t cand = __ballot(valid && t >= t0 + dt) & rest;
            if (!cand) break;
            const int l = __ffsll((unsigned long long)cand) - 1;
            kept |= 1ull << l;
            t0 = __shfl(t, l, 64);
            rest = l == 63 ? 0ull : rest & ~((2ull << l) - 1ull);
        }
        if (valid) keep[i] = (kept >> lane) & 1ull ? 1 : 0;
    }
}

__device__ __forceinline__ int group_of(const int64_t *__restrict__ off, int G, long long p) {
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (off[mid + 1] > p) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ double time_at(const double *__restrict__ times,
                                          const int64_t *__restrict__ order, long long p) {
    return times[order[p]];
}

// First position in [a, b) whose time exceeds x (times sorted within a group).
__device__ __forceinline__ long long first_above(const double *__restrict__ times,
                                                 const int64_t *__restrict__ order, long long a,
                                                 long long b, double x) {
    while (a < b) {
        const long long mid = (a + b) >> 1;
        if (time_at(times, order, mid) > x) b = mid;
        else a = mid + 1;
    }
    return a;
}

// One thread per grouped position j; a block whose 256 positions share one
// label counts in LDS and adds its nonzero bins once.
__global__ __launch_bounds__(256) void acg_kernel(const double *__restrict__ times,
                                                  const int64_t *__restrict__ order, long long n,
                                                  const int64_t *__restrict__ off, int G,
                                                  const uint8_t *__restrict__ active, int nbins,
                                                  double tbin, uint32_t *__restrict__ counts) {
    __shared__ uint32_t s_h[2 * kMaxBins + 1];
    const int L = 2 * nbins + 1;
    const long long p0 = (long long)blockIdx.x * 256;
    const long long plast = (p0 + 256 < n ? p0 + 256 : n) - 1;
    const int g0 = group_of(off, G, p0), g1 = group_of(off, G, plast);
    const bool single = g0 == g1;
    if (single) {
        if (!active[g0]) return;
        for (int b = threadIdx.x; b < L; b += 256) s_h[b] = 0;
        __syncthreads();
    }
    const long long p = p0 + threadIdx.x;
    if (p < n) {
        const int g = single ? g0 : group_of(off, G, p);
        if (active[g]) {
            const long long a = off[g], b = off[g + 1];
            const double dt = nbins * tbin;
            const double tj = time_at(times, order, p);
            const long long kl = first_above(times, order, a, b, tj - dt);
            const long long kh = first_above(times, order, kl, b, tj + dt);
            uint32_t *row = counts + (size_t)g * L;
            for (long long k = kl; k < kh; ++k) {
                const double d = (tj - time_at(times, order, k)) / tbin;
                int ib = (int)rint(d) + nbins;
                ib = ib < 0 ? 0 : (ib >= L ? L - 1 : ib);
                if (single) atomicAdd(&s_h[ib], 1u);
                else atomicAdd(&row[ib], 1u);
            }
        }
    }
    if (single) {
        __syncthreads();
        uint32_t *row = counts + (size_t)g0 * L;
        for (int b = threadIdx.x; b < L; b += 256)
            if (s_h[b]) atomicAdd(&row[b], s_h[b]);
    }
}

// Cross-correlogram of two sorted trains: one thread per spike j of t2; the
// block counts in LDS and adds its nonzero bins once.
__global__ __launch_bounds__(256) void ccg_kernel(const double *__restrict__ t1, long long n1,
                                                  const double *__restrict__ t2, long long n2,
                                                  int nbins, double tbin,
                                                  uint32_t *__restrict__ counts) {
    __shared__ uint32_t s_h[2 * kMaxBins + 1];
    const int L = 2 * nbins + 1;
    for (int b = threadIdx.x; b < L; b += 256) s_h[b] = 0;
    __syncthreads();
    const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
    if (j < n2) {
        const double dt = nbins * tbin;
        const double tj = t2[j];
        long long a = 0, b = n1;
        const double lo_x = tj - dt, hi_x = tj + dt;
        while (a < b) {                       // first t1 > tj - dt
            const long long mid = (a + b) >> 1;
            if (t1[mid] > lo_x) b = mid;
            else a = mid + 1;
        }
        const long long kl = a;
        b = n1;
        while (a < b) {                       // first t1 > tj + dt
            const long long mid = (a + b) >> 1;
            if (t1[mid] > hi_x) b = mid;
            else a = mid + 1;
        }
        for (long long k = kl; k < a; ++k) {
            int ib = (int)rint((tj - t1[k]) / tbin) + nbins;
            ib = ib < 0 ? 0 : (ib >= L ? L - 1 : ib);
            atomicAdd(&s_h[ib], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < L; b += 256)
        if (s_h[b]) atomicAdd(&counts[b], s_h[b]);
}

// X[rows[i]][r][:] = X[rows[i]][r][:] @ M for every row r < R (M [P][P],
// float32, products summed in p order); one thread per (spike, row).
__global__ void pc_map_kernel(float *__restrict__ X, const int64_t *__restrict__ rows,
                              long long nrows, int R, int P, const float *__restrict__ M) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrows * R) return;
    const long long i = e / R;
    const int r = (int)(e - i * R);
    float *x = X + ((size_t)rows[i] * R + r) * P;
    float v[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) v[p] = p < P ? x[p] : 0.f;
    for (int q = 0; q < P; ++q) {
        float acc = 0.f;
#pragma unroll
        for (int p = 0; p < 16; ++p)
            if (p < P) acc = __builtin_fmaf(v[p], M[p * P + q], acc);
        x[q] = acc;
    }
}

// Per-template sums of the spikes' feature rows, by a counting sort of the
// spikes by template and one block per template summing its spikes' rows in
// float64 -- each feature row is read once, with no floating-point atomics.
// The sort: each of NB blocks takes a contiguous range of spikes and counts
// them per template in LDS (H[b][t]); per template, the column prefix over
// the blocks (O[b][t]) and the total; the exclusive prefix of the totals
// over templates (off[t]); then each block places its spikes at off[t] +
// O[b][t] + their rank in the block (an LDS cursor).  float64 sums of
// float32 values are exact while the values' exponents and the count span
// fewer than 29 bits, so the order of a template's spikes within a block
// (set by the LDS cursor) does not change the sums.
constexpr int kTsThreads = 1024;
constexpr int kTsMaxK = 16384;              // LDS: one int per template (64 KB)

__global__ __launch_bounds__(kTsThreads) void ts_hist_kernel(const int64_t *__restrict__ tmpl,
                                                             long long n, long long per, int K,
                                                             int *__restrict__ H) {
    extern __shared__ int s_h[];
    for (int k = threadIdx.x; k < K; k += kTsThreads) s_h[k] = 0;
    __syncthreads();
    const long long s0 = (long long)blockIdx.x * per, s1 = min(n, s0 + per);
    for (long long s = s0 + threadIdx.x; s < s1; s += kTsThreads) atomicAdd(&s_h[tmpl[s]], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += kTsThreads) H[(size_t)blockIdx.x * K + k] = s_h[k];
}

// O[b][t] = sum of H[b'][t] over b' < b, tot[t] = the sum over all blocks:
// a wave per template, lanes = blocks, 64 at a time (a shuffle scan).
__global__ __launch_bounds__(256) void ts_columns_kernel(const int *__restrict__ H, int NB, int K,
                                                         int *__restrict__ O,
                                                         int *__restrict__ tot) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= K) return;                        // wave-uniform
    int run = 0;
    for (int b0 = 0; b0 < NB; b0 += 64) {
        const int b = b0 + lane;
        const int h = b < NB ? H[(size_t)b * K + t] : 0;
        int incl = h;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        if (b < NB) O[(size_t)b * K + t] = run + incl - h;
        run += __shfl(incl, 63, 64);
    }
    if (lane == 0) tot[t] = run;
}

// off[t] = spikes of templates before t (off[K] = n).  One block of 1024
// threads, each taking consecutive templates.
__global__ __launch_bounds__(1024) void ts_scan_kernel(const int *__restrict__ cnt, int K,
                                                       int *__restrict__ off) {
    __shared__ int s_w[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int per = (K + 1023) / 1024;
    const int k0 = tid * per, k1 = min(K, k0 + per);
    int mine = 0;
    for (int k = k0; k < k1; ++k) mine += cnt[k];
    int incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    int base = 0;
    for (int i = 0; i < w; ++i) base += s_w[i];
    int run = base + incl - mine;
    for (int k = k0; k < k1; ++k) {
        off[k] = run;
        run += cnt[k];
    }
    if (tid == 1023) off[K] = base + incl;
}

__global__ __launch_bounds__(kTsThreads) void ts_place_kernel(const int64_t *__restrict__ tmpl,
                                                              long long n, long long per, int K,
                                                              const int *__restrict__ off,
                                                              const int *__restrict__ O,
                                                              int *__restrict__ list) {
    extern __shared__ int s_c[];
    for (int k = threadIdx.x; k < K; k += kTsThreads) s_c[k] = 0;
    __syncthreads();
    const long long s0 = (long long)blockIdx.x * per, s1 = min(n, s0 + per);
    const int *ob = O + (size_t)blockIdx.x * K;
    for (long long s = s0 + threadIdx.x; s < s1; s += kTsThreads) {
        const int t = (int)tmpl[s];
        list[off[t] + ob[t] + atomicAdd(&s_c[t], 1)] = (int)s;
    }
}

// Block t: sums[t][r] = sum of tF[s][r] over the spikes s of template t
// (lanes = r, up to 64 at a time; the block's 8 waves take every 8th spike,
// 8 rows in flight per wave, spike indices staged in LDS), the 8 partial
// sums added in wave order; counts[t] and presence[labels[s]][t].
constexpr int kTsWaves = 8;
constexpr int kTsStage = 2048;
__global__ __launch_bounds__(kTsWaves * 64) void ts_reduce_kernel(
    const float *__restrict__ tF, const int *__restrict__ list, const int *__restrict__ off,
    const int32_t *__restrict__ labels, int R, int K, double *__restrict__ sums,
    unsigned long long *__restrict__ counts, uint8_t *__restrict__ presence) {
    __shared__ int s_list[kTsStage];
    __shared__ double s_part[kTsWaves][64];
    const int t = blockIdx.x;
    const int b = off[t], e = off[t + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int r0 = 0; r0 < R; r0 += 64) {
        const int r = r0 + lane;
        const bool in = r < R;
        const int rr = in ? r : 0;
        double acc = 0.0;
        for (int c0 = b; c0 < e; c0 += kTsStage) {
            const int cn = min(kTsStage, e - c0);
            __syncthreads();                       // s_list of the previous stage is read
            for (int i = tid; i < cn; i += kTsWaves * 64) s_list[i] = list[c0 + i];
            __syncthreads();
            int i = w;
            for (; i + 7 * kTsWaves < cn; i += 8 * kTsWaves) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = tF[(size_t)s_list[i + u * kTsWaves] * R + rr];
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += (double)v[u];
            }
            for (; i < cn; i += kTsWaves) acc += (double)tF[(size_t)s_list[i] * R + rr];
        }
        s_part[w][lane] = in ? acc : 0.0;
        __syncthreads();
        if (w == 0 && in) {
            double tot = 0.0;
            for (int k = 0; k < kTsWaves; ++k) tot += s_part[k][lane];
            sums[(size_t)t * R + r] = tot;
        }
        __syncthreads();
    }
    if (tid == 0) counts[t] = (unsigned long long)(e - b);
    if (presence)
        for (int i = b + tid; i < e; i += kTsWaves * 64)
            presence[(size_t)labels[list[i]] * K + t] = 1;
}

__global__ void gather_features_kernel(const float *__restrict__ tF,
                                       const int64_t *__restrict__ tmpl, long long n, int nC,
                                       int P, const int32_t *__restrict__ sigma,
                                       float *__restrict__ out, int layout) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * P) return;
    const long long s = e / P;
    const int p = (int)(e - s * P);
    const int32_t *sg = sigma + (size_t)tmpl[s] * nC;
    const float *src = tF + (size_t)s * nC * P + p;
    for (int k = 0; k < nC; ++k) {
        const int q = sg[k];
        const float v = q >= 0 ? src[(size_t)q * P] : 0.f;
        if (layout == 0) out[((size_t)s * nC + k) * P + p] = v;
        else out[((size_t)s * P + p) * nC + k] = v;
    }
}

// The cluster-by-cluster channel ranking of make_pc_features, one block
// walking the clusters in ascending order (each depends on the earlier ones
// through the shared templates' maps).  Per cluster g: the templates it holds
// (presence), the sum over their spikes of each current feature row placed at
// its channel (acc[c][p], float64, from the per-template sums and each
// template's map sigma), the norm of the mean per channel, the n_chans
// channels of largest norm (ties: lowest channel), and each held template's
// map composed with the choice.
constexpr int kRankThreads = 1024;
constexpr int kMaxRankC = 2048;                 // channels a thread's candidates cover

constexpr int kPer = (kMaxRankC + kRankThreads - 1) / kRankThreads;   // channels per thread

// (v2, c2) comes before (v, c): larger norm, or the same norm and a lower
// channel
__device__ __forceinline__ int first(double v2, int c2, double v, int c) {
    return (v2 > v || (v2 == v && c2 < c)) ? 1 : 0;
}

__device__ __forceinline__ double readlane_f64(double x, int m) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), m);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), m);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(kRankThreads) void pc_rank_kernel(
    const double *__restrict__ sums, const unsigned long long *__restrict__ counts,
    const uint8_t *__restrict__ presence, const int64_t *__restrict__ iC, int K, int nC, int P,
    int C, int G, int32_t *__restrict__ sigma, int32_t *__restrict__ sigma_new,
    uint32_t *__restrict__ feature_ind) {
    extern __shared__ double s_dyn[];
    double *s_acc = s_dyn;                                        // [C][P]
    double *s_norm = s_acc + (size_t)C * P;                       // [C]
    int *s_mark = reinterpret_cast<int *>(s_norm + C);            // [C]
    short *s_ic = reinterpret_cast<short *>(s_mark + C);          // [K][nC] channels
    short *s_held = s_ic + (size_t)K * nC;                        // [K] held templates
    __shared__ int s_wc[kRankThreads / 64];
    __shared__ double s_cv[(kRankThreads / 64) * 32];      // per-wave top channels
    __shared__ int s_cc[(kRankThreads / 64) * 32];
    __shared__ int s_top[32];
    __shared__ int s_nheld;
    __shared__ unsigned long long s_N;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int KN = K * nC;
    for (int e = tid; e < KN; e += kRankThreads) {
        const int t = e / nC, j = e - t * nC;
        sigma[e] = j;
        s_ic[e] = (short)iC[(size_t)j * K + t];
    }
    __syncthreads();
    for (int g = 0; g < G; ++g) {
        const uint8_t *pres = presence + (size_t)g * K;
        for (int e = tid; e < C * P; e += kRankThreads) s_acc[e] = 0.0;
        for (int c = tid; c < C; c += kRankThreads) s_mark[c] = 0;
        if (tid == 0) {
            s_N = 0;
            s_nheld = 0;
        }
        __syncthreads();
        // the held templates, compacted in ascending order wave by wave
        for (int t0 = 0; t0 < K; t0 += kRankThreads) {
            const int t = t0 + tid;
            const bool h = t < K && pres[t];
            const uint64_t m = __ballot(h);
            if (lane == 0) s_wc[wv] = __popcll(m);
            __syncthreads();
            int base = s_nheld;
            for (int w = 0; w < wv; ++w) base += s_wc[w];
            if (h) s_held[base + __popcll(m & ((1ull << lane) - 1ull))] = (short)t;
            __syncthreads();
            if (tid == 0) {
                int tot = 0;
                for (int w = 0; w < kRankThreads / 64; ++w) tot += s_wc[w];
                s_nheld += tot;
            }
            __syncthreads();
        }
        const int nh = s_nheld;
        if (nh == 0) continue;                   // label without spikes (block-uniform)
        for (int e = tid; e < nh * nC; e += kRankThreads) {
            const int i = e / nC, j = e - i * nC;
            const int t = s_held[i];
            const int c = s_ic[t * nC + j];
            s_mark[c] = 1;
            const int q = sigma[t * nC + j];
            if (q >= 0)
                for (int p = 0; p < P; ++p)
                    atomicAdd(&s_acc[c * P + p], sums[((size_t)t * nC + q) * P + p]);
            if (j == 0) atomicAdd(&s_N, counts[t]);
        }
        __syncthreads();
        const double dN = (double)s_N;
        for (int c = tid; c < C; c += kRankThreads) {
            double a = 0.0;
            for (int p = 0; p < P; ++p) {
                const double m = s_acc[c * P + p] / dN;
                a += m * m;
            }
            s_norm[c] = s_mark[c] ? sqrt(a) : -1.0;
        }
        __syncthreads();
        // top nC channels by (norm descending, channel ascending), a strict
        // order (NaN norms rank below every channel, above none): each wave
        // ranks its channels by counting, over its lanes read by readlane,
        // those that come first, and keeps ranks < nC in order; then wave 0
        // ranks those candidates the same way
        {
            const int nwr = C > kRankThreads ? kRankThreads / 64 : (C + 63) / 64;
            if (wv < nwr) {                      // wave-uniform
                double lv[kPer];
                int lc[kPer];
#pragma unroll
                for (int u = 0; u < kPer; ++u) {
                    const int c = tid + u * kRankThreads;
                    const double nv = c < C ? s_norm[c] : -3.0;
                    lv[u] = nv != nv ? -2.5 : nv;
                    lc[u] = c;                   // distinct, padding past C
                }
                int rk[kPer] = {};
                const int nu = C > kRankThreads ? kPer : 1;
                for (int u2 = 0; u2 < nu; ++u2)
#pragma unroll 4
                    for (int m = 0; m < 64; ++m) {
                        const double vm = readlane_f64(lv[u2], m);
                        const int cm = __builtin_amdgcn_readlane(lc[u2], m);
#pragma unroll
                        for (int u = 0; u < kPer; ++u) rk[u] += first(vm, cm, lv[u], lc[u]);
                    }
#pragma unroll
                for (int u = 0; u < kPer; ++u)
                    if (rk[u] < nC && (u == 0 || nu == kPer)) {
                        s_cv[wv * 32 + rk[u]] = lv[u];
                        s_cc[wv * 32 + rk[u]] = lc[u];
                    }
            }
            __syncthreads();
            if (wv == 0) {
                const int M = nwr * nC;          // candidates, wave w's at w * 32 + r
                for (int e = lane; e - lane < M; e += 64) {
                    const bool ok = e < M;
                    const int w = ok ? e / nC : 0, r = ok ? e - w * nC : 0;
                    const double v = s_cv[w * 32 + r];
                    const int c = s_cc[w * 32 + r];
                    int rank = 0;
                    for (int m = 0; m < M; ++m) {
                        const int wm = m / nC, idx = wm * 32 + (m - wm * nC);
                        rank += first(s_cv[idx], s_cc[idx], v, c);
                    }
                    if (ok && rank < nC) {
                        s_top[rank] = c;
                        feature_ind[(size_t)g * nC + rank] = (uint32_t)c;
                    }
                }
            }
            __syncthreads();
        }
        for (int e = tid; e < nh * nC; e += kRankThreads) {
            const int i = e / nC, k = e - i * nC;
            const int t = s_held[i];
            const int c = s_top[k];
            int src = -1;
            for (int j = 0; j < nC; ++j)
                if (s_ic[t * nC + j] == c) src = j;
            sigma_new[t * nC + k] = src >= 0 ? sigma[t * nC + src] : -1;
        }
        __syncthreads();
        for (int e = tid; e < nh * nC; e += kRankThreads) {
            const int t = s_held[e / nC];
            const int k = e - (e / nC) * nC;
            sigma[t * nC + k] = sigma_new[t * nC + k];
        }
        __syncthreads();
    }
}

struct GroupWs {
    uint32_t *counts;
    int64_t *total;
    long long nb, chunk;
    size_t bytes;
};

GroupWs group_layout(void *base, long long n, int G) {
    GroupWs w{};
    w.chunk = group_chunk(n, G);
    w.nb = n > 0 ? (n + w.chunk - 1) / w.chunk : 0;
    const size_t c_bytes = ((size_t)w.nb * G * sizeof(uint32_t) + 255) & ~(size_t)255;
    w.bytes = c_bytes + (size_t)G * sizeof(int64_t);
    uint8_t *p = static_cast<uint8_t *>(base);
    w.counts = reinterpret_cast<uint32_t *>(p);
    w.total = reinterpret_cast<int64_t *>(p + c_bytes);
    return w;
}

}  // namespace

extern "C" {

size_t ksm_group_workspace(long long n, int n_groups) {
    if (n < 0 || n > kMaxSpikes || n_groups <= 0 || n_groups > kMaxGroups) return 0;
    return group_layout(nullptr, n, n_groups).bytes;
}

int ksm_group(const int32_t *labels, long long n, int n_groups, int64_t *offsets, int64_t *order,
              void *workspace, size_t ws_bytes, void *stream) {
    if (n < 0 || n > kMaxSpikes || n_groups <= 0 || n_groups > kMaxGroups || !offsets ||
        (n > 0 && (!labels || !order)))
        return ksm::fail(KSM_ERR_INVALID,
                         "ksm_group: bad arguments (0 <= n < 2^31, 1 <= n_groups <= %d)",
                         kMaxGroups);
    const GroupWs w = group_layout(workspace, n, n_groups);
    if (!workspace || ws_bytes < w.bytes)
        return ksm::fail(KSM_ERR_INVALID, "ksm_group: workspace of %zu bytes needed", w.bytes);
    hipStream_t s = ksm::as_stream(stream);
    const size_t lds = (size_t)n_groups * sizeof(uint32_t);
    if (w.nb > 0) {
        group_count_kernel<<<(unsigned)w.nb, kGrpThreads, lds, s>>>(labels, n, n_groups, w.chunk,
                                                                   w.counts);
        KSM_LAUNCH_CHECK("group_count_kernel");
    }
    group_scan_kernel<<<(n_groups + 255) / 256, 256, 0, s>>>(w.counts, w.nb, n_groups, w.total);
    KSM_LAUNCH_CHECK("group_scan_kernel");
    group_offsets_kernel<<<1, 1024, 0, s>>>(w.total, n_groups, offsets);
    KSM_LAUNCH_CHECK("group_offsets_kernel");
    if (w.nb > 0) {
        int bits = 0;
        while ((1 << bits) < n_groups) ++bits;
        group_place_kernel<<<(unsigned)w.nb, kGrpThreads, lds, s>>>(
            labels, n, n_groups, bits, w.chunk, w.counts, offsets, order);
        KSM_LAUNCH_CHECK("group_place_kernel");
    }
    return KSM_OK;
}

int ksm_remove_duplicates(const int64_t *times, long long n, int n_groups, const int64_t *offsets,
                          const int64_t *order, long long dt, uint8_t *keep, void *stream) {
    if (n < 0 || n_groups <= 0 || n_groups > kMaxGroups || !offsets ||
        (n > 0 && (!times || !order || !keep)))
        return ksm::fail(KSM_ERR_INVALID, "ksm_remove_duplicates: bad arguments");
    if (n == 0) return KSM_OK;
    dedup_kernel<<<n_groups, 64, 0, ksm::as_stream(stream)>>>(times, offsets, order, n_groups,
                                                              dt, keep);
    KSM_LAUNCH_CHECK("dedup_kernel");
    return KSM_OK;
}

int ksm_acg(const double *times, long long n, int n_groups, const int64_t *offsets,
            const int64_t *order, const uint8_t *active, int nbins, double tbin, uint32_t *counts,
            void *stream) {
    if (n < 0 || n_groups <= 0 || n_groups > kMaxGroups || !offsets || !active || !counts ||
        (n > 0 && (!times || !order)) || nbins <= 0 || nbins > kMaxBins || !(tbin > 0.0))
        return ksm::fail(KSM_ERR_INVALID, "ksm_acg: bad arguments (1 <= nbins <= %d)", kMaxBins);
    hipStream_t s = ksm::as_stream(stream);
    const int rc = zero_async(counts, (size_t)n_groups * (2 * nbins + 1) * sizeof(uint32_t), s);
    if (rc != KSM_OK || n == 0) return rc;
    acg_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(times, order, n, offsets, n_groups,
                                                          active, nbins, tbin, counts);
    KSM_LAUNCH_CHECK("acg_kernel");
    return KSM_OK;
}

int ksm_ccg(const double *t1, long long n1, const double *t2, long long n2, int nbins,
            double tbin, uint32_t *counts, void *stream) {
    if (n1 < 0 || n2 < 0 || (n1 > 0 && !t1) || (n2 > 0 && !t2) || !counts || nbins <= 0 ||
        nbins > kMaxBins || !(tbin > 0.0))
        return ksm::fail(KSM_ERR_INVALID, "ksm_ccg: bad arguments (1 <= nbins <= %d)", kMaxBins);
    hipStream_t s = ksm::as_stream(stream);
    const int rc = zero_async(counts, (size_t)(2 * nbins + 1) * sizeof(uint32_t), s);
    if (rc != KSM_OK || n1 == 0 || n2 == 0) return rc;
    ccg_kernel<<<(unsigned)((n2 + 255) / 256), 256, 0, s>>>(t1, n1, t2, n2, nbins, tbin, counts);
    KSM_LAUNCH_CHECK("ccg_kernel");
    return KSM_OK;
}

int ksm_pc_map(float *X, const int64_t *rows, long long nrows, int R, int P, const float *M,
               void *stream) {
    if (nrows < 0 || R <= 0 || P <= 0 || P > 16 || !M || (nrows > 0 && (!X || !rows)))
        return ksm::fail(KSM_ERR_INVALID, "ksm_pc_map: bad arguments (P <= 16)");
    if (nrows == 0) return KSM_OK;
    const long long total = nrows * R;
    pc_map_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ksm::as_stream(stream)>>>(
        X, rows, nrows, R, P, M);
    KSM_LAUNCH_CHECK("pc_map_kernel");
    return KSM_OK;
}

namespace {
// Blocks of the counting sort for n spikes (each a contiguous range).
int ts_blocks(long long n) {
    return (int)std::max<long long>(1, std::min<long long>(256, (n + 8191) / 8192));
}
}  // namespace

size_t ksm_template_sums_workspace(long long n, int K) {
    if (n < 0 || K <= 0) return 0;
    const size_t NB = (size_t)ts_blocks(n);
    return ksm::align_up(NB * K * 4, 256) * 2 + ksm::align_up((size_t)K * 4, 256) +
           ksm::align_up(((size_t)K + 1) * 4, 256) + ksm::align_up((size_t)n * 4, 256) + 256;
}

int ksm_template_sums(const float *tF, const int64_t *tmpl, const int32_t *labels, long long n,
                      int nC, int P, int K, int n_groups, double *sums, uint64_t *counts,
                      uint8_t *presence, void *workspace, size_t ws_bytes, void *stream) {
    if (n < 0 || n >= (1LL << 31) || nC <= 0 || P <= 0 || K <= 0 || K > kTsMaxK || !sums ||
        !counts || (presence && (n_groups <= 0 || (double)n_groups * K > (double)(1LL << 30))) ||
        (n > 0 && (!tF || !tmpl || (presence && !labels))))
        return ksm::fail(KSM_ERR_INVALID, "ksm_template_sums: bad arguments (K <= %d)", kTsMaxK);
    if (!workspace || ws_bytes < ksm_template_sums_workspace(n, K))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_template_sums: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const int R = nC * P;
    const int NB = ts_blocks(n);
    const long long per = (n + NB - 1) / NB;
    ksm::Carver cv(workspace);
    int *H = cv.take<int>((size_t)NB * K);
    int *O = cv.take<int>((size_t)NB * K);
    int *tot = cv.take<int>(K);
    int *off = cv.take<int>((size_t)K + 1);
    int *list = cv.take<int>((size_t)n);
    if (presence) {
        const int rc = zero_async(presence, (size_t)n_groups * K, s);
        if (rc != KSM_OK) return rc;
    }
    const size_t lds = (size_t)K * sizeof(int);
    ts_hist_kernel<<<NB, kTsThreads, lds, s>>>(tmpl, n, per, K, H);
    KSM_LAUNCH_CHECK("ts_hist_kernel");
    ts_columns_kernel<<<(K + 3) / 4, 256, 0, s>>>(H, NB, K, O, tot);
    KSM_LAUNCH_CHECK("ts_columns_kernel");
    ts_scan_kernel<<<1, 1024, 0, s>>>(tot, K, off);
    KSM_LAUNCH_CHECK("ts_scan_kernel");
    ts_place_kernel<<<NB, kTsThreads, lds, s>>>(tmpl, n, per, K, off, O, list);
    KSM_LAUNCH_CHECK("ts_place_kernel");
    ts_reduce_kernel<<<(unsigned)K, kTsWaves * 64, 0, s>>>(
        tF, list, off, labels, R, K, sums, reinterpret_cast<unsigned long long *>(counts),
        presence);
    KSM_LAUNCH_CHECK("ts_reduce_kernel");
    return KSM_OK;
}

int ksm_pc_channel_rank(const double *sums, const uint64_t *counts, const uint8_t *presence,
                        const int64_t *iC, int K, int nC, int P, int C, int n_groups,
                        int32_t *sigma, int32_t *sigma_tmp, uint32_t *feature_ind,
                        void *stream) {
    const size_t lds = (size_t)C * P * sizeof(double) + (size_t)C * sizeof(double) +
                       (size_t)C * sizeof(int) + ((size_t)K * nC + K) * sizeof(short);
    if (!sums || !counts || !presence || !iC || !sigma || !sigma_tmp || !feature_ind || K <= 0 ||
        nC <= 0 || nC > 32 || P <= 0 || C <= 0 || C > kMaxRankC || K > 32767 ||
        n_groups <= 0 || lds > 65536 - 8192)
        return ksm::fail(KSM_ERR_INVALID,
                         "ksm_pc_channel_rank: bad arguments (nC <= 32, C <= 2048, "
                         "C (8 P + 12) + 2 K (nC + 1) <= 57344)");
    pc_rank_kernel<<<1, kRankThreads, lds, ksm::as_stream(stream)>>>(
        sums, reinterpret_cast<const unsigned long long *>(counts), presence, iC, K, nC, P, C,
        n_groups, sigma, sigma_tmp, feature_ind);
    KSM_LAUNCH_CHECK("pc_rank_kernel");
    return KSM_OK;
}

int ksm_gather_features(const float *tF, const int64_t *tmpl, long long n, int nC, int P,
                        const int32_t *sigma, float *out, int layout, void *stream) {
    if (n < 0 || nC <= 0 || P <= 0 || (layout != 0 && layout != 1) || !sigma ||
        (n > 0 && (!tF || !tmpl || !out)))
        return ksm::fail(KSM_ERR_INVALID, "ksm_gather_features: bad arguments");
    if (n > 0 && (const void *)out == (const void *)tF)
        return ksm::fail(KSM_ERR_INVALID, "ksm_gather_features: out must not alias tF");
    if (n == 0) return KSM_OK;
    const long long total = n * P;
    gather_features_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ksm::as_stream(stream)>>>(
        tF, tmpl, n, nC, P, sigma, out, layout);
    KSM_LAUNCH_CHECK("gather_features_kernel");
    return KSM_OK;
}

}  // extern "C"
