// Universal-template spike detection: Kilosort4's first spike-finding pass.
//
// Reference: kilosort/spikedetect.py:133-178 (template_match), :190-197
// (yweighted) and :256-281 (the per-batch features of run).
//
// Per padded batch X [C][T]:
//   B[k][c][t] = sum_m wTEMP[k, m] X[c, t + m - nt/2]          (conv_pc, FP32)
//   A[i, k, f, t] = sum_j weigh[i, j, f] B[k][iC[j, f]][t]       (score_kernel)
//   As[f, t] = max_{q = i nk + k} |A|, first q on ties;
//   Im[f, t] = (1 + q) sign(A[q]) -- stored by the score pass for other
//     shapes than the reference's defaults; at the defaults the score pass
//     keeps only As and the detected spikes' q and sign are recomputed
//     (emit_one), which spares the pass its index bookkeeping
//   Amaxs[f, t] = max over the nC2 nearest templates f' = iC2[n, f] and over
//     |dt| <= nt0 of As[f', t + dt], where columns t + dt outside
//     [nt, T - nt) count as 0 (the reference zeroes them before its
//     max_pool1d; the maxima are exact, so their order does not matter)
//   spikes: As == Amaxs and As > Th, in (f, t) order (torch.nonzero).
// Kernels: score_kernel (per template group and column tile, the group's
// channel range of B staged in LDS), window_max_kernel (R[f, t] = the time-window maximum of row f),
// mask_kernel (per (f, t): the maximum of R over the neighbours, stopping at
// the first neighbour above As; one 64-bit mask word per 64 columns),
// row_count_kernel + row_scan_kernel (offsets per row), emit_kernel (one wave per
// template row: the spikes in time order with amp, imax, adist, the
// channel-weighted y position and the PC features of the residual window).
#include "ksm_common.h"
#include "ksm_profile.h"
#include "ksm_project_internal.h"

namespace {

constexpr int kScoreG = 32;       // templates per score block
constexpr int kScoreT = 32;       // columns per score block
constexpr int kScoreLds = 24576;  // bytes of LDS for the staged channel range
constexpr int kMaxW = 64;         // nsizes * nC
constexpr int kMaxSizes = 8;      // template sizes (nsizes)
constexpr int kMaxNk = 8;         // universal waveforms (nk)
constexpr int kMaxNC = 16;        // nearest channels (nC)
constexpr int kWinT = 1024;       // columns per window-max block
constexpr int kMaxNt0 = 256;

// As[f][t], Im[f][t] for a group of kScoreG consecutive templates and
// kScoreT columns.  Consecutive template positions are neighbours on the
// probe, so the group's channels lie in a short index range [cmin, cmax]:
// B[k][cmin..cmax][tile] is staged once in LDS and every template of the
// group reads its nC channels from there (a group whose range exceeds
// max_span channels reads B from global memory instead).  Thread: one column, the
// group's templates slot, slot + 8, ...  Sizes passed at run time; the
// reference's defaults take score_tiles_kernel below.
__global__ __launch_bounds__(256) void score_kernel(
    const float *__restrict__ B, int C, int ldc, const int32_t *__restrict__ iC, int nC_,
    const float *__restrict__ weigh, int nsizes_, int nk, int Nfilt, int T, int max_span,
    float *__restrict__ As, int16_t *__restrict__ Im) {
    const int nC = nC_;
    const int nsizes = nsizes_;
    constexpr int MC = kMaxNC;
    constexpr int MS = kMaxSizes;
    extern __shared__ float s_b[];                  // [nk][span][kScoreT], span <= max_span
    __shared__ float s_w[kScoreG][kMaxW];
    __shared__ int s_c[kScoreG][kMaxNC];
    __shared__ int s_lo, s_hi;
    // blocks of neighbouring groups (sharing channels) are launched together
    const int f0 = blockIdx.x * kScoreG, t0 = blockIdx.y * kScoreT;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_lo = 0x7fffffff;
        s_hi = -1;
    }
    __syncthreads();
    int lo = 0x7fffffff, hi = -1;
    for (int e = tid; e < kScoreG * nC; e += 256) {
        const int g = e / nC, j = e - g * nC;
        const int f = f0 + g;
        const int c = f < Nfilt ? iC[(size_t)j * Nfilt + f] : -1;
        s_c[g][j] = c;
        if (c >= 0) {
            lo = min(lo, c);
            hi = max(hi, c);
        }
    }
    for (int e = tid; e < kScoreG * nsizes * nC; e += 256) {
        const int g = e / (nsizes * nC), r = e - g * nsizes * nC;
        const int i = r / nC, j = r - i * nC;
        const int f = f0 + g;
        s_w[g][i * nC + j] = f < Nfilt ? weigh[((size_t)i * nC + j) * Nfilt + f] : 0.f;
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if ((tid & 63) == 0) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
    }
    __syncthreads();
    const int cmin = s_lo, span = s_hi - s_lo + 1;
    const bool staged = span <= max_span;
    if (staged) {                                   // thread: column tid % 32, rows tid / 32 + 8 n
        const int uu = tid & (kScoreT - 1), r0 = tid / kScoreT;
        const int t = t0 + uu;
        for (int k = 0; k < nk; ++k)
            for (int c = r0; c < span; c += 256 / kScoreT)
                s_b[(k * span + c) * kScoreT + uu] =
                    t < ldc ? B[((size_t)k * C + cmin + c) * ldc + t] : 0.f;
    }
    __syncthreads();
    const int u = tid & (kScoreT - 1), slot = tid / kScoreT;
    const int t = t0 + u;
    if (t >= T) return;
    for (int g = slot; g < kScoreG; g += 256 / kScoreT) {
        const int f = f0 + g;
        if (f >= Nfilt) break;
        float best = -1.f, bval = 0.f;
        int bq = 0;
        for (int k = 0; k < nk; ++k) {
            float b[MC];
            // the two sources in separate loops: a select between an LDS and a
            // global address would become generic (flat) loads
            if (staged) {
#pragma unroll
                for (int j = 0; j < MC; ++j)
                    if (j < nC) {
                        b[j] = s_b[(k * span + s_c[g][j] - cmin) * kScoreT + u];
                    }
            } else {
#pragma unroll
                for (int j = 0; j < MC; ++j)
                    if (j < nC) {
                        b[j] = B[((size_t)k * C + s_c[g][j]) * ldc + t];
                    }
            }
#pragma unroll
            for (int i = 0; i < MS; ++i) {
                if (i >= nsizes) break;
                float a = 0.f;
#pragma unroll
                for (int j = 0; j < MC; ++j)
                    if (j < nC) a = __builtin_fmaf(s_w[g][i * nC + j], b[j], a);
                const float m = fabsf(a);
                const int q = i * nk + k;
                // the first q on ties (torch.max); NaN wins and stays
                if (m > best || (m == best && q < bq) || (m != m && best == best)) {
                    best = m;
                    bq = q;
                    bval = a;
                }
            }
        }
        const int sg = bval > 0.f ? 1 : (bval < 0.f ? -1 : 0);
        As[(size_t)f * ldc + t] = best;
        Im[(size_t)f * ldc + t] = (int16_t)((1 + bq) * sg);
    }
}

// The score pass at compile-time sizes (the reference's defaults): as
// score_kernel, but a block keeps its group of kScoreG positions over a run
// of `tiles` column tiles of kScoreT.  The group's weights and channel table
// are set up once; its distinct channels (at most kTileU, found through a
// channel bitmap, so that a group where the position order wraps across the
// probe stages too) are staged per tile in LDS, and the next tile's rows are
// loaded into registers while the current tile is scored.  Each sum takes
// the channels in order as a chain of FMAs from 0 (channels outer, NS x NK
// sums); only the maximum |A| is kept (NaN skipped, as a strictly larger |A|
// would skip it), the index and sign of its first occurrence being
// recomputed for the detected spikes alone (emit_one).
// Compute units of the current device (cached per host thread and device).
int device_cus() {
    static thread_local int c_dev = -1, c_cus = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev != c_dev || c_cus <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        c_dev = dev;
        c_cus = cus;
    }
    return c_cus;
}

constexpr int kTileU = 32;        // distinct channels a staged group may use
constexpr int kTileBits = 4096;   // channel bitmap: C <= kTileBits
template <int NS, int NCH, int NK>
__global__ __launch_bounds__(256) void score_tiles_kernel(
    const float *__restrict__ B, int C, int ldc, const int32_t *__restrict__ iC,
    const float *__restrict__ weigh, int Nfilt, int T, int tiles, float *__restrict__ As) {
    constexpr int kPf = (NK * kTileU * kScoreT + 255) / 256;   // prefetch registers
    __shared__ float s_b[NK * kTileU * kScoreT];
    __shared__ float s_w[kScoreG][NS * NCH];
    __shared__ int s_slot[kScoreG][NCH];   // staged: slot of (g, j); else its channel
    __shared__ int s_ch[kTileU];
    __shared__ unsigned s_bits[kTileBits / 32];
    __shared__ int s_pre[kTileBits / 32];
    __shared__ int s_nu;
    const int tid = threadIdx.x, lane = tid & 63;
    const int f0 = blockIdx.x * kScoreG;
    for (int e = tid; e < kTileBits / 32; e += 256) s_bits[e] = 0u;
    __syncthreads();
    for (int e = tid; e < kScoreG * NCH; e += 256) {
        const int g = e / NCH, j = e - g * NCH, f = f0 + g;
        if (f < Nfilt) {
            const int c = iC[(size_t)j * Nfilt + f];
            atomicOr(&s_bits[c >> 5], 1u << (c & 31));
        }
    }
    for (int e = tid; e < kScoreG * NS * NCH; e += 256) {
        const int g = e / (NS * NCH), r = e - g * (NS * NCH);
        const int i = r / NCH, j = r - i * NCH, f = f0 + g;
        s_w[g][r] = f < Nfilt ? weigh[((size_t)i * NCH + j) * Nfilt + f] : 0.f;
    }
    __syncthreads();
    if (tid < 64) {                          // slots: prefix of the words' popcounts
        int run = 0;
        for (int w0 = 0; w0 < kTileBits / 32; w0 += 64) {
            const int pc = __popc(s_bits[w0 + lane]);
            int incl = pc;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if (lane >= o) incl += v;
            }
            s_pre[w0 + lane] = run + incl - pc;
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) s_nu = run;
    }
    __syncthreads();
    const int nu = s_nu;
    const bool staged = nu <= kTileU;        // block-uniform
    for (int e = tid; e < kScoreG * NCH; e += 256) {
        const int g = e / NCH, j = e - g * NCH, f = f0 + g;
        int sl = -1;
        if (f < Nfilt) {
            const int c = iC[(size_t)j * Nfilt + f];
            const int slot = s_pre[c >> 5] + __popc(s_bits[c >> 5] & ((1u << (c & 31)) - 1u));
            sl = staged ? slot : c;
            if (staged) s_ch[slot] = c;      // every writer of a slot stores the same c
        }
        s_slot[g][j] = sl;
    }
    __syncthreads();
    const int u = tid & (kScoreT - 1), slot_t = tid / kScoreT;
    const int ntiles = (T + kScoreT - 1) / kScoreT;
    const int tile0 = blockIdx.y * tiles, tile1 = min(ntiles, tile0 + tiles);
    // staging map: thread tid holds column tid % 32 of rows r = tid / 32 + 8 m
    // (m < kPf), row r = k kTileU + slot, so k = m / 4 and slot = 8 (m % 4) +
    // tid / 32; rows with slot >= nu are not loaded
    static_assert(kTileU == 32 && kScoreT == 32 && kPf == NK * 4, "staging map");
    const int uu = tid & 31, r0 = tid >> 5;
    int chq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) chq[q] = (staged && 8 * q + r0 < nu) ? s_ch[8 * q + r0] : -1;
    float pf[kPf];
    auto load = [&](int tile) {
        const int t = tile * kScoreT + uu;
#pragma unroll
        for (int m = 0; m < kPf; ++m) {
            const int c = chq[m & 3];
            pf[m] = (c >= 0 && t < ldc) ? B[((size_t)(m >> 2) * C + c) * ldc + t] : 0.f;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int m = 0; m < kPf; ++m)
            if (chq[m & 3] >= 0) s_b[((m >> 2) * kTileU + 8 * (m & 3) + r0) * kScoreT + uu] = pf[m];
    };
    if (staged && tile0 < tile1) {
        load(tile0);
        store();
    }
    __syncthreads();
    for (int tile = tile0; tile < tile1; ++tile) {
        const bool more = staged && tile + 1 < tile1;     // block-uniform
        if (more) load(tile + 1);
        const int t = tile * kScoreT + u;
        if (t < T) {
#pragma unroll 1
            for (int g = slot_t; g < kScoreG; g += 256 / kScoreT) {
                const int f = f0 + g;
                if (f >= Nfilt) break;
                float acc[NS][NK];
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int k = 0; k < NK; ++k) acc[i][k] = 0.f;
                auto fold = [&](int j, const float (&bj)[NK]) {
                    float wj[NS];
#pragma unroll
                    for (int i = 0; i < NS; ++i) wj[i] = s_w[g][i * NCH + j];
#pragma unroll
                    for (int i = 0; i < NS; ++i)
#pragma unroll
                        for (int k = 0; k < NK; ++k)
                            acc[i][k] = __builtin_fmaf(wj[i], bj[k], acc[i][k]);
                };
                // two channels' loads in flight at a time: unrolling all NCH
                // hoists every load and costs occupancy
                if (staged) {
#pragma unroll 2
                    for (int j = 0; j < NCH; ++j) {
                        const int sl = s_slot[g][j];
                        float bj[NK];
#pragma unroll
                        for (int k = 0; k < NK; ++k) bj[k] = s_b[(k * kTileU + sl) * kScoreT + u];
                        fold(j, bj);
                    }
                } else {
#pragma unroll 2
                    for (int j = 0; j < NCH; ++j) {
                        const int c = s_slot[g][j];
                        float bj[NK];
#pragma unroll
                        for (int k = 0; k < NK; ++k) bj[k] = B[((size_t)k * C + c) * ldc + t];
                        fold(j, bj);
                    }
                }
                float best = -1.f;
#pragma unroll
                for (int i = 0; i < NS; ++i)
#pragma unroll
                    for (int k = 0; k < NK; ++k) best = fmaxf(best, fabsf(acc[i][k]));
                As[(size_t)f * ldc + t] = best;
            }
        }
        __syncthreads();                     // the tile's rows are read
        if (more) store();
        __syncthreads();
    }
}

// R[f][t] = max of As[f][s] over s in [t - nt0, t + nt0] and [nt, T - nt),
// 0 if that range is empty (the reference's zeroed columns).  Block: row f,
// kWinT columns (four per thread), the window staged in LDS.
__global__ __launch_bounds__(256) void window_max_kernel(const float *__restrict__ As, int ldc,
                                                         int T, int nt, int nt0,
                                                         float *__restrict__ R) {
    __shared__ float s_a[kWinT + 2 * kMaxNt0];
    const int f = blockIdx.y, t0 = blockIdx.x * kWinT;
    const float *row = As + (size_t)f * ldc;
    const int w = kWinT + 2 * nt0;
    for (int e = threadIdx.x; e < w; e += 256) {
        const int s = t0 - nt0 + e;
        s_a[e] = (s >= nt && s < T - nt) ? row[s] : 0.f;
    }
    __syncthreads();
    for (int r = 0; r < kWinT / 256; ++r) {
        const int u = threadIdx.x + 256 * r, t = t0 + u;
        if (t >= T) return;
        float m = 0.f;
        for (int d = 0; d <= 2 * nt0; ++d) m = fmaxf(m, s_a[u + d]);
        R[(size_t)f * ldc + t] = m;
    }
}

// mask[f][w] bit b: column t = 64 w + b is a spike of template f.  Block:
// row f, 1024 columns; each wave writes four mask words.
__global__ __launch_bounds__(256) void mask_kernel(const float *__restrict__ As,
                                                   const float *__restrict__ R, int ldc, int T,
                                                   const int32_t *__restrict__ iC2, int nC2,
                                                   int Nfilt, float th,
                                                   unsigned long long *__restrict__ mask,
                                                   int nw) {
    const int f = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the wave's four score loads are issued together
    float av[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int t = 64 * (blockIdx.x * 16 + wave * 4 + r) + lane;
        av[r] = t < T ? As[(size_t)f * ldc + t] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int w = blockIdx.x * 16 + wave * 4 + r;       // mask word
        const int t = 64 * w + lane;
        bool hit = false;
        if (t < T) {
            const float a = av[r];
            if (a > th) {
                float m = -1.f;
                for (int n = 0; n < nC2; ++n) {
                    const float v = R[(size_t)iC2[(size_t)n * Nfilt + f] * ldc + t];
                    m = fmaxf(m, v);
                    if (m > a) break;        // already not the maximum
                }
                hit = m == a;
            }
        }
        const unsigned long long bits = __ballot(hit);
        if (lane == 0 && w < nw) mask[(size_t)f * nw + w] = bits;
    }
}

// cnt[f] = the spikes of row f (one wave per row).
__global__ __launch_bounds__(256) void row_count_kernel(const unsigned long long *__restrict__ mask,
                                                        int nw, int Nfilt, int *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= Nfilt) return;
    int c = 0;
    for (int w = lane; w < nw; w += 64) c += __popcll(mask[(size_t)f * nw + w]);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) cnt[f] = c;
}

// off[f] = exclusive prefix of cnt over rows; *count = the total.
__global__ __launch_bounds__(1024) void row_scan_kernel(const int *__restrict__ cnt, int Nfilt,
                                                        int *__restrict__ off,
                                                        int *__restrict__ count) {
    __shared__ int s_part[1024];
    const int tid = threadIdx.x;
    const int per = (Nfilt + 1023) / 1024;    // a contiguous run of rows per thread
    const int r0 = tid * per, r1 = min(Nfilt, r0 + per);
    int acc = 0;
    for (int f = r0; f < r1; ++f) acc += cnt[f];
    s_part[tid] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {     // inclusive scan over threads
        const int v = tid >= o ? s_part[tid - o] : 0;
        __syncthreads();
        s_part[tid] += v;
        __syncthreads();
    }
    int run = tid > 0 ? s_part[tid - 1] : 0;
    for (int f = r0; f < r1; ++f) {
        off[f] = run;
        run += cnt[f];
    }
    if (tid == 1023) *count = s_part[1023];
}

// The outputs of spike s = (f, t) except its position, written by the whole
// wave (lane j < nC: adist and y; lane j * P + p: feature p of channel j).
// The index q of the first maximum |A| and the sign of that A come from Im,
// or, where the score pass kept only the maximum (Im null), are recomputed
// here: lane q = i nk + k forms A[i, k] with the score pass's operands in
// its order (so bit for bit the same value), NaN never chosen, and the
// lanes' first maximum (largest |A|, then smallest q) is the sequential
// scan's "strictly larger |A|" choice.
__device__ __forceinline__ void emit_one(
    int f, int t, int s, const float *__restrict__ As, const int16_t *__restrict__ Im,
    const float *__restrict__ weigh, int nsizes,
    const float *__restrict__ B, int C, int ldc, const float *__restrict__ X, int ldx, int T,
    const int32_t *__restrict__ iC, int nC, int Nfilt, int nk, const float *__restrict__ yc,
    const float *__restrict__ wPCA, int P, int nt, int h, int cap, int32_t *__restrict__ imax,
    float *__restrict__ amp, float *__restrict__ adist, float *__restrict__ yct,
    float *__restrict__ tF, int lane) {
    int q;
    float sg;
    if (Im) {
        const int im = Im[(size_t)f * ldc + t];
        q = (im < 0 ? -im : im) - 1;
        sg = im > 0 ? 1.f : (im < 0 ? -1.f : 0.f);
    } else {
        float a = 0.f, key = -2.f;
        int ql = 64;
        if (lane < nsizes * nk) {
            const int i = lane / nk, kk = lane - i * nk;
            for (int j = 0; j < nC; ++j) {
                const int c = iC[(size_t)j * Nfilt + f];
                a = __builtin_fmaf(weigh[((size_t)i * nC + j) * Nfilt + f],
                                   B[((size_t)kk * C + c) * ldc + t], a);
            }
            key = a != a ? -1.f : fabsf(a);
            ql = lane;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float ok = __shfl_xor(key, o, 64);
            const int oq = __shfl_xor(ql, o, 64);
            if (ok > key || (ok == key && oq < ql)) {
                key = ok;
                ql = oq;
            }
        }
        q = ql;
        const float bv = __shfl(a, q, 64);
        sg = bv > 0.f ? 1.f : (bv < 0.f ? -1.f : 0.f);
    }
    const int k = q % nk;
    float ad = 0.f, yj = 0.f;
    if (lane < nC) {
        const int c = iC[(size_t)lane * Nfilt + f];
        ad = B[((size_t)k * C + c) * ldc + t] * sg;
        yj = yc[c];
        adist[(size_t)lane * cap + s] = ad;
    }
    const float rl = lane < nC ? (ad != ad ? ad : fmaxf(ad, 0.f)) : 0.f;   // relu keeps NaN
    float S = 0.f;                                  // sum of relu(adist) in channel order
    for (int j = 0; j < nC; ++j) S += __shfl(rl, j, 64);
    float y = 0.f;
    for (int j = 0; j < nC; ++j) y += (__shfl(rl, j, 64) / S) * __shfl(yj, j, 64);
    if (lane < nC * P) {
        const int j = lane / P, p = lane - j * P;
        const int c = iC[(size_t)j * Nfilt + f];
        const float *xr = X + (size_t)c * ldx;
        float acc = 0.f;
        for (int m = 0; m < nt; ++m) {
            const int u = t + m - h;                // inside [0, T) for nt0 <= nt - nt/2 - 1
            const float x = (u >= 0 && u < T) ? xr[u] : 0.f;
            acc = __builtin_fmaf(x, wPCA[p * nt + m], acc);
        }
        tF[((size_t)s * nC + j) * P + p] = acc;
    }
    if (lane == 0) {
        imax[s] = q;
        amp[s] = As[(size_t)f * ldc + t];
        yct[s] = y;
    }
}

// One wave per spike s < min(*count, cap) (grid-stride): its outputs from
// the position xy[s] written by emit_kernel.
__global__ __launch_bounds__(256) void spike_features_kernel(
    const int *__restrict__ count, const float *__restrict__ As, const int16_t *__restrict__ Im,
    const float *__restrict__ weigh, int nsizes, const float *__restrict__ B, int C, int ldc, const float *__restrict__ X, int ldx, int T,
    const int32_t *__restrict__ iC, int nC, int Nfilt, int nk, const float *__restrict__ yc,
    const float *__restrict__ wPCA, int P, int nt, int cap, const int32_t *__restrict__ xy,
    int32_t *__restrict__ imax, float *__restrict__ amp, float *__restrict__ adist,
    float *__restrict__ yct, float *__restrict__ tF) {
    const int lane = threadIdx.x & 63;
    const int n = min(*count, cap);
    for (int s = blockIdx.x * 4 + (threadIdx.x >> 6); s < n; s += gridDim.x * 4) {
        const int f = xy[2 * (size_t)s], t = xy[2 * (size_t)s + 1];
        emit_one(f, t, s, As, Im, weigh, nsizes, B, C, ldc, X, ldx, T, iC, nC, Nfilt, nk, yc, wPCA,
                 P, nt, nt / 2, cap, imax, amp, adist, yct, tF, lane);
    }
}

// One wave per template row f: the positions of its spikes, in time order,
// at off[f] ... (the mask words of the row are scanned 64 at a time).  The
// other outputs (spike_features_kernel):
// Outputs (index s < cap): xy[2 s] = f, xy[2 s + 1] = t; imax[s] = q;
// amp[s] = As; adist[j * cap + s] = B[q % nk][iC[j, f]][t] * sign;
// yct[s] = sum_j (relu(adist_j) / sum relu(adist)) * yc[iC[j, f]];
// tF[(s * nC + j) * P + p] = sum_m X[iC[j, f], t + m - nt/2] wPCA[p, m].
__global__ __launch_bounds__(256) void emit_kernel(
    const unsigned long long *__restrict__ mask, int nw, const int *__restrict__ off,
    const float *__restrict__ As, const int16_t *__restrict__ Im, const float *__restrict__ B,
    int C, int ldc, const float *__restrict__ X, int ldx, int T,
    const int32_t *__restrict__ iC, int nC, int Nfilt, int nk, const float *__restrict__ yc,
    const float *__restrict__ wPCA, int P, int nt, int cap, int32_t *__restrict__ xy,
    int32_t *__restrict__ imax, float *__restrict__ amp, float *__restrict__ adist,
    float *__restrict__ yct, float *__restrict__ tF) {
    const int lane = threadIdx.x & 63;
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= Nfilt) return;
    const int h = nt / 2;
    int base = off[f];
    for (int w0 = 0; w0 < nw; w0 += 64) {
        // lanes load 64 mask words; exclusive prefix of their popcounts
        const int wl = w0 + lane;
        const unsigned long long myw = wl < nw ? mask[(size_t)f * nw + wl] : 0ull;
        const int pc = __popcll(myw);
        int incl = pc;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        const int excl = incl - pc;
        const int total = __shfl(incl, 63, 64);
        unsigned long long busy = __ballot(myw != 0ull);
        while (busy) {
            const int src = __builtin_ctzll(busy);
            busy &= busy - 1ull;
            unsigned long long bits = __shfl(myw, src, 64);
            int s = base + __shfl(excl, src, 64);
            while (bits) {
                const int b = __builtin_ctzll(bits);
                bits &= bits - 1ull;
                const int t = 64 * (w0 + src) + b;
                if (s < cap && lane == 0) {
                    xy[2 * (size_t)s] = f;
                    xy[2 * (size_t)s + 1] = t;
                }
                ++s;
            }
        }
        base += total;
    }
}

// ---------------------------------------------------------------------------
// Isolated single-channel peaks and their snippets (extract_snippets,
// spikedetect.py:34-51): peak[c][t] = |X| equals the maximum of |X| over
// channels c +- lc and samples t +- lt and exceeds th (max_pool2d); isolated
// = a peak with no other peak over c +- ic, t +- it (avg_pool2d x window
// size == 1, computed as the reference's float32 division and product) and
// nt <= t < T - nt.

constexpr int kSnT = 64;          // columns per peak block
constexpr int kSnC = 64;          // channels per peak block
constexpr int kMaxLc = 16, kMaxLt = 16;

__global__ __launch_bounds__(256) void snippet_peak_kernel(const float *__restrict__ X, int ldx,
                                                           int C, int T, int lc, int lt,
                                                           float th,
                                                           uint8_t *__restrict__ peak) {
    __shared__ float s_a[kSnC + 2 * kMaxLc][kSnT + 2 * kMaxLt + 1];
    __shared__ float s_r[kSnC + 2 * kMaxLc][kSnT + 1];
    const int c0 = blockIdx.y * kSnC, t0 = blockIdx.x * kSnT;
    const int rows = kSnC + 2 * lc, cols = kSnT + 2 * lt;
    for (int e = threadIdx.x; e < rows * cols; e += 256) {
        const int r = e / cols, j = e - r * cols;
        const int c = c0 - lc + r, t = t0 - lt + j;
        s_a[r][j] = (c >= 0 && c < C && t >= 0 && t < T) ? fabsf(X[(size_t)c * ldx + t])
                                                         : -INFINITY;
    }
    __syncthreads();
    // maximum over t +- lt per row, then over c +- lc (separable)
    for (int e = threadIdx.x; e < rows * kSnT; e += 256) {
        const int r = e / kSnT, j = e - r * kSnT;
        float m = -INFINITY;
        for (int d = 0; d <= 2 * lt; ++d) m = fmaxf(m, s_a[r][j + d]);
        s_r[r][j] = m;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kSnC * kSnT; e += 256) {
        const int r = e / kSnT, j = e - r * kSnT;
        const int c = c0 + r, t = t0 + j;
        if (c >= C || t >= T) continue;
        float m = -INFINITY;
        for (int d = 0; d <= 2 * lc; ++d) m = fmaxf(m, s_r[r + d][j]);
        const float a = s_a[r + lc][j + lt];
        peak[(size_t)c * T + t] = (a == m && a > th) ? 1 : 0;
    }
}

// iso[c][w] bit b: t = 64 w + b is an isolated peak of channel c.
__global__ __launch_bounds__(256) void snippet_iso_kernel(const uint8_t *__restrict__ peak,
                                                          int C, int T, int ic, int it, int nt,
                                                          unsigned long long *__restrict__ iso,
                                                          int nw) {
    const int c = blockIdx.y;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool hit = false;
    if (t < T && peak[(size_t)c * T + t]) {
        int cnt = 0;
        for (int dc = -ic; dc <= ic; ++dc) {
            const int cc = c + dc;
            if (cc < 0 || cc >= C) continue;
            const uint8_t *row = peak + (size_t)cc * T;
            for (int dt = -it; dt <= it; ++dt) {
                const int tt = t + dt;
                if (tt >= 0 && tt < T) cnt += row[tt];
            }
        }
        const float size = (float)((2 * ic + 1) * (2 * it + 1));
        const float v = ((float)cnt / size) * size;
        hit = v == 1.f && t >= nt && t < T - nt;
    }
    const unsigned long long bits = __ballot(hit);
    const int w = (blockIdx.x * 256 + (threadIdx.x & ~63)) >> 6;
    if (lane == 0 && w < nw) iso[(size_t)c * nw + w] = bits;
}

// One wave per channel: snippets X[c][t - twav_min + k], k < nt, of its
// isolated peaks in time order at off[c] ...
__global__ __launch_bounds__(256) void snippet_emit_kernel(
    const unsigned long long *__restrict__ iso, int nw, const int *__restrict__ off,
    const float *__restrict__ X, int ldx, int C, int nt, int twav_min, int cap,
    float *__restrict__ clips, int32_t *__restrict__ xy) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= C) return;
    int s = off[c];
    for (int w = 0; w < nw; ++w) {
        unsigned long long bits = iso[(size_t)c * nw + w];
        while (bits) {
            const int b = __builtin_ctzll(bits);
            bits &= bits - 1ull;
            const int t = 64 * w + b;
            if (s < cap) {
                for (int k = lane; k < nt; k += 64)
                    clips[(size_t)s * nt + k] = X[(size_t)c * ldx + t - twav_min + k];
                if (lane == 0) {
                    xy[2 * (size_t)s] = c;
                    xy[2 * (size_t)s + 1] = t;
                }
            }
            ++s;
        }
    }
}

__global__ void clear_int_kernel(int *__restrict__ p) { *p = 0; }

struct DetectLayout {
    int ldc, nw;
};

DetectLayout detect_layout(int T) {
    DetectLayout L;
    L.ldc = (int)ksm::align_up(T, 256);
    L.nw = (T + 63) / 64;
    return L;
}

}  // namespace

extern "C" {

size_t ksm_detect_workspace(int C, int T, int nk, int Nfilt) {
    if (C <= 0 || T <= 0 || nk <= 0 || Nfilt <= 0) return 0;
    const DetectLayout L = detect_layout(T);
    return ksm::align_up((size_t)nk * C * L.ldc * sizeof(float), 256) +
           2 * ksm::align_up((size_t)Nfilt * L.ldc * sizeof(float), 256) +
           ksm::align_up((size_t)Nfilt * L.ldc * sizeof(int16_t), 256) +
           ksm::align_up((size_t)Nfilt * L.nw * sizeof(unsigned long long), 256) +
           2 * ksm::align_up((size_t)Nfilt * sizeof(int), 256) + 512;
}

int ksm_detect(const float *X, int ldx, int C, int T, const float *wTEMP, int nk, int nt,
               const int32_t *iC, int nC, const int32_t *iC2, int nC2, const float *weigh,
               int nsizes, int Nfilt, int nt0, float th, const float *yc, const float *wPCA,
               int P, int cap, int32_t *xy, int32_t *imax, float *amp, float *adist, float *yct,
               float *tF, int *count, void *workspace, size_t ws_bytes, void *stream) {
    if (!X || !wTEMP || !iC || !iC2 || !weigh || !yc || !wPCA || !count || C <= 0 || T <= 0 ||
        ldx < T || nk <= 0 || nt <= 0 || (nt % 2) == 0 || Nfilt <= 0 || nC2 <= 0 || P <= 0 ||
        cap < 0 || (cap > 0 && (!xy || !imax || !amp || !adist || !yct || !tF)))
        return ksm::fail(KSM_ERR_INVALID, "ksm_detect: bad arguments");
    if (nC <= 0 || nC > kMaxNC || nsizes <= 0 || nsizes > kMaxSizes || nk > kMaxNk ||
        nsizes * nC > kMaxW ||
        nC * P > 64 || nt0 < 0 || nt0 > kMaxNt0 || 2 * nt > T)
        return ksm::fail(KSM_ERR_INVALID,
                         "ksm_detect: limits nC <= %d, nsizes <= %d, nsizes*nC <= %d, nk <= %d, "
                         "nC*P <= 64, nt0 <= %d, 2 nt <= T",
                         kMaxNC, kMaxSizes, kMaxW, kMaxNk, kMaxNt0);
    if (!workspace || ws_bytes < ksm_detect_workspace(C, T, nk, Nfilt))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_detect: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const DetectLayout L = detect_layout(T);
    ksm::Carver cv(workspace);
    float *B = cv.take<float>((size_t)nk * C * L.ldc);
    float *As = cv.take<float>((size_t)Nfilt * L.ldc);
    float *R = cv.take<float>((size_t)Nfilt * L.ldc);
    int16_t *Im = cv.take<int16_t>((size_t)Nfilt * L.ldc);
    unsigned long long *mask = cv.take<unsigned long long>((size_t)Nfilt * L.nw);
    int *off = cv.take<int>(Nfilt);
    int *cnt = cv.take<int>(Nfilt);
    int *nonfinite = cv.take<int>(1);

    // the reference's defaults take the score pass that keeps only max |A|
    const bool fast = nsizes == 5 && nC == 10 && nk == 6 && C <= kTileBits;
    clear_int_kernel<<<1, 1, 0, s>>>(nonfinite);
    KSM_LAUNCH_CHECK("clear_int_kernel");
    int rc = ksm::conv_pc(X, ldx, wTEMP, C, T, nk, nt, B, L.ldc, nonfinite, s);
    if (rc != KSM_OK) return rc;
    {
        ksm::ProfileScope prof("detect_score", ksm::kProfProject, s);
        const int max_span = kScoreLds / (nk * kScoreT * (int)sizeof(float));
        const unsigned shm = (unsigned)(nk * max_span * kScoreT * sizeof(float));
        const dim3 grid((Nfilt + kScoreG - 1) / kScoreG, (T + kScoreT - 1) / kScoreT);
        if (fast) {                                     // the reference's defaults
            // as many blocks as are resident at once, each over a run of tiles
            const int gx = (Nfilt + kScoreG - 1) / kScoreG;
            const int ntiles = (T + kScoreT - 1) / kScoreT;
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, score_tiles_kernel<5, 10, 6>, 256, 0) != hipSuccess || per_cu <= 0)
                per_cu = 2;
            const int gy = std::max(1, std::min(ntiles, per_cu * device_cus() / gx));
            const int tiles = (ntiles + gy - 1) / gy;
            prof.launch(score_tiles_kernel<5, 10, 6>, dim3(gx, (ntiles + tiles - 1) / tiles),
                        dim3(256), 0u, static_cast<const float *>(B), C, L.ldc, iC, weigh, Nfilt,
                        T, tiles, As);
        } else
            prof.launch(score_kernel, grid, dim3(256), shm,
                        static_cast<const float *>(B), C, L.ldc, iC, nC, weigh, nsizes, nk, Nfilt,
                        T, max_span, As, Im);
    }
    KSM_LAUNCH_CHECK("score_kernel");
    window_max_kernel<<<dim3((T + kWinT - 1) / kWinT, Nfilt), 256, 0, s>>>(As, L.ldc, T, nt, nt0,
                                                                         R);
    KSM_LAUNCH_CHECK("window_max_kernel");
    {
        ksm::ProfileScope prof("detect_mask", ksm::kProfProject, s);
        prof.launch(mask_kernel, dim3((T + 1023) / 1024, Nfilt), dim3(256), 0u,
                    static_cast<const float *>(As), static_cast<const float *>(R), L.ldc, T, iC2,
                    nC2, Nfilt, th, mask, L.nw);
    }
    KSM_LAUNCH_CHECK("mask_kernel");
    row_count_kernel<<<(Nfilt + 3) / 4, 256, 0, s>>>(mask, L.nw, Nfilt, cnt);
    KSM_LAUNCH_CHECK("row_count_kernel");
    row_scan_kernel<<<1, 1024, 0, s>>>(cnt, Nfilt, off, count);
    KSM_LAUNCH_CHECK("row_scan_kernel");
    {
        ksm::ProfileScope prof("detect_emit", ksm::kProfProject, s);
        prof.launch(emit_kernel, dim3((Nfilt + 3) / 4), dim3(256), 0u,
                    static_cast<const unsigned long long *>(mask), L.nw,
                    static_cast<const int *>(off), static_cast<const float *>(As),
                    static_cast<const int16_t *>(Im), static_cast<const float *>(B), C, L.ldc, X,
                    ldx, T, iC, nC, Nfilt, nk, yc, wPCA, P, nt, cap, xy, imax, amp, adist, yct,
                    tF);
    }
    KSM_LAUNCH_CHECK("emit_kernel");
    if (cap > 0) {
        ksm::ProfileScope prof("detect_features", ksm::kProfProject, s);
        const int nblk = (min(cap, 16384) + 3) / 4;
        prof.launch(spike_features_kernel, dim3(nblk), dim3(256), 0u,
                    static_cast<const int *>(count), static_cast<const float *>(As),
                    static_cast<const int16_t *>(fast ? nullptr : Im), weigh, nsizes,
                    static_cast<const float *>(B), C, L.ldc, X,
                    ldx, T, iC, nC, Nfilt, nk, yc, wPCA, P, nt, cap,
                    static_cast<const int32_t *>(xy), imax, amp, adist, yct, tF);
        KSM_LAUNCH_CHECK("spike_features_kernel");
    }
    return KSM_OK;
}

size_t ksm_snippets_workspace(int C, int T) {
    if (C <= 0 || T <= 0) return 0;
    const int nw = (T + 63) / 64;
    return ksm::align_up((size_t)C * T, 256) +
           ksm::align_up((size_t)C * nw * sizeof(unsigned long long), 256) +
           2 * ksm::align_up((size_t)C * sizeof(int), 256) + 256;
}

int ksm_snippets(const float *X, int ldx, int C, int T, int nt, int twav_min, float th, int lc,
                 int lt, int ic, int it, int cap, float *clips, int32_t *xy, int *count,
                 void *workspace, size_t ws_bytes, void *stream) {
    if (!X || !count || C <= 0 || T <= 0 || ldx < T || nt <= 0 || twav_min < 0 ||
        twav_min > nt || cap < 0 || (cap > 0 && (!clips || !xy)) || lc < 0 || lt < 0 ||
        lc > kMaxLc || lt > kMaxLt || ic < 0 || it < 0 || 2 * nt > T)
        return ksm::fail(KSM_ERR_INVALID, "ksm_snippets: bad arguments (lc <= %d, lt <= %d)",
                         kMaxLc, kMaxLt);
    if (!workspace || ws_bytes < ksm_snippets_workspace(C, T))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_snippets: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const int nw = (T + 63) / 64;
    ksm::Carver cv(workspace);
    uint8_t *peak = cv.take<uint8_t>((size_t)C * T);
    unsigned long long *iso = cv.take<unsigned long long>((size_t)C * nw);
    int *off = cv.take<int>(C);
    int *cnt = cv.take<int>(C);
    snippet_peak_kernel<<<dim3((T + kSnT - 1) / kSnT, (C + kSnC - 1) / kSnC), 256, 0, s>>>(
        X, ldx, C, T, lc, lt, th, peak);
    KSM_LAUNCH_CHECK("snippet_peak_kernel");
    snippet_iso_kernel<<<dim3((T + 255) / 256, C), 256, 0, s>>>(peak, C, T, ic, it, nt, iso, nw);
    KSM_LAUNCH_CHECK("snippet_iso_kernel");
    row_count_kernel<<<(C + 3) / 4, 256, 0, s>>>(iso, nw, C, cnt);
    KSM_LAUNCH_CHECK("row_count_kernel");
    row_scan_kernel<<<1, 1024, 0, s>>>(cnt, C, off, count);
    KSM_LAUNCH_CHECK("row_scan_kernel");
    snippet_emit_kernel<<<(C + 3) / 4, 256, 0, s>>>(iso, nw, off, X, ldx, C, nt, twav_min, cap,
                                                    clips, xy);
    KSM_LAUNCH_CHECK("snippet_emit_kernel");
    return KSM_OK;
}

}  // extern "C"
