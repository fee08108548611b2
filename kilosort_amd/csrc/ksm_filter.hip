// Batch preprocessing: raw samples -> padded float batch -> channel map, sign,
// channel means, common median reference -> high-pass filter -> artifact
// test -> whitening.
//
// Reference: kilosort/io.py:789-831 (padded batch), io.py:990-1026 (filter),
// kilosort/preprocessing.py:139-154 (the filter in the Fourier domain).
//
// Kernels (one pass each over the batch):
//   chan_sums_kernel   channel sums of the converted samples (double partials
//                      per block of rows, added in a fixed order) -> means;
//   car_kernel         a tile of columns x all channels in LDS: subtract the
//                      means, select the lower median of each column (radix
//                      select on order-preserving keys, one wave per column),
//                      subtract it; written with a circular halo of J columns
//                      on each side for the filter;
//   mm_mfma_kernel     FP32 MFMA tiles (v_mfma_f32_32x32x2f32) for two
//                      products: the high-pass filter as a banded Toeplitz
//                      product (the taps come from LDS, the sliding window of
//                      each channel from the haloed batch), and the whitening
//                      product W @ Z.
// The Fourier-domain filter of the reference equals a circular correlation
// with the filter centred at T/2 (the conjugate spectrum correlates, fftshift
// rolls by T/2):  out[c, n] = sum_{j=-J..J} g[j] x[c, (n + j) mod T]; taps
// beyond J (where the impulse response has decayed below float32 resolution)
// are cut by the host.
#include "ksm_common.h"
#include "ksm_profile.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kSumRows = 256;      // rows (samples) per chan_sums block
constexpr int kMaxFilterC = 1536;  // channels (median keys: 24 per lane)
constexpr int kMaxJ = 4096;        // taps on each side of the filter's centre
constexpr int MB = 128, NB = 64, KB = 32;    // product tile (rows, columns, depth)
constexpr int ALD = KB + 1;                   // LDS row of the A tile (odd stride)
constexpr int BLD = NB + 4;                   // LDS row of the B tile

struct Src {
    const void *raw;
    int fmt;              // ksm_raw
    long long ldraw;      // row stride in elements (time-major: >= n_chan_bin)
    int nsamp, off;       // batch column off + r holds raw row r, r < nsamp
    int edge;             // KSM_FILTER_EDGE_FIRST / _LAST bits
    int flags;            // KSM_FILTER_* (invert, scale, shift)
    float scale, shift;
    const int32_t *chan_map;
};

// Converted sample of batch row c (mapped channel) at batch column t:
// io.py:797-831 (uint16 shift, scale, shift, edge replication or zeros),
// then io.py:993-997 (channel map, sign).
__device__ __forceinline__ float sample(const Src &s, int c, int t) {
    int r = t - s.off;
    if (r < 0) r = (s.edge & KSM_FILTER_EDGE_FIRST) ? 0 : -1;
    else if (r >= s.nsamp) r = (s.edge & KSM_FILTER_EDGE_LAST) ? s.nsamp - 1 : -1;
    if (r < 0) return 0.f;
    const int ch = s.chan_map ? s.chan_map[c] : c;
    float v;
    if (s.fmt == KSM_RAW_F32_CM) {
        v = static_cast<const float *>(s.raw)[(size_t)ch * s.ldraw + r];
    } else {
        const size_t i = (size_t)r * s.ldraw + ch;
        if (s.fmt == KSM_RAW_I16) v = (float)static_cast<const int16_t *>(s.raw)[i];
        else if (s.fmt == KSM_RAW_U16) v = (float)static_cast<const uint16_t *>(s.raw)[i] - 32768.f;
        else v = static_cast<const float *>(s.raw)[i];
    }
    if (s.flags & KSM_FILTER_SCALE) v = v * s.scale;
    if (s.flags & KSM_FILTER_SHIFT) v = v + s.shift;
    if (s.flags & KSM_FILTER_INVERT) v = -v;
    return v;
}

// psum[b][c] = sum of the samples of rows [b kSumRows, (b+1) kSumRows) of
// batch row c; block (b, channel group of 64).  Time-major sources: lanes
// along channels (coalesced rows), four row phases per block; channel-major:
// lanes along time, four channels per pass.  Each thread keeps 8 loads in
// flight; the block's phases meet in LDS and are added in a fixed order.
__global__ __launch_bounds__(256) void chan_sums_kernel(Src s, int C, int T,
                                                        double *__restrict__ psum) {
    __shared__ double s_p[4][64];
    const int b = blockIdx.x, t0 = b * kSumRows, c0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, ph = threadIdx.x >> 6;
    if (s.fmt != KSM_RAW_F32_CM) {
        const int c = c0 + lane;
        double acc = 0.0;
        if (c < C) {
            for (int i0 = 0; i0 < kSumRows / 4; i0 += 8) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int t = t0 + ph + 4 * (i0 + u);
                    v[u] = t < T ? sample(s, c, t) : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += (double)v[u];
            }
        }
        s_p[ph][lane] = acc;
    } else {
        // channel c0 + 4 i + ph, rows t0 + lane + 64 k
        for (int i = 0; i < 16; ++i) {
            const int c = c0 + 4 * i + ph;
            double acc = 0.0;
            if (c < C) {
                float v[kSumRows / 64];
#pragma unroll
                for (int k = 0; k < kSumRows / 64; ++k) {
                    const int t = t0 + lane + 64 * k;
                    v[k] = t < T ? sample(s, c, t) : 0.f;
                }
#pragma unroll
                for (int k = 0; k < kSumRows / 64; ++k) acc += (double)v[k];
            }
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (lane == 0) s_p[0][4 * i + ph] = acc;
        }
        __syncthreads();
        if (ph == 0) {
            const int c = c0 + lane;
            if (c < C) psum[(size_t)b * C + c] = s_p[0][lane];
        }
        return;
    }
    __syncthreads();
    if (ph == 0) {
        const int c = c0 + lane;
        if (c < C) psum[(size_t)b * C + c] = ((s_p[0][lane] + s_p[1][lane]) + s_p[2][lane]) +
                                             s_p[3][lane];
    }
}

// mean[c] = (sum over blocks of psum[b][c]) / T: one wave per channel, lanes
// over blocks, then a fixed-order tree over the lanes.
__global__ __launch_bounds__(256) void chan_means_kernel(const double *__restrict__ psum, int nb,
                                                         int C, int T, float *__restrict__ mean) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= C) return;
    double acc = 0.0;
    for (int b = lane; b < nb; b += 64) acc += psum[(size_t)b * C + c];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) mean[c] = (float)(acc / (double)T);
}

__device__ __forceinline__ unsigned order_key(float v) {
    const unsigned u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_value(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Lower medians (element (C-1)/2 of the sorted column, NaN if the column
// holds one) of columns j of s_y [C][ld], two columns per pass of a wave
// (independent dependency chains): the key of each is found bit by bit from
// the top as the smallest key with more than (C-1)/2 keys <= it.  NQ =
// ceil(C/64) keys per lane.
template <int NQ>
__device__ __forceinline__ void column_medians(const float *s_y, int ld, int C, int ncol,
                                               float *s_med, int wave, int lane) {
    const int k = (C - 1) / 2;
    for (int j = wave; j < ncol; j += 8) {        // columns j and j + 4 (ncol % 8 == 0)
        unsigned ka[NQ], kb[NQ];
        bool na = false, nb = false;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int c = q * 64 + lane;
            const bool ok = c < C;
            const float va = ok ? s_y[c * ld + j] : 0.f;
            const float vb = ok ? s_y[c * ld + j + 4] : 0.f;
            na |= va != va;
            nb |= vb != vb;
            ka[q] = ok ? order_key(va) : 0xffffffffu;
            kb[q] = ok ? order_key(vb) : 0xffffffffu;
        }
        unsigned aa = 0u, ab = 0u;
        for (int b = 31; b >= 0; --b) {
            const unsigned m = (1u << b) - 1u;
            const unsigned ca = aa | m, cb = ab | m;
            int na_le = 0, nb_le = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                na_le += __popcll(__ballot(ka[q] <= ca));
                nb_le += __popcll(__ballot(kb[q] <= cb));
            }
            if (na_le <= k) aa |= 1u << b;
            if (nb_le <= k) ab |= 1u << b;
        }
        const bool nana = __ballot(na) != 0ull, nanb = __ballot(nb) != 0ull;
        if (lane == 0) {
            s_med[j] = nana ? __builtin_nanf("") : key_value(aa);
            s_med[j + 4] = nanb ? __builtin_nanf("") : key_value(ab);
        }
    }
}

// NQ = ceil(C/64).  Columns [t0, t0 + ncol) of all C rows: y = x - mean[c]; with CAR, the
// lower median m of each column (torch.median: element (C-1)/2 of the sorted
// column; NaN if the column holds one) and out = y - m.  Written to
// dst[c * ldd + halo + t], and for t < halo or t >= T - halo also to the
// circular halo copies.  With `thr` finite, |out| >= thr raises *flag.
template <int NQ>
__global__ __launch_bounds__(256) void car_kernel(Src s, int C, int T, int ncol,
                                                  const float *__restrict__ mean, int car,
                                                  float *__restrict__ dst, int ldd, int halo,
                                                  float thr, int *__restrict__ flag) {
    extern __shared__ float s_y[];          // [C][ncol + 1]
    __shared__ float s_med[32];
    const int ld = ncol + 1;
    const int t0 = blockIdx.x * ncol;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // lanes along channels within a sample (time-major sources) or along time
    // within a row (channel-major), 8 loads in flight per thread
    const int cstep = 256 / ncol, jrow = tid % ncol, crow = tid / ncol;
    if (s.fmt != KSM_RAW_F32_CM) {
        const int nq = (C + 63) / 64;
        for (int i0 = 0; i0 < nq * (ncol / 4); i0 += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u, q = i / (ncol / 4), j = wave + 4 * (i % (ncol / 4));
                const int c = 64 * q + lane, t = t0 + j;
                v[u] = (i < nq * (ncol / 4) && c < C && t < T) ? sample(s, c, t) - mean[c] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u, q = i / (ncol / 4), j = wave + 4 * (i % (ncol / 4));
                const int c = 64 * q + lane;
                if (i < nq * (ncol / 4) && c < C) s_y[c * ld + j] = v[u];
            }
        }
    } else {
        for (int c0 = crow; c0 < C; c0 += 8 * cstep) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = c0 + u * cstep, t = t0 + jrow;
                v[u] = (c < C && t < T) ? sample(s, c, t) - mean[c] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = c0 + u * cstep;
                if (c < C) s_y[c * ld + jrow] = v[u];
            }
        }
    }
    __syncthreads();
    if (car) {
        column_medians<NQ>(s_y, ld, C, ncol, s_med, wave, lane);
    } else if (tid < ncol) {
        s_med[tid] = 0.f;
    }
    __syncthreads();
    bool big = false;
    for (int c = crow; c < C; c += cstep) {
        const int j = jrow, t = t0 + j;
        if (t >= T) continue;
        const float v = car ? s_y[c * ld + j] - s_med[j] : s_y[c * ld + j];
        big |= fabsf(v) >= thr;
        float *row = dst + (size_t)c * ldd + halo;
        row[t] = v;
        if (t < halo) row[T + t] = v;
        if (t >= T - halo) row[t - T] = v;
    }
    if (__ballot(big) != 0ull && lane == 0) atomicOr(flag, 1);
}

// Zero columns [from, ldd) of rows [0, C) and all of rows [C, Cp): what the
// filter product reads beyond the haloed batch (times 0 taps, or rows that
// are not stored, so they must be finite).
__global__ void zero_pad_kernel(float *__restrict__ y, int ldd, int from, int C, int Cp) {
    const int w = ldd - from;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n1 = (long long)C * w;
    if (e < n1) {
        const int c = (int)(e / w), j = (int)(e - (long long)c * w);
        y[(size_t)c * ldd + from + j] = 0.f;
    } else if (e - n1 < (long long)(Cp - C) * ldd) {
        y[(size_t)C * ldd + (e - n1)] = 0.f;
    }
}

__global__ void clear_flag_kernel(int *__restrict__ f) { *f = 0; }

// dst[r][c] = src[r][c] for r, c < n, else 0 (dst [rows][ld]).
__global__ void pad_matrix_kernel(const float *__restrict__ src, int n, float *__restrict__ dst,
                                  int rows, int ld) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * ld) return;
    const int r = e / ld, c = e - r * ld;
    dst[e] = (r < n && c < n) ? src[(size_t)r * n + c] : 0.f;
}

__global__ void zero_if_flag_kernel(float *__restrict__ X, int ldx, int C, int T,
                                    const int *__restrict__ flag) {
    if (*flag == 0) return;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)C * T) return;
    const int c = (int)(e / T), t = (int)(e - (long long)c * T);
    X[(size_t)c * ldx + t] = 0.f;
}

__device__ __forceinline__ int xcd_remap_f(int bid, int nb) {
    const int xcd = bid & 7, pos = bid >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// out[m][n] = sum_k A(m, k) B(k, n) on FP32 MFMA, block tile 128 x 64, four
// waves of 32 x 64 (1 x 2 tiles of 32x32x2).  Tiles of 64 columns keep the
// launch at about two rounds of resident blocks (383 x 60122: 2820 tiles),
// so that the last round is not a small tail.
//   TOEPLITZ: A(m, k) = a[m][n0 + k] (the sliding window of row m; a holds
//             the haloed batch, its rows zero padded to a multiple of 128),
//             B(k, n) = taps[k - (n - n0)] for
//             0 <= k - (n - n0) <= 2J (from LDS), depth 2J + NB;
//   else:     A(m, k) = a[m][k] (k < K), B(k, n) = b[k][n] (k < K); with
//             `veca`, a is zero padded to whole tiles (rows to a multiple of
//             128, columns to a multiple of 32, lda % 4 == 0) and read as
//             float4 without bounds tests; rows of b are read as float4 when
//             `vecb` (b 16-byte aligned, ldb % 4 == 0).
// Rows m < M and columns n < N are stored; with `flag` set (an artifact),
// zeros are stored instead.  With `thr` finite, |out| >= thr raises *flag
// (read by a later launch).
template <bool TOEPLITZ>
__global__ __launch_bounds__(256, 2) void mm_mfma_kernel(
    const float *__restrict__ a, int lda, const float *__restrict__ b, int ldb,
    const float *__restrict__ taps, int J, float *__restrict__ out, int ldo, int M, int N,
    int K, int nTm, int nTn, float thr, int *__restrict__ flag, const int *__restrict__ zflag,
    int veca, int vecb) {
    extern __shared__ float s_g[];               // Toeplitz taps, zero padded
    __shared__ float As[2][MB * ALD];
    __shared__ float Bs[TOEPLITZ ? 1 : 2][TOEPLITZ ? 1 : KB * BLD];
    const int logical = xcd_remap_f(blockIdx.x, nTm * nTn);
    const int tm = logical % nTm, tn = logical / nTm;
    const int m0 = tm * MB, n0 = tn * NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int depth = TOEPLITZ ? 2 * J + NB : K;
    const int nks = (depth + KB - 1) / KB;
    if (TOEPLITZ) {
        // s_g[u] = taps[u - (NB - 1)] for 0 <= u - (NB - 1) <= 2J, else 0
        const int ng = 2 * J + 2 * NB + KB;
        for (int u = tid; u < ng; u += 256) {
            const int q = u - (NB - 1);
            s_g[u] = (q >= 0 && q <= 2 * J) ? taps[q] : 0.f;
        }
    }
    // A tile: 128 rows x 32 depth; thread -> row tid/2, 16 consecutive columns
    const int ar = tid >> 1, ac = (tid & 1) * 16;
    const int arow = m0 + ar;
    const float *ap = a + (size_t)(arow < M || veca || TOEPLITZ ? arow : 0) * lda +
                      (TOEPLITZ ? n0 : 0) + ac;
    // B tile (general): 32 rows x 64 columns; thread -> rows tid/16 and +16
    const int br = tid >> 4, bc = (tid & 15) * 4;
    float4 ra0, ra1, ra2, ra3, rb0, rb1;
    // (always inlined: captured registers must not be given stack addresses)
    auto load = [&](int st) __attribute__((always_inline)) {
        const int k0 = st * KB;
        if constexpr (TOEPLITZ) {              // rows padded with zeros to whole tiles
            ra0 = *reinterpret_cast<const float4 *>(ap + k0);
            ra1 = *reinterpret_cast<const float4 *>(ap + k0 + 4);
            ra2 = *reinterpret_cast<const float4 *>(ap + k0 + 8);
            ra3 = *reinterpret_cast<const float4 *>(ap + k0 + 12);
        } else {
            if (veca) {                      // zero-padded, aligned A
                ra0 = *reinterpret_cast<const float4 *>(ap + k0);
                ra1 = *reinterpret_cast<const float4 *>(ap + k0 + 4);
                ra2 = *reinterpret_cast<const float4 *>(ap + k0 + 8);
                ra3 = *reinterpret_cast<const float4 *>(ap + k0 + 12);
            } else {
                float v[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int k = k0 + ac + e;
                    v[e] = (arow < M && k < K) ? ap[k0 + e] : 0.f;
                }
                ra0 = make_float4(v[0], v[1], v[2], v[3]);
                ra1 = make_float4(v[4], v[5], v[6], v[7]);
                ra2 = make_float4(v[8], v[9], v[10], v[11]);
                ra3 = make_float4(v[12], v[13], v[14], v[15]);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = k0 + br + 16 * h;
                const int n = n0 + bc;
                float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
                if (k < K) {
                    const float *bp = b + (size_t)k * ldb + n;
                    if (vecb && n + 3 < N) {
                        w = *reinterpret_cast<const float4 *>(bp);
                    } else {
                        w.x = n + 0 < N ? bp[0] : 0.f;
                        w.y = n + 1 < N ? bp[1] : 0.f;
                        w.z = n + 2 < N ? bp[2] : 0.f;
                        w.w = n + 3 < N ? bp[3] : 0.f;
                    }
                }
                if (h == 0) rb0 = w;
                else rb1 = w;
            }
        }
    };
    auto store = [&](int buf) __attribute__((always_inline)) {
        float *d = &As[buf][ar * ALD + ac];
        d[0] = ra0.x; d[1] = ra0.y; d[2] = ra0.z; d[3] = ra0.w;
        d[4] = ra1.x; d[5] = ra1.y; d[6] = ra1.z; d[7] = ra1.w;
        d[8] = ra2.x; d[9] = ra2.y; d[10] = ra2.z; d[11] = ra2.w;
        d[12] = ra3.x; d[13] = ra3.y; d[14] = ra3.z; d[15] = ra3.w;
        if constexpr (!TOEPLITZ) {
            *reinterpret_cast<float4 *>(&Bs[buf][br * BLD + bc]) = rb0;
            *reinterpret_cast<float4 *>(&Bs[buf][(br + 16) * BLD + bc]) = rb1;
        }
    };

    f32x16 acc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[j][r] = 0.f;

    load(0);
    store(0);
    __syncthreads();
    const int kr = lane >> 5, kc = lane & 31;
    for (int st = 0; st < nks; ++st) {
        const int buf = st & 1;
        const bool more = st + 1 < nks;
        if (more) load(st + 1);
#pragma unroll
        for (int kk = 0; kk < KB; kk += 2) {
            const float av = As[buf][(wave * 32 + kc) * ALD + kk + kr];
            float bv[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if constexpr (TOEPLITZ)
                    bv[j] = s_g[st * KB + kk + kr - (j * 32 + kc) + NB - 1];
                else
                    bv[j] = Bs[buf][(kk + kr) * BLD + j * 32 + kc];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv[j], acc[j], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
    }
    const bool zero = zflag != nullptr && *zflag != 0;
    bool big = false;
    // D layout of 32x32 MFMA: lane l, register r -> row 8(r/4) + 4(l/32) + r%4,
    // column l%32
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = m0 + wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int n = n0 + j * 32 + (lane & 31);
            if (m < M && n < N) {
                const float v = zero ? 0.f : acc[j][r];
                big |= fabsf(v) >= thr;
                out[(size_t)m * ldo + n] = v;
            }
        }
    if (flag != nullptr && __ballot(big) != 0ull && lane == 0) atomicOr(flag, 1);
}

struct FilterLayout {
    int nsb;              // chan_sums blocks
    int ldy;              // haloed batch row (filter input)
    int ldz;              // filter / median output row before whitening
};

FilterLayout filter_layout(int T, int J) {
    FilterLayout L;
    L.nsb = (T + kSumRows - 1) / kSumRows;
    const int depth = 2 * J + NB;
    L.ldy = (int)ksm::align_up(ksm::align_up(T, NB) + ksm::align_up(depth, KB), 4);
    L.ldz = (int)ksm::align_up(T, 4);
    return L;
}

int car_columns(int C) {
    for (int n = 32; n >= 8; n >>= 1)
        if ((size_t)C * (n + 1) * sizeof(float) <= 64 * 1024) return n;
    return 0;
}

int launch_mm(bool toeplitz, const float *a, int lda, const float *b, int ldb, const float *taps,
              int J, float *out, int ldo, int M, int N, int K, float thr, int *flag,
              const int *zflag, hipStream_t s, const char *name, int veca = 0) {
    const int nTm = (M + MB - 1) / MB, nTn = (N + NB - 1) / NB;
    ksm::ProfileScope prof(name, ksm::kProfProject, s);
    if (toeplitz) {
        const unsigned shm = (unsigned)((2 * J + 2 * NB + KB) * sizeof(float));
        prof.launch(mm_mfma_kernel<true>, dim3(nTm * nTn), dim3(256), shm, a, lda, b, ldb, taps,
                    J, out, ldo, M, N, K, nTm, nTn, thr, flag, zflag, 0, 0);
    } else {
        const int vecb = (ldb % 4) == 0 && (reinterpret_cast<uintptr_t>(b) & 15u) == 0;
        prof.launch(mm_mfma_kernel<false>, dim3(nTm * nTn), dim3(256), 0u, a, lda, b, ldb, taps,
                    J, out, ldo, M, N, K, nTm, nTn, thr, flag, zflag, veca, vecb);
    }
    KSM_LAUNCH_CHECK(name);
    return KSM_OK;
}

}  // namespace

extern "C" {

size_t ksm_filter_workspace(int C, int T, int J, int flags) {
    if (C <= 0 || T <= 0 || J < -1) return 0;
    const FilterLayout L = filter_layout(T, J < 0 ? 0 : J);
    size_t n = ksm::align_up((size_t)L.nsb * C * sizeof(double), 256) +
               ksm::align_up((size_t)C * sizeof(float), 256) + 256;
    if (J >= 0) n += ksm::align_up(ksm::align_up(C, MB) * L.ldy * sizeof(float), 256);
    if (flags & KSM_FILTER_WHITEN)
        n += ksm::align_up((size_t)C * L.ldz * sizeof(float), 256) +
             ksm::align_up(ksm::align_up(C, MB) * ksm::align_up(C, KB) * sizeof(float), 256);
    return n + 256;
}

int ksm_filter_batch(const void *raw, int fmt, int64_t ldraw, int nsamp, int off,
                     const int32_t *chan_map, int C, int T, int flags, float scale, float shift,
                     const float *taps, int J, float artifact_threshold, const float *whiten,
                     float *X, int ldx, void *workspace, size_t ws_bytes, void *stream) {
    if (!raw || !X || C <= 0 || T <= 0 || nsamp <= 0 || ldx < T || ldraw <= 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_filter_batch: bad arguments");
    if (fmt < KSM_RAW_I16 || fmt > KSM_RAW_F32_CM)
        return ksm::fail(KSM_ERR_INVALID, "ksm_filter_batch: unknown raw format %d", fmt);
    if (C > kMaxFilterC)
        return ksm::fail(KSM_ERR_INVALID, "ksm_filter_batch: C=%d exceeds %d", C, kMaxFilterC);
    const bool hp = taps != nullptr;
    if (hp && (J < 0 || J > kMaxJ || 2 * J + 1 > T))
        return ksm::fail(KSM_ERR_INVALID, "ksm_filter_batch: J=%d (need 0 <= J <= %d, 2J+1 <= T)",
                         J, kMaxJ);
    if (!hp) J = -1;
    const bool skip = (flags & KSM_FILTER_SKIP) != 0;
    const bool wh = whiten != nullptr && !skip;
    if (wh != ((flags & KSM_FILTER_WHITEN) != 0) && !skip)
        return ksm::fail(KSM_ERR_INVALID,
                         "ksm_filter_batch: KSM_FILTER_WHITEN must be set exactly when whiten "
                         "is given");
    if (!workspace || ws_bytes < ksm_filter_workspace(C, T, J, flags))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_filter_batch: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    const FilterLayout L = filter_layout(T, hp ? J : 0);
    ksm::Carver cv(workspace);
    double *psum = cv.take<double>((size_t)L.nsb * C);
    float *mean = cv.take<float>(C);
    int *flag = cv.take<int>(1);
    const int Cp = (int)ksm::align_up(C, MB);
    float *Y = hp ? cv.take<float>((size_t)Cp * L.ldy) : nullptr;
    float *Z = (flags & KSM_FILTER_WHITEN) ? cv.take<float>((size_t)C * L.ldz) : nullptr;
    const int wrows = (int)ksm::align_up(C, MB), wld = (int)ksm::align_up(C, KB);
    float *Wp = (flags & KSM_FILTER_WHITEN) ? cv.take<float>((size_t)wrows * wld) : nullptr;

    Src src;
    src.raw = raw;
    src.fmt = fmt;
    src.ldraw = ldraw;
    src.nsamp = nsamp;
    src.off = off;
    src.edge = flags & (KSM_FILTER_EDGE_FIRST | KSM_FILTER_EDGE_LAST);
    src.flags = flags;
    src.scale = scale;
    src.shift = shift;
    src.chan_map = chan_map;

    const bool check = !skip && artifact_threshold < INFINITY;
    const float thr = check ? artifact_threshold : INFINITY;
    clear_flag_kernel<<<1, 1, 0, s>>>(flag);
    KSM_LAUNCH_CHECK("clear_flag_kernel");
    chan_sums_kernel<<<dim3(L.nsb, (C + 63) / 64), 256, 0, s>>>(src, C, T, psum);
    KSM_LAUNCH_CHECK("chan_sums_kernel");
    chan_means_kernel<<<(C + 3) / 4, 256, 0, s>>>(psum, L.nsb, C, T, mean);
    KSM_LAUNCH_CHECK("chan_means_kernel");

    // destination of the median stage
    float *d1;
    int ld1, halo1;
    if (skip) {
        d1 = X, ld1 = ldx, halo1 = 0;
    } else if (hp) {
        d1 = Y, ld1 = L.ldy, halo1 = J;
    } else if (wh) {
        d1 = Z, ld1 = L.ldz, halo1 = 0;
    } else {
        d1 = X, ld1 = ldx, halo1 = 0;
    }
    const int ncol = car_columns(C);
    {
        ksm::ProfileScope prof("filter_car", ksm::kProfProject, s);
        const unsigned shm = (unsigned)((size_t)C * (ncol + 1) * sizeof(float));
        decltype(&car_kernel<1>) kern = nullptr;
        switch ((C + 63) / 64) {
#define KSM_CAR_CASE(n) \
    case n: kern = car_kernel<n>; break;
            KSM_CAR_CASE(1) KSM_CAR_CASE(2) KSM_CAR_CASE(3) KSM_CAR_CASE(4)
            KSM_CAR_CASE(5) KSM_CAR_CASE(6) KSM_CAR_CASE(7) KSM_CAR_CASE(8)
            KSM_CAR_CASE(9) KSM_CAR_CASE(10) KSM_CAR_CASE(11) KSM_CAR_CASE(12)
            KSM_CAR_CASE(13) KSM_CAR_CASE(14) KSM_CAR_CASE(15) KSM_CAR_CASE(16)
            KSM_CAR_CASE(17) KSM_CAR_CASE(18) KSM_CAR_CASE(19) KSM_CAR_CASE(20)
            KSM_CAR_CASE(21) KSM_CAR_CASE(22) KSM_CAR_CASE(23) KSM_CAR_CASE(24)
#undef KSM_CAR_CASE
            default: return ksm::fail(KSM_ERR_INVALID, "ksm_filter_batch: C=%d", C);
        }
        prof.launch(kern, dim3((T + ncol - 1) / ncol), dim3(256), shm, src, C, T, ncol,
                    static_cast<const float *>(mean), (flags & KSM_FILTER_CAR) ? 1 : 0, d1, ld1,
                    halo1, hp ? INFINITY : thr, flag);
    }
    KSM_LAUNCH_CHECK("car_kernel");
    if (skip) return KSM_OK;
    if (hp) {
        const int from = T + 2 * J;          // < ldy (ldy >= T + 2J + 64)
        const long long n = (long long)C * (L.ldy - from) + (long long)(Cp - C) * L.ldy;
        zero_pad_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(Y, L.ldy, from, C, Cp);
        KSM_LAUNCH_CHECK("zero_pad_kernel");
        float *d2 = wh ? Z : X;
        const int ld2 = wh ? L.ldz : ldx;
        const int rc = launch_mm(true, Y, L.ldy, nullptr, 0, taps, J, d2, ld2, C, T, 0, thr,
                                 check ? flag : nullptr, nullptr, s, "filter_highpass");
        if (rc != KSM_OK) return rc;
    }
    if (wh) {
        pad_matrix_kernel<<<(wrows * wld + 255) / 256, 256, 0, s>>>(whiten, C, Wp, wrows, wld);
        KSM_LAUNCH_CHECK("pad_matrix_kernel");
        const int rc = launch_mm(false, Wp, wld, Z, L.ldz, nullptr, 0, X, ldx, C, T, C, INFINITY,
                                 nullptr, check ? flag : nullptr, s, "filter_whiten", 1);
        if (rc != KSM_OK) return rc;
    } else if (check) {
        const long long n = (long long)C * T;
        zero_if_flag_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(X, ldx, C, T, flag);
        KSM_LAUNCH_CHECK("zero_if_flag_kernel");
    }
    return KSM_OK;
}

int ksm_gemm_f32(const float *A, int lda, const float *B, int ldb, float *Cm, int ldc, int M,
                 int N, int K, void *stream) {
    if (!A || !B || !Cm || M <= 0 || N <= 0 || K <= 0 || lda < K || ldb < N || ldc < N)
        return ksm::fail(KSM_ERR_INVALID, "ksm_gemm_f32: bad arguments");
    return launch_mm(false, A, lda, B, ldb, nullptr, 0, Cm, ldc, M, N, K, INFINITY, nullptr,
                     nullptr, ksm::as_stream(stream), "gemm_f32_rowmajor");
}

}  // extern "C"
