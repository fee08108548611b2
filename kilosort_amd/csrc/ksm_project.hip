// Template scores B = U * (W (*) X): the temporal-PC correlation of every
// channel (VALU, register-blocked) followed by the channel projection as a
// GEMM on the matrix cores.
//
// Reference: kilosort/template_matching.py:180-181 (run_matching):
//   B = conv1d(X[:, None], W[:, None], padding=nt//2)     -> Bc [C, P, T]
//   B = einsum('ijk, kjl -> il', U, B)                     -> B  [K, T]
// Here Bc is stored [P*C (+pad)][ldc] with row r = p*C + c, so that U viewed
// as [K][P*C] is the left GEMM operand without reshuffling.
#include "ksm_common.h"
#include "ksm_profile.h"
#include "ksm_project_internal.h"
#include "ksm_graph.h"

namespace {

constexpr int kConvThreads = 256;
constexpr int kConvTT = 8;                          // outputs per thread
constexpr int kConvSpan = kConvThreads * kConvTT;   // 2048 columns per block

// Skewed LDS index: one pad word every 8 words so that lanes reading 8-word
// strided windows hit distinct banks.
__host__ __device__ constexpr int skew(int e) { return e + (e >> 3); }

// Bc[(p*C + c)*ldc + t] = sum_k W[p,k] * X[c, t + k - NTAP/2], t < ldc.
// Columns t >= T are written as 0.
template <int P, int NTAP>
__global__ __launch_bounds__(kConvThreads) void conv_pc_kernel(
    const float *__restrict__ X, int ldx, const float *__restrict__ W, int C, int T,
    int ldc, float *__restrict__ Bc, int *__restrict__ nonfinite) {
    constexpr int H = NTAP / 2;
    constexpr int NWIN = kConvTT + NTAP - 1;
    constexpr int NLOAD = kConvSpan + NTAP - 1;
    __shared__ float xs[skew(NLOAD) + 8];
    const int c = blockIdx.y;
    const int t0 = blockIdx.x * kConvSpan;
    const float *xr = X + (size_t)c * ldx;
    bool bad = false;
    for (int e = threadIdx.x; e < NLOAD; e += kConvThreads) {
        const int t = t0 - H + e;
        const float v = (t >= 0 && t < T) ? xr[t] : 0.f;
        bad |= !__builtin_isfinite(v);
        xs[skew(e)] = v;
    }
    if (__builtin_amdgcn_ballot_w64(bad) != 0ull && (threadIdx.x & 63) == 0)
        atomicOr(nonfinite, 1);
    __syncthreads();
    const int base = threadIdx.x * kConvTT;
    float win[NWIN];
#pragma unroll
    for (int e = 0; e < NWIN; ++e) win[e] = xs[skew(base + e)];
    float acc[kConvTT][P];
#pragma unroll
    for (int u = 0; u < kConvTT; ++u)
#pragma unroll
        for (int p = 0; p < P; ++p) acc[u][p] = 0.f;
#pragma unroll
    for (int k = 0; k < NTAP; ++k) {
        float w[P];
#pragma unroll
        for (int p = 0; p < P; ++p) w[p] = W[p * NTAP + k];
#pragma unroll
        for (int u = 0; u < kConvTT; ++u)
#pragma unroll
            for (int p = 0; p < P; ++p) acc[u][p] = __builtin_fmaf(win[u + k], w[p], acc[u][p]);
    }
    const int tb = t0 + base;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float *dst = Bc + ((size_t)p * C + c) * ldc + tb;
        if (tb + kConvTT <= ldc) {
            float4 v0, v1;
            v0.x = tb + 0 < T ? acc[0][p] : 0.f;
            v0.y = tb + 1 < T ? acc[1][p] : 0.f;
            v0.z = tb + 2 < T ? acc[2][p] : 0.f;
            v0.w = tb + 3 < T ? acc[3][p] : 0.f;
            v1.x = tb + 4 < T ? acc[4][p] : 0.f;
            v1.y = tb + 5 < T ? acc[5][p] : 0.f;
            v1.z = tb + 6 < T ? acc[6][p] : 0.f;
            v1.w = tb + 7 < T ? acc[7][p] : 0.f;
            reinterpret_cast<float4 *>(dst)[0] = v0;
            reinterpret_cast<float4 *>(dst)[1] = v1;
        } else {
#pragma unroll
            for (int u = 0; u < kConvTT; ++u)
                if (tb + u < ldc) dst[u] = tb + u < T ? acc[u][p] : 0.f;
        }
    }
}

// Generic fallback: one output per thread, runtime P and nt.
__global__ __launch_bounds__(256) void conv_pc_generic(const float *__restrict__ X, int ldx,
                                                       const float *__restrict__ W, int P,
                                                       int nt, int C, int T, int ldc,
                                                       float *__restrict__ Bc,
                                                       int *__restrict__ nonfinite) {
    const int c = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ldc) return;
    if (t < T && !__builtin_isfinite(X[(size_t)c * ldx + t])) atomicOr(nonfinite, 1);
    const int H = nt / 2;
    const float *xr = X + (size_t)c * ldx;
    for (int p = 0; p < P; ++p) {
        float acc = 0.f;
        if (t < T)
            for (int k = 0; k < nt; ++k) {
                const int s = t + k - H;
                const float x = (s >= 0 && s < T) ? xr[s] : 0.f;
                acc = __builtin_fmaf(x, W[p * nt + k], acc);
            }
        Bc[((size_t)p * C + c) * ldc + t] = acc;
    }
}

// At[r][Kp] = U[perm[k]][r] (k < K), 0 for K <= k < Kp or r >= R
// (perm == nullptr: identity).
__global__ void transpose_pad_kernel(const float *__restrict__ U, const int32_t *__restrict__ perm,
                                     int K, int R, int Kp, int Rp, float *__restrict__ At) {
    const size_t n = (size_t)Rp * Kp;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % Kp);
        const int r = (int)(e / Kp);
        const int src = (k < K && perm) ? perm[k] : k;
        At[e] = (k < K && r < R) ? U[(size_t)src * R + r] : 0.f;
    }
}

__global__ void zero_rows_kernel(float *__restrict__ p, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        p[e] = 0.f;
}

__global__ void set_flag_kernel(int *__restrict__ p) { *p = 0; }

// The contraction steps (GBK rows of the reduction index r = p*C + c) that a
// template tile needs: steps where the tile's 64 columns of At are not all
// zero.  Templates are zero outside their channel footprint, and the tiles of
// templates sorted by peak channel touch a band of channels, so most steps
// are skipped; a skipped step would add only exact zeros.  With a non-finite
// value in X, W or U every step is kept (0 * inf must give NaN as in the
// dense product).  steps[tm][0 .. counts[tm]) in increasing order.
constexpr int kStepsMax = 4096;        // reduction steps per tile handled sparsely
__global__ __launch_bounds__(256) void tile_steps_kernel(
    const float *__restrict__ At, int Kp, int nk, const float *__restrict__ W, int PW,
    const int *__restrict__ nonfinite, int *__restrict__ steps, int *__restrict__ counts) {
    __shared__ unsigned char s_need[kStepsMax];
    __shared__ int s_dense;
    const int tm = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_dense = *nonfinite != 0 || nk > kStepsMax;
    __syncthreads();
    bool wbad = false;
    for (int e = tid; e < PW; e += 256) wbad |= !__builtin_isfinite(W[e]);
    if (wbad) s_dense = 1;                 // benign race: any writer stores 1
    __syncthreads();
    const bool dense = s_dense != 0;
    if (!dense) {
        const float *a = At + (size_t)tm * 64 + lane;
        for (int st = wave; st < nk; st += 4) {
            bool nz = false;
            for (int i = 0; i < 16; ++i) {
                const float v = a[((size_t)st * 16 + i) * Kp];
                nz |= v != 0.f;                    // NaN counts as nonzero
            }
            const bool any = __builtin_amdgcn_ballot_w64(nz) != 0ull;
            if (lane == 0) s_need[st] = any;
        }
    }
    __syncthreads();
    if (wave == 0) {
        int cnt = 0;
        for (int c0 = 0; c0 < nk; c0 += 64) {
            const int st = c0 + lane;
            const bool f = st < nk && (dense || s_need[st]);
            const unsigned long long m = __builtin_amdgcn_ballot_w64(f);
            const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (f) steps[(size_t)tm * nk + pos] = st;
            cnt += __popcll(m);
        }
        if (lane == 0) counts[tm] = cnt;
    }
}

// ---------------------------------------------------------------------------
// FP32 GEMM on v_mfma_f32_32x32x2f32: Bout[m][n] = sum_r At[r][m] * Bc[r][n].
// Block tile 64 (templates) x 256 (time), depth step 16, 4 waves each owning
// 64x64 as 2x2 MFMA tiles (64-row template tiles pad 700 templates to 704,
// not 768).  Operand tiles are double buffered in LDS with the next step's
// global loads in flight during the MFMAs.

constexpr int GBM = 64, GBN = 256, GBK = 16;
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Maps the hardware block id to a logical tile so that each XCD (blocks are
// dispatched round-robin over 8 XCDs) works on a contiguous run of tiles;
// tiles are ordered template-tile fastest, so the blocks that share one Bc
// column panel run on the same XCD and share its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int xcd = bid & 7, pos = bid >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

__global__ __launch_bounds__(256, 2) void gemm_f32_mfma(
    const float *__restrict__ At, int Kp, const float *__restrict__ Bc, int ldc, int nkt,
    const int *__restrict__ steps, const int *__restrict__ counts, float *__restrict__ Bout,
    int ldb, int K, int nmax, int nTm, int nTn) {
    __shared__ float As[2][GBK][GBM];
    __shared__ float Bs[2][GBK][GBN];
    const int logical = xcd_remap(blockIdx.x, nTm * nTn);
    const int tm = logical % nTm, tn = logical / nTm;
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wn = wave;
    // global->LDS mapping: A 16 rows x 16 float4 (one per thread); B 16 rows x
    // 64 float4 (four per thread, rows lb + 4 i)
    const int la = tid >> 4, lca = (tid & 15) * 4;
    const int lb = tid >> 6, lcb = (tid & 63) * 4;

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    // the tile's reduction steps (rows GBK*step ..), in increasing order
    const int nk = counts[tm];
    const int *lst = steps + (size_t)tm * nkt;
    const float *ga0 = At + (size_t)la * Kp + m0 + lca;
    const float *gb0 = Bc + (size_t)lb * ldc + n0 + lcb;
    const size_t sb = (size_t)4 * ldc;
    const int st0 = nk > 0 ? lst[0] : 0;
    const float *ga = ga0 + (size_t)st0 * GBK * Kp;
    const float *gb = gb0 + (size_t)st0 * GBK * ldc;
    float4 ra = *reinterpret_cast<const float4 *>(ga);
    float4 rb0 = *reinterpret_cast<const float4 *>(gb);
    float4 rb1 = *reinterpret_cast<const float4 *>(gb + sb);
    float4 rb2 = *reinterpret_cast<const float4 *>(gb + 2 * sb);
    float4 rb3 = *reinterpret_cast<const float4 *>(gb + 3 * sb);
    *reinterpret_cast<float4 *>(&As[0][la][lca]) = ra;
    *reinterpret_cast<float4 *>(&Bs[0][lb][lcb]) = rb0;
    *reinterpret_cast<float4 *>(&Bs[0][lb + 4][lcb]) = rb1;
    *reinterpret_cast<float4 *>(&Bs[0][lb + 8][lcb]) = rb2;
    *reinterpret_cast<float4 *>(&Bs[0][lb + 12][lcb]) = rb3;
    __syncthreads();
    const int kr = lane >> 5, kc = lane & 31;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            const int st = lst[kt + 1];
            ga = ga0 + (size_t)st * GBK * Kp;
            gb = gb0 + (size_t)st * GBK * ldc;
            ra = *reinterpret_cast<const float4 *>(ga);
            rb0 = *reinterpret_cast<const float4 *>(gb);
            rb1 = *reinterpret_cast<const float4 *>(gb + sb);
            rb2 = *reinterpret_cast<const float4 *>(gb + 2 * sb);
            rb3 = *reinterpret_cast<const float4 *>(gb + 3 * sb);
        }
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 2) {
            const float a0 = As[buf][kk + kr][kc];
            const float a1 = As[buf][kk + kr][32 + kc];
            const float b0 = Bs[buf][kk + kr][wn * 64 + kc];
            const float b1 = Bs[buf][kk + kr][wn * 64 + 32 + kc];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (more) {
            const int nb = buf ^ 1;
            *reinterpret_cast<float4 *>(&As[nb][la][lca]) = ra;
            *reinterpret_cast<float4 *>(&Bs[nb][lb][lcb]) = rb0;
            *reinterpret_cast<float4 *>(&Bs[nb][lb + 4][lcb]) = rb1;
            *reinterpret_cast<float4 *>(&Bs[nb][lb + 8][lcb]) = rb2;
            *reinterpret_cast<float4 *>(&Bs[nb][lb + 12][lcb]) = rb3;
        }
        __syncthreads();
    }
    // D layout of 32x32 MFMA: lane l, register r -> row 8*(r/4) + 4*(l/32) + r%4,
    // column l%32.  The scores are stored transposed, Bt[n][m]: registers
    // 4g..4g+3 of a lane are four consecutive templates of one time column.
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int m = m0 + i * 32 + 8 * g + 4 * (lane >> 5);
                const int n = n0 + wn * 64 + j * 32 + (lane & 31);
                if (n >= nmax) continue;
                float *dst = Bout + (size_t)n * ldb + m;
                if (m + 3 < K) {
                    float4 v;
                    v.x = acc[i][j][4 * g + 0];
                    v.y = acc[i][j][4 * g + 1];
                    v.z = acc[i][j][4 * g + 2];
                    v.w = acc[i][j][4 * g + 3];
                    *reinterpret_cast<float4 *>(dst) = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (m + e < K) dst[e] = acc[i][j][4 * g + e];
                }
            }
}

struct ProjectLayout {
    int R, Rp, Kp, ldc, nTm, nk;
    size_t bc_bytes, at_bytes;
};

ProjectLayout project_layout(int C, int T, int K, int P) {
    ProjectLayout L;
    L.R = P * C;
    L.Rp = (int)ksm::align_up(L.R, GBK);
    L.Kp = (int)ksm::align_up(K, GBM);
    L.ldc = (int)ksm::align_up(T, GBN);
    L.bc_bytes = (size_t)L.Rp * L.ldc * sizeof(float);
    L.at_bytes = (size_t)L.Rp * L.Kp * sizeof(float);
    L.nTm = L.Kp / GBM;
    L.nk = L.Rp / GBK;
    return L;
}


}  // namespace

namespace ksm {

int conv_pc(const float *X, int ldx, const float *W, int C, int T, int P, int nt, float *Bc,
            int ldc, int *nonfinite, hipStream_t s) {
    if (P == 6 && nt == 61) {
        ProfileScope prof("conv_pc", kProfProject, s);
        dim3 grid((ldc + kConvSpan - 1) / kConvSpan, C);
        prof.launch(conv_pc_kernel<6, 61>, grid, dim3(kConvThreads), 0u, X, ldx, W, C, T, ldc, Bc,
                    nonfinite);
    } else {
        ProfileScope prof("conv_pc", kProfProject, s);
        dim3 grid((ldc + 255) / 256, C);
        prof.launch(conv_pc_generic, grid, dim3(256), 0u, X, ldx, W, P, nt, C, T, ldc, Bc,
                    nonfinite);
    }
    KSM_LAUNCH_CHECK("conv_pc_kernel");
    return KSM_OK;
}

}  // namespace ksm

extern "C" {

size_t ksm_project_workspace(int C, int T, int K, int P, int nt, int precision) {
    (void)nt;
    if (C <= 0 || T <= 0 || K <= 0 || P <= 0) return 0;
    if (precision == KSM_PREC_BF16 || precision == KSM_PREC_BF16X3)
        return ksm::project_bf16_workspace(C, T, K, P, precision);
    if (precision == KSM_PREC_FP8) return ksm::project_fp8_workspace(C, T, K, P);
    const ProjectLayout L = project_layout(C, T, K, P);
    return ksm::align_up(L.bc_bytes, 256) + ksm::align_up(L.at_bytes, 256) +
           ksm::align_up((size_t)L.nTm * L.nk * sizeof(int), 256) +
           ksm::align_up((size_t)L.nTm * sizeof(int), 256) + 256 + 512;
}

int ksm_project(const float *X, int ldx, const float *W, const float *U, const int32_t *perm,
                int C, int T, int K, int P, int nt, float *Bt, int ldbt, int precision,
                void *workspace, size_t ws_bytes, void *stream) {
    return ksm_project_bounded(X, ldx, W, U, perm, C, T, K, P, nt, Bt, ldbt, precision, 0,
                               nullptr, 0, workspace, ws_bytes, stream);
}

int ksm_project_bounded(const float *X, int ldx, const float *W, const float *U,
                        const int32_t *perm, int C, int T, int K, int P, int nt, float *Bt,
                        int ldbt, int precision, int tile_m, void *bound, size_t bound_bytes,
                        void *workspace, size_t ws_bytes, void *stream) {
    if (!X || !W || !U || !Bt || C <= 0 || T <= 0 || K <= 0 || P <= 0 || nt <= 0 ||
        (nt % 2) == 0 || ldx < T)
        return ksm::fail(KSM_ERR_INVALID, "ksm_project: bad arguments");
    if (precision != KSM_PREC_FP32 && precision != KSM_PREC_BF16 &&
        precision != KSM_PREC_BF16X3 && precision != KSM_PREC_FP8)
        return ksm::fail(KSM_ERR_INVALID, "ksm_project: unknown precision %d", precision);
    if (tile_m != 0 && tile_m != 64 && tile_m != 128)
        return ksm::fail(KSM_ERR_INVALID, "ksm_project_bounded: tile_m %d (0, 64 or 128)", tile_m);
    if (bound && precision == KSM_PREC_FP32)
        return ksm::fail(KSM_ERR_INVALID,
                         "ksm_project_bounded: the FP32 form has no reduced-precision bound");
    if (bound && bound_bytes < ksm_bound_bytes(C, T, K, P, nt))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_project_bounded: bound buffer too small");
    const ProjectLayout L = project_layout(C, T, K, P);
    if (ldbt < K || (ldbt % 4) != 0)
        return ksm::fail(KSM_ERR_INVALID, "ksm_project: ldbt %d must be >= K and a multiple of 4",
                         ldbt);
    if (!workspace || ws_bytes < ksm_project_workspace(C, T, K, P, nt, precision))
        return ksm::fail(KSM_ERR_WORKSPACE, "ksm_project: workspace too small");
    hipStream_t s = ksm::as_stream(stream);
    if (precision == KSM_PREC_FP8)
        return ksm::project_fp8(X, ldx, W, U, perm, C, T, K, P, nt, Bt, ldbt, bound, workspace,
                                s);
    if (precision != KSM_PREC_FP32)
        return ksm::project_bf16(X, ldx, W, U, perm, C, T, K, P, nt, Bt, ldbt, precision, tile_m,
                                 bound, workspace, s);
    ksm::Carver cv(workspace);
    float *Bc = cv.take<float>((size_t)L.Rp * L.ldc);
    float *At = cv.take<float>((size_t)L.Rp * L.Kp);
    int *steps = cv.take<int>((size_t)L.nTm * L.nk);
    int *counts = cv.take<int>(L.nTm);
    int *nonfinite = cv.take<int>(1);

    ksm::launch(set_flag_kernel, 1, 1, 0, s, nonfinite);
    KSM_LAUNCH_CHECK("set_flag_kernel");
    {
        const int rc = ksm::conv_pc(X, ldx, W, C, T, P, nt, Bc, L.ldc, nonfinite, s);
        if (rc != KSM_OK) return rc;
    }
    if (L.Rp > L.R) {
        ksm::launch(zero_rows_kernel, 64, 256, 0, s, Bc + (size_t)L.R * L.ldc,
                    (size_t)(L.Rp - L.R) * L.ldc);
        KSM_LAUNCH_CHECK("zero_rows_kernel");
    }
    ksm::launch(transpose_pad_kernel, 512, 256, 0, s, U, perm, K, L.R, L.Kp, L.Rp, At);
    KSM_LAUNCH_CHECK("transpose_pad_kernel");
    ksm::launch(tile_steps_kernel, L.nTm, 256, 0, s, At, L.Kp, L.nk, W, P * nt, nonfinite, steps,
                counts);
    KSM_LAUNCH_CHECK("tile_steps_kernel");
    const int nTm = L.nTm, nTn = L.ldc / GBN;
    ksm::ProfileScope prof("gemm_f32", ksm::kProfProject, s);
    prof.launch(gemm_f32_mfma, dim3(nTm * nTn), dim3(256), 0u, At, L.Kp, Bc, L.ldc, L.nk, steps,
                counts, Bt, ldbt, K, (int)ksm::align_up(T, 128), nTm, nTn);
    KSM_LAUNCH_CHECK("gemm_f32_mfma");
    return KSM_OK;
}

}  // extern "C"
