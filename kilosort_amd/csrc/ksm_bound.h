// Error bound of reduced-precision scores: the layout of the buffer that
// ksm_project_bounded writes and the peel reads (see ksm.h).
#pragma once

#include "ksm_common.h"

namespace ksm {

// Channels per sub-block of the per-template bound (half a 64-lane wave).
constexpr int kBoundSub = 32;
// Largest channel count of the bounded forms (LDS of the weight kernel).
constexpr int kBoundMaxC = 4096;

enum BoundMode { kBoundLinear = 0, kBoundSqrt = 1 };

// Sections, each 256-byte aligned:
//   hdr   int32 [64]:  mode (BoundMode), nsb
//   colD  float [T]:   max over templates of D_r(t) + colA[t]
//   colA  float [T]:   per-column additive term
//   S     float [nsb][T]: per sub-block column statistic
//   Wt    float [nsb][K]: per sub-block template weight (peel order)
//   srng  int32 [K][2]:   significant channels [lo, hi) of template position r
//   Ttab  float [K][C][nt]: T_r[c][k] = sum_p U[perm[r],p,c] W[p,k] for the
//                           significant channels (row c - lo of position r)
// D_r(t) = sum_b Wt[b][r] S[b][t] (linear) or sqrt(sum_b Wt[b][r] S[b][t])
// (sqrt), and |B~[r,t] - B[r,t]| / sqrt(nm_r) <= D_r(t) + colA[t].
struct BoundLayout {
    int nsb;
    size_t o_hdr, o_colD, o_colA, o_S, o_Wt, o_srng, o_T, bytes;
};

inline BoundLayout bound_layout(int C, int T, int K, int nt) {
    BoundLayout L;
    L.nsb = (C + kBoundSub - 1) / kBoundSub;
    size_t o = 0;
    auto take = [&o](size_t n) {
        const size_t at = o;
        o = align_up(o + n, 256);
        return at;
    };
    L.o_hdr = take(64 * 4);
    L.o_colD = take((size_t)T * 4);
    L.o_colA = take((size_t)T * 4);
    L.o_S = take((size_t)L.nsb * T * 4);
    L.o_Wt = take((size_t)L.nsb * K * 4);
    L.o_srng = take((size_t)2 * K * 4);
    L.o_T = take((size_t)K * C * nt * 4);
    L.bytes = o;
    return L;
}

// Device view of a bound buffer.
struct BoundView {
    int mode, nsb;
    const float *colD, *colA, *S, *Wt, *T;
    const int *srng;
};

inline BoundView bound_view(const void *buf, int C, int T, int K, int nt, int mode) {
    const BoundLayout L = bound_layout(C, T, K, nt);
    const char *b = static_cast<const char *>(buf);
    BoundView v;
    v.mode = mode;
    v.nsb = L.nsb;
    v.colD = reinterpret_cast<const float *>(b + L.o_colD);
    v.colA = reinterpret_cast<const float *>(b + L.o_colA);
    v.S = reinterpret_cast<const float *>(b + L.o_S);
    v.Wt = reinterpret_cast<const float *>(b + L.o_Wt);
    v.srng = reinterpret_cast<const int *>(b + L.o_srng);
    v.T = reinterpret_cast<const float *>(b + L.o_T);
    return v;
}

// Fills a bound buffer from the correlation kernels' per-column statistics
// stats[(j * nsb2 + b) * ldst + t] over 32-channel sub-blocks b < nsb2
// (j = 0: sum of Bc^2, 1: largest Bc^2, 2: sum of (Bc / s)^2 -- FP8 only).
//   mode kBoundLinear (BF16 forms, proven):  Wt = kappa sqrt(E_b / nm),
//     S = sqrt(sum Bc^2), colA = 2^-14 ||Bc||;
//   mode kBoundSqrt (FP8, statistical):  Wt = zc^2 E_b / nm, S = max Bc^2,
//     colA = (2^-14 + acc) ||Bc|| + abs_terms[0] + abs_terms[1] ||Bc / s||.
// E_b = the template's energy on sub-block b, nm = its total.
int bound_finish(const float *stats, int nsb2, int ldst, const float *U, const float *W,
                 const int32_t *perm, int C, int T, int K, int P, int nt, int mode, double kappa,
                 double zc, double acc, const float *abs_terms, void *bound, hipStream_t s);

}  // namespace ksm
