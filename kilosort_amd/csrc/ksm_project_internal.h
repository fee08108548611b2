// Entry points shared between the projection sources.
#pragma once

#include "ksm_common.h"

namespace ksm {

// Temporal correlation Bc[(p*C + c)*ldc + t] = sum_k W[p,k] X[c, t+k-nt/2]
// (zero padding; columns T <= t < ldc hold 0), FP32 as the projection's;
// raises *nonfinite on a non-finite sample.  ldc: a multiple of 256.
int conv_pc(const float *X, int ldx, const float *W, int C, int T, int P, int nt, float *Bc,
            int ldc, int *nonfinite, hipStream_t s);

// Workspace bytes of the BF16 projection forms.
size_t project_bf16_workspace(int C, int T, int K, int P, int precision);

// BF16 / split-BF16 projection into Bt (see ksm_project_bf16.hip).
int project_bf16(const float *X, int ldx, const float *W, const float *U, const int32_t *perm,
                 int C, int T, int K, int P, int nt, float *Bt, int ldbt, int precision,
                 int tile_m, void *bound, void *workspace, hipStream_t s);

// Workspace bytes of the FP8 projection form.
size_t project_fp8_workspace(int C, int T, int K, int P);

// FP8 (scaled e4m3) projection into Bt (see ksm_project_fp8.hip).
int project_fp8(const float *X, int ldx, const float *W, const float *U, const int32_t *perm,
                int C, int T, int K, int P, int nt, float *Bt, int ldbt, void *bound,
                void *workspace, hipStream_t s);

}  // namespace ksm
