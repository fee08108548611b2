/*
 * ksm.h -- C ABI of the MI355X (gfx950) learned-template matching library
 * (libksm.so).
 *
 * The library replaces the tensor operations of Kilosort4's matching-pursuit
 * spike extraction, kilosort/template_matching.py:14-244, with hand-written
 * HIP kernels.  Every entry point takes plain device pointers, sizes and a
 * stream (a hipStream_t passed as void*; NULL = the default stream), returns
 * a ksm_status code, and is asynchronous with respect to the host unless
 * stated otherwise.  No entry point allocates device memory: scratch space is
 * passed in as a workspace whose size the matching *_workspace query returns.
 *
 * Conventions (reference symbols in brackets):
 *   K  templates [Nfilt]      P  temporal PCs [n_pcs]      C  channels [Nchan]
 *   nt waveform length (odd)  T  padded batch length [NT + 2*nt]
 *   U    float32 [K][P][C]   spatial components, the layout run_matching takes
 *   W    float32 [P][nt]     temporal basis ops['wPCA']
 *   X    float32 [C][ldx]    one padded, whitened batch (T valid columns)
 *   ctc  float32 [K][K][ldl] template Gram, stored firing-template major:
 *        ctc[(j*K + i)*ldl + l] = reference ctc[i, j, l], l = 0..2*nt
 *        (the reference tensor is the permuted view (1, 0, 2) of this array)
 */
#ifndef KSM_H
#define KSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ksm_status {
    KSM_OK = 0,
    KSM_ERR_INVALID = 1,        /* bad sizes, pointers or parameters        */
    KSM_ERR_CAPACITY = 2,       /* more spikes than the st capacity         */
    KSM_ERR_HIP = 3,            /* a HIP runtime call or launch failed      */
    KSM_ERR_WORKSPACE = 4,      /* workspace smaller than the query result  */
    KSM_ERR_TIES = 5            /* reserved: no longer returned (ties of any
                                   size are peeled)                          */
} ksm_status;

typedef enum ksm_dtype {
    KSM_DTYPE_F32 = 0,          /* float32 arrays                            */
    KSM_DTYPE_F64 = 1           /* float64 arrays (gradient checks)          */
} ksm_dtype;

typedef enum ksm_precision {
    KSM_PREC_FP32 = 0,          /* FP32 correlation + FP32 MFMA projection   */
    KSM_PREC_BF16 = 1,          /* BF16 operands on MFMA, FP32 accumulation  */
    KSM_PREC_BF16X3 = 2,        /* split BF16 (hi*hi + hi*lo + lo*hi) on MFMA */
    KSM_PREC_FP8 = 3            /* scaled OCP e4m3 operands on MFMA, FP32 accumulation */
} ksm_precision;

/* Library version (major*10000 + minor*100 + patch). */
int ksm_version(void);

/* Static text for a status code. */
const char *ksm_status_string(int status);

/* Detail of the last error raised on the calling host thread ("" if none). */
const char *ksm_last_error(void);

/*
 * Peak channel and peak-channel PC features of every template.
 * Replaces template_matching.py:50-51 (prepare_extract):
 *   iU[i]       = argmax_c sum_p U[i,p,c]^2   (first index on ties)
 *   Ucc[n,i,p]  = U[i, p, iCC[n, iU[i]]]
 * iCC: int64 [nC][C] nearest-channel table (computed on the host with the
 * reference's own numpy argsort, template_matching.py:44-46).
 * Outputs iU int64 [K], Ucc float32 [nC][K][P].
 */
int ksm_prepare_extract(const float *U, int K, int P, int C,
                        const int64_t *iCC, int nC,
                        int64_t *iU, float *Ucc, void *stream);

/*
 * Template norms nm[i] = sum_p sum_c U[i,p,c]^2, accumulated in float64 and
 * rounded once.  Replaces template_matching.py:176 (run_matching).
 */
int ksm_template_norms(const float *U, int K, int P, int C, float *nm,
                       void *stream);

/* Workspace bytes of ksm_prepare_matching. */
size_t ksm_prepare_matching_workspace(int K, int P, int C, int nt);

/*
 * Template Gram ctc, formed in float64 and rounded once to float32.
 * Replaces template_matching.py:155-167 (prepare_matching):
 *   ctc[i,j,d+nt] = sum_c sum_k T_i[c,k] T_j[c,k+d],  T_i = U_i^T W,
 * written firing-template major (see the header comment), ldl >= 2*nt+1;
 * padding lags l in [2*nt+1, ldl) are written as 0.
 */
int ksm_prepare_matching(const float *U, const float *W, int K, int P, int C,
                         int nt, float *ctc, int ldl, void *workspace,
                         size_t ws_bytes, void *stream);

/*
 * Nonzero-row lists of the Gram: for each firing template j, the increasing
 * list rows[j*K + q], q < nrows[j], of templates i whose ctc[i, j, :] has a
 * nonzero entry.  Rows left out are exactly zero, so skipping them in the
 * peel update B -= amp*ctc changes nothing (template_matching.py:238).
 */
int ksm_ctc_rows(const float *ctc, int K, int L, int ldl, int32_t *rows,
                 int32_t *nrows, void *stream);

/*
 * Peel order of the templates: sorted by peak channel (argmax over channels
 * of sum_p U[j,p,c]^2), stable in the template id.  perm[r] = template at
 * position r, rank[j] = position of template j (int32 [K] each).  With
 * channels in probe order this makes the nonzero rows of every template's
 * Gram column a short contiguous range of positions.  No reference
 * counterpart: an internal layout choice of the peel.
 */
int ksm_template_order(const float *U, int K, int P, int C, int32_t *perm,
                       int32_t *rank, void *stream);

/*
 * Row ranges of the compact Gram: lohi[2r], lohi[2r+1] = [lo, hi) = the
 * smallest range of positions covering every template i with a nonzero
 * ctc[i, perm[r], :] (from ksm_ctc_rows), widened to multiples of 4 (hi may
 * exceed K by up to 3; those rows are zero), and off int64 [K+1]:
 * off[r] = sum_{r'<r} L*(hi-lo), off[K] = total floats of the compact Gram.
 */
int ksm_gram_ranges(const int32_t *rows, const int32_t *nrows, const int32_t *rank,
                    int K, int L, int32_t *lohi, int64_t *off, void *stream);

/*
 * Compact lag-major Gram for the peel:
 *   cc[off[r] + l*(hi-lo) + (q-lo)] = reference ctc[perm[q], perm[r], l],
 * q in [lo, hi) of position r (0 for q >= K), l = 0..L-1 (L = 2*nt+1).
 */
int ksm_gram_compact(const float *ctc, int ldl, const int32_t *perm,
                     const int32_t *lohi, const int64_t *off, int K, int L,
                     float *cc, void *stream);

/* Workspace bytes of ksm_project for the given sizes and precision. */
size_t ksm_project_workspace(int C, int T, int K, int P, int nt, int precision);

/*
 * Scores of every template at every time: the temporal-PC correlation of
 * each channel followed by the projection on the spatial components.
 * Replaces template_matching.py:180-181 (run_matching):
 *   B[i,t] = sum_p sum_c U[i,p,c] * sum_k W[p,k] * X[c, t + k - nt/2]
 * (zero padding outside [0, T)).  Written column-major in peel order:
 *   Bt[t*ldbt + r] = B[perm[r], t]   (perm = NULL: identity),
 * Bt float32 with T rounded up to 128 rows (rows >= T hold scores of the zero
 * padding), ldbt >= K and a multiple of 4.
 */
int ksm_project(const float *X, int ldx, const float *W, const float *U,
                const int32_t *perm, int C, int T, int K, int P, int nt, float *Bt,
                int ldbt, int precision, void *workspace, size_t ws_bytes,
                void *stream);

/* Bytes of the bound buffer of ksm_project_bounded. */
size_t ksm_bound_bytes(int C, int T, int K, int P, int nt);

/*
 * ksm_project for the reduced-precision forms (BF16, BF16X3, FP8), plus the
 * error bound of their scores, for the peel's decisions (no reference
 * counterpart: the reference scores templates in FP32 only).  bound: a
 * device buffer of ksm_bound_bytes(C, T, K, P, nt) bytes, handed to ksm_peel
 * with the same sizes.  Its sections, each 256-byte aligned: a 256-byte
 * header; colD float [T]; colA float [T]; S float [nsb][T]; Wt float
 * [nsb][K]; srng int32 [K][2]; Ttab float [K][C][nt] (nsb = ceil(C / 32)
 * sub-blocks of 32 channels; template positions r in perm order).  For every
 * template position r and column t, with D_r(t) = sum_b Wt[b][r] S[b][t]
 * (BF16 forms) or sqrt(sum_b Wt[b][r] S[b][t]) (FP8):
 *     |B~[r,t] - B[r,t]| / sqrt(nm_r) <= D_r(t) + colA[t] <= colD[t],
 * B~ the reduced-precision score, B the exact one, nm_r = sum U[perm[r]]^2.
 *   BF16 / BF16X3: a proof -- kappa sum |U||Bc| by Cauchy-Schwarz over each
 *     sub-block (kappa = 2u + u^2 + gamma_n, u = 2^-8; 3u^2(1+u)^3 + gamma_n
 *     for the split form), plus 2^-14 ||Bc[., t]|| for the FP32 rounding of
 *     later subtractions;
 *   FP8: a statistical margin -- 6 standard deviations of the modelled
 *     independent e4m3 rounding errors (variance <= (2/3) 2^-8 sum_b E_rb
 *     max_b Bc^2), plus worst-case accumulation and subnormal terms.
 * srng / Ttab: each template's significant channels [lo, hi) (energy at
 * least 2^-70 of its largest channel's) and its waveforms there, Ttab[r][c -
 * lo][k] = sum_p U[perm[r],p,c] W[p,k], from which the peel re-scores in FP32.
 * Same workspace as ksm_project.  KSM_PREC_FP32 is rejected; C <= 4096.
 * tile_m: template rows per tile of the BF16 forms' GEMM, 64 or 128 (0: 128),
 * an autotuning choice (same results: only exact-zero steps are skipped).
 */
int ksm_project_bounded(const float *X, int ldx, const float *W, const float *U,
                        const int32_t *perm, int C, int T, int K, int P, int nt, float *Bt,
                        int ldbt, int precision, int tile_m, void *bound, size_t bound_bytes,
                        void *workspace, size_t ws_bytes, void *stream);

/* Workspace bytes of ksm_peel. */
size_t ksm_peel_workspace(int K, int T, int max_peels);

/*
 * Matching pursuit (the peel).  Replaces template_matching.py:183-244:
 * up to max_peels iterations of
 *   Cf = relu(B)^2 / nm, first and last nt columns zeroed;
 *   accept columns x where max over +-nt of max_i Cf[i, .] exceeds Th^2 and
 *   equals max_i Cf[i, x]; record (x, argmax_i), amps = B/nm,
 *   th_amps = sqrt(Cmax); subtract amps * (U^T W) from Xres and
 *   amps * ctc[:, id, :] from B in two interleaved phases (even-ranked
 *   spikes first), stopping early when nothing is accepted.
 * Bt (from ksm_project with the same perm) and Xres (a copy of X) are updated
 * in place; nm float32 [K] by template id (ksm_template_norms); perm/rank from
 * ksm_template_order; cc/lohi/off from ksm_gram_ranges + ksm_gram_compact.
 * flags: KSM_PEEL_RESCORE (below) with the bound buffer of
 * ksm_project_bounded, or 0 with bound NULL (FP32 scores, decided as they
 * are).
 * Outputs: st int64 [capacity][2] (time column, template id), amps and
 * th_amps float32 [capacity], and info int32 [4] on the device:
 *   info[0] = spike count k, info[1] = ksm_status of the device run,
 *   info[2] = iterations that accepted spikes, info[3] = (column, template)
 *   scores re-scored in FP32 (KSM_PEEL_RESCORE).
 * Spikes are ordered as the reference orders them: by iteration, then time.
 * Limits: K <= 8192, P <= 8, C <= 65535, odd nt <= 255.
 */
int ksm_peel(float *Bt, int ldbt, float *Xres, int ldx, const float *U,
             const float *W, const float *nm, const int32_t *perm,
             const int32_t *rank, const float *cc, const int32_t *lohi,
             const int64_t *off, int K, int P, int C, int T, int nt, float Th,
             int max_peels, int flags, const void *bound, int64_t *st, float *amps,
             float *th_amps, int capacity, int32_t *info, void *workspace,
             size_t ws_bytes, void *stream);

/*
 * ksm_peel without a workspace, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as ksm_peel; the peel state lives in one static device array of
 * ksm_peel_static_capacity() bytes (32 MiB) per GPU, held until the process
 * ends, whose device address is looked up once per device on the first call.
 * Sizes whose ksm_peel_workspace exceeds it return KSM_ERR_WORKSPACE (e.g.
 * T = 60122 holds K <= 2048).  The call allocates nothing (no device or pinned
 * host memory, no timing events), so every iteration up to max_peels is
 * enqueued (the device skips them once the peel has ended).  Every call on
 * one GPU shares its array: calls that overlap on one GPU (other streams or
 * host threads) race and must be serialised by the caller.
 */
int ksm_peel_static(float *Bt, int ldbt, float *Xres, int ldx, const float *U,
                    const float *W, const float *nm, const int32_t *perm,
                    const int32_t *rank, const float *cc, const int32_t *lohi,
                    const int64_t *off, int K, int P, int C, int T, int nt, float Th,
                    int max_peels, int flags, const void *bound, int64_t *st,
                    float *amps, float *th_amps, int capacity, int32_t *info, void *stream);

/* Bytes of the static peel state of ksm_peel_static. */
size_t ksm_peel_static_capacity(void);

/*
 * Iterations [it_begin, it_end) of the peel of ksm_peel (no reference
 * counterpart: a split of the same loop), for callers that capture the peel
 * in a graph or run it in parts.  it_begin == 0 starts a peel (state,
 * tables, first pass over Bt); a later range continues from the state the
 * previous range left in `workspace` (same workspace, Bt, Xres, outputs and
 * sizes).  Nothing is polled or timed: every iteration of the range is
 * enqueued and those after the end return at once.  info as ksm_peel's after
 * the range; the loop has ended when info[2] < it_end (an iteration accepted
 * nothing) or it_end == max_peels.
 */
int ksm_peel_range(float *Bt, int ldbt, float *Xres, int ldx, const float *U,
                   const float *W, const float *nm, const int32_t *perm,
                   const int32_t *rank, const float *cc, const int32_t *lohi,
                   const int64_t *off, int K, int P, int C, int T, int nt, float Th,
                   int max_peels, int flags, const void *bound, int64_t *st, float *amps,
                   float *th_amps, int capacity, int32_t *info, void *workspace,
                   size_t ws_bytes, int it_begin, int it_end, void *stream);

/* ksm_peel flags. */
#define KSM_PEEL_RESCORE 1  /* the scores come from a reduced-precision
                               projection (ksm_project_bounded, same perm) and
                               bound is its bound buffer: every decision the
                               bound leaves open is taken on scores re-scored
                               in FP32 from the residual, and accepted spikes
                               take their FP32 score for amps and th_amps.
                               Without it bound must be NULL. */

/*
 * PC features of each spike on the nC channels nearest its template's peak
 * channel: the final residual projected on W plus the spike's own template.
 * Replaces template_matching.py:85-86 (extract):
 *   tF[s,n,p] = sum_k Xres[iCC[n, iU[id_s]], t_s + k - nt/2] W[p,k]
 *             + amps[s] * Ucc[n, id_s, p]
 * st int64 [nsp][2] as produced by ksm_peel; output tF float32 [nsp][nC][P].
 * Limits: nC <= 32, nt <= 255.
 */
int ksm_features(const float *Xres, int ldx, const float *W,
                 const int64_t *st, const float *amps, int nsp,
                 const int64_t *iCC, const int64_t *iU, const float *Ucc,
                 int nC, int K, int P, int C, int nt, float *tF, void *stream);

/*
 * ksm_features for a spike count on the device (a peel's info[0]; no
 * reference counterpart): features of spikes s < min(info[0], capacity),
 * from a fixed grid, so that the call can be captured in a graph with the
 * peel before the count is known on the host.
 */
int ksm_features_counted(const float *Xres, int ldx, const float *W, const int64_t *st,
                         const float *amps, const int32_t *info, int capacity,
                         const int64_t *iCC, const int64_t *iU, const float *Ucc, int nC, int K,
                         int P, int C, int nt, float *tF, void *stream);

/*
 * Batch graphs (no reference counterpart: the reference enqueues each batch's
 * work op by op, template_matching.py:79-110).  Between ksm_graph_begin(s)
 * and ksm_graph_end(s, &g), on one host thread, the work that these entry
 * points would enqueue on stream s is recorded as the nodes of one graph, in
 * call order, and does not run: ksm_project, ksm_project_bounded,
 * ksm_peel_range, ksm_peel_static, ksm_features, ksm_features_counted,
 * ksm_copy_async.  (Other entry points run at once; ksm_peel, whose host
 * follows the device's progress, returns KSM_ERR_INVALID on s.)  The graph is built node by node, not by stream capture, so
 * other host threads' runtime calls meanwhile do not interact with it.
 * ksm_graph_end instantiates it (g = NULL and an error when a recorded call
 * failed or nothing was recorded, and the recording is discarded either
 * way); ksm_graph_launch(g, s) runs it on s; ksm_graph_destroy(g) frees it
 * once no launch of it is pending.  Argument values are taken at recording:
 * every launch works on the same buffers.
 */
int ksm_graph_begin(void *stream);
int ksm_graph_end(void *stream, void **graph_exec);
int ksm_graph_launch(void *graph_exec, void *stream);
int ksm_graph_destroy(void *graph_exec);

/* ksm_copy_async kinds. */
#define KSM_COPY_DEVICE_TO_DEVICE 0
#define KSM_COPY_DEVICE_TO_HOST   1   /* dst: pinned host memory */

/* An ordered copy of `bytes` on `stream` (recorded when the stream is). */
int ksm_copy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);

/*
 * Per-kernel timing (benchmark and test use; no reference counterpart).
 * ksm_profile_enable(mask) turns on HIP event pairs on the launch stream
 * around the kernels of the selected groups: bit 0 = projection
 * ("conv_pc", "gemm_f32"), bit 1 = peel ("colmax_all", "peel_select",
 * "peel_update"); 0 turns timing off.  ksm_profile_query synchronises the
 * recorded events and returns the summed milliseconds and launch count of
 * one kernel name.  ksm_profile_launch returns launch i of every recorded
 * launch in launch order: its kernel name, its start relative to the first
 * recorded launch's start and its duration, in milliseconds (the start
 * offsets are comparable between launches of one stream).
 */
int ksm_profile_enable(int mask);
int ksm_profile_reset(void);
int ksm_profile_query(const char *name, double *total_ms, int *launches);
int ksm_profile_launch(int i, char *name, int name_cap, double *start_ms, double *dur_ms);

/*
 * Template refinement: the residual of a fixed spike set and its gradients.
 * The reference refines templates by averaging spike features
 * (clustering_qr.py:524-527) and has no gradients; the model is its own
 * subtraction (template_matching.py:235-236): spike s = (x_s, id_s) = st[s]
 * with amplitude amps[s] contributes amps[s] * sum_p U[id,p,c] * W[p,k] at
 * sample x_s - nt/2 + k.  Arrays are float32 or float64 (dtype); U [K][P][C],
 * W [P][nt], X and R [C][ld*], st int64 [nsp][2], ids in [0, K).
 */

/* Workspace bytes of ksm_residual and ksm_residual_backward. */
size_t ksm_residual_workspace(int nsp, int K, int P, int C, int T, int nt, int dtype);

/*
 * R = X - sum_s amps[s] T_{id_s}(. - x_s) (R may be NULL) and loss[0] =
 * sum R^2 in float64 (loss may be NULL).  order: spike indices sorted by
 * x (non-decreasing); contributions are summed in that order.
 */
int ksm_residual(const void *X, int ldx, const void *U, const void *W, const int64_t *st,
                 const void *amps, const int32_t *order, int nsp, int K, int P, int C, int T,
                 int nt, void *R, int ldr, double *loss, int dtype, void *workspace,
                 size_t ws_bytes, void *stream);

/*
 * Gradients of scale/2 * sum R^2 (scale = 2 dL for the plain loss) with
 * spike times and ids fixed, from the residual R of ksm_residual:
 *   dU [K][P][C], dW [P][nt], damp [nsp] (by spike index).
 * order_k: spike indices sorted by (id, x); seg [K+1]: template k's spikes are
 * order_k[seg[k] .. seg[k+1]).  Deterministic (fixed summation orders).
 */
int ksm_residual_backward(const void *R, int ldr, const void *U, const void *W,
                          const int64_t *st, const void *amps, const int32_t *order_k,
                          const int32_t *seg, int nsp, int K, int P, int C, int T, int nt,
                          double scale, void *dU, void *dW, void *damp, int dtype,
                          void *workspace, size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Batch preprocessing (the step before the matching path).
 */

/* Raw sample layouts of ksm_filter_batch. */
typedef enum ksm_raw {
    KSM_RAW_I16 = 0,            /* int16   [nsamp][ldraw], time-major (binary file)   */
    KSM_RAW_U16 = 1,            /* uint16  [nsamp][ldraw], shifted by -2^15          */
    KSM_RAW_F32 = 2,            /* float32 [nsamp][ldraw], time-major                */
    KSM_RAW_F32_CM = 3          /* float32 [n_chan][ldraw], channel-major (a batch)  */
} ksm_raw;

#define KSM_FILTER_INVERT     1    /* invert_sign                                    */
#define KSM_FILTER_CAR        2    /* do_CAR: subtract the median across channels    */
#define KSM_FILTER_SKIP       4    /* skip_preproc: stop after the median stage      */
#define KSM_FILTER_SCALE      8    /* multiply the raw values by `scale`             */
#define KSM_FILTER_SHIFT      16   /* then add `shift`                               */
#define KSM_FILTER_EDGE_FIRST 32   /* columns before `off` repeat raw row 0          */
#define KSM_FILTER_EDGE_LAST  64   /* columns after the data repeat the last row     */
#define KSM_FILTER_WHITEN     128  /* `whiten` is given (workspace for its input)    */

/* Workspace bytes of ksm_filter_batch (J = -1: no high-pass filter). */
size_t ksm_filter_workspace(int C, int T, int J, int flags);

/*
 * One padded batch through the reference's filter.  Replaces
 * BinaryRWFile.padded_batch_to_torch (io.py:789-831) followed by
 * BinaryFiltered.filter (io.py:990-1026):
 *   column t of the batch holds raw row t - off (0 <= t - off < nsamp); other
 *   columns repeat row 0 / the last row (EDGE_FIRST / EDGE_LAST) or are 0;
 *   X[c, t] = converted raw[., chan_map[c]] (chan_map NULL: identity), times -1
 *   with INVERT; minus the mean over t of each row; minus the lower median
 *   over c of each column (CAR); high-pass
 *   out[c, n] = sum_{j=-J..J} taps[J + j] * x[c, (n + j) mod T]
 *   (taps NULL: none; the reference's Fourier-domain filter equals this
 *   circular correlation with its impulse response centred at T/2); if any
 *   |out| >= artifact_threshold the batch becomes 0 (INFINITY: no test);
 *   X = whiten @ out (whiten [C][C] row-major, already multiplied by the
 *   batch's drift matrix when there is one; NULL: none).
 * X float32 [C][ldx] on the device.  Limits: C <= 1536, J <= 4096, 2J+1 <= T.
 */
int ksm_filter_batch(const void *raw, int fmt, int64_t ldraw, int nsamp, int off,
                     const int32_t *chan_map, int C, int T, int flags, float scale,
                     float shift, const float *taps, int J, float artifact_threshold,
                     const float *whiten, float *X, int ldx, void *workspace,
                     size_t ws_bytes, void *stream);

/*
 * Row-major FP32 product C[M][N] = A[M][K] @ B[K][N] on MFMA (FP32
 * accumulation), for the drift-matrix products (Kyx @ iKxx, M @ Wrot) of
 * preprocessing.py:67 and io.py:1021-1023.
 */
int ksm_gemm_f32(const float *A, int lda, const float *B, int ldb, float *Cm, int ldc,
                 int M, int N, int K, void *stream);

/* ---------------------------------------------------------------------------
 * Universal-template spike detection (the pass before clustering).
 */

/* Workspace bytes of ksm_detect. */
size_t ksm_detect_workspace(int C, int T, int nk, int Nfilt);

/*
 * Replaces spikedetect.template_match (spikedetect.py:133-178) and the
 * per-batch features of spikedetect.run (:256-281, yweighted :190-197):
 *   B[k][c][t] = sum_m wTEMP[k, m] X[c, t + m - nt/2];
 *   A[i, k, f, t] = sum_j weigh[i, j, f] B[k][iC[j, f]][t];
 *   As = max over q = i*nk + k of |A| (first q on ties), Im = (1 + q) sign(A_q);
 *   spikes where As > th and As equals the maximum of As over the nC2
 *   templates iC2[:, f] and over |dt| <= nt0 (columns outside [nt, T-nt)
 *   count as 0), in (f, t) order.
 * iC int32 [nC][Nfilt], iC2 int32 [nC2][Nfilt], weigh [nsizes][nC][Nfilt],
 * yc [C], wPCA [P][nt].  Outputs for spikes s < cap: xy [cap][2] = (f, t),
 * imax [cap] = q, amp [cap] = As, adist [nC][cap] = B[q % nk][iC[j, f]][t] *
 * sign, yct [cap] (relu(adist)-weighted y of the channels), tF
 * [cap][nC][P] (window t - nt/2 .. t + nt/2 of X on the channels, times wPCA);
 * *count (device) = the number of spikes, also when it exceeds cap.
 * Limits: nC <= 16, nsizes <= 8, nsizes*nC <= 64, nk <= 8, nC*P <= 64, nt0 <= 256.
 */
int ksm_detect(const float *X, int ldx, int C, int T, const float *wTEMP, int nk, int nt,
               const int32_t *iC, int nC, const int32_t *iC2, int nC2, const float *weigh,
               int nsizes, int Nfilt, int nt0, float th, const float *yc, const float *wPCA,
               int P, int cap, int32_t *xy, int32_t *imax, float *amp, float *adist,
               float *yct, float *tF, int *count, void *workspace, size_t ws_bytes,
               void *stream);

/* Workspace bytes of ksm_snippets. */
size_t ksm_snippets_workspace(int C, int T);

/*
 * Replaces spikedetect.extract_snippets (spikedetect.py:34-51): isolated
 * single-channel peaks -- |X[c,t]| equal to the maximum of |X| over channels
 * c +- lc, samples t +- lt (outside: -inf) and above th, with no other such
 * peak over c +- ic, t +- it, and nt <= t < T - nt -- in (channel, time)
 * order; clips [cap][nt] = X[c][t - twav_min + k], xy [cap][2] = (c, t);
 * *count (device) = the number of peaks, also beyond cap.
 * Limits: lc <= 16, lt <= 16, 2 nt <= T.
 */
int ksm_snippets(const float *X, int ldx, int C, int T, int nt, int twav_min, float th, int lc,
                 int lt, int ic, int it, int cap, float *clips, int32_t *xy, int *count,
                 void *workspace, size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Template Gram consumers (post-processing of learned templates).
 */

/* Workspace bytes of ksm_similarity. */
size_t ksm_similarity_workspace(int K, int C, int P, int nt);

/*
 * Replaces CCG.similarity (CCG.py:94-102) and the per-step Gram rows of
 * merging_function (template_matching.py:285-293):
 *   Wnorm[k] = Wall[k] / (1e-6 + ||Wall[k]||),  Wall float32 [K][C][P];
 *   S[i][j] = max over lags l of sum_{p,q} UtU[i,j,p,q] WtW[p,q,l]
 *           (the ctc of ksm_prepare_matching of Wnorm, float64 sums);
 *   lag[i][j] = the first lag index (0 .. 2nt) attaining it (NULL: not written).
 */
int ksm_similarity(const float *Wall, const float *W, int K, int C, int P, int nt, float *S,
                   int32_t *lag, void *workspace, size_t ws_bytes, void *stream);

/* Workspace bytes of ksm_align_templates. */
size_t ksm_align_workspace(int K, int C, int P, int nt, int nk);

/*
 * Replaces template_matching.align_U (template_matching.py:128-141):
 * Wall float32 [K][C][P], wPCA [P][nt], wTEMP [nk][nt];
 *   Y[k][c][t] = sum_p Wall[k][c][p] wPCA[p][t];
 *   imax[k] = the first t maximising max over waveforms q and channels c of
 *     |sum_m wTEMP[q][m] Y'[c][k nt + t + m - nt/2]|, Y' the channel rows of Y
 *     laid end to end over templates (zero outside);
 *   Unew[k][p][c] = sum_t wPCA[p][t] Y[k][c][(t - nt/2 + imax[k]) mod nt].
 * nt odd, <= 256.
 */
int ksm_align_templates(const float *Wall, const float *wPCA, const float *wTEMP, int K, int C,
                        int P, int nt, int nk, float *Unew, int32_t *imax, void *workspace,
                        size_t ws_bytes, void *stream);

/*
 * Replaces postprocessing.compute_spike_positions (postprocessing.py:33-62):
 * per spike s with template ids[s] and peak channel u = iU[ids[s]]
 * (iU as from ksm_prepare_extract, in [0, C)), weights ||tF[s, j, :]||_2 *
 * mask[j][u] normalised over j, and (xs, ys)[s] = the weighted mean of
 * (xc, yc)[iCC[j][u]].  iCC int64 and mask uint8 are [nC][C] (nearest
 * channels of every channel and their position mask); tF float32
 * [n][nC][P]; all on the device.  The caller keeps ids in range.  nC <= 32.
 * A spike with no unmasked feature mass gets NaN, as the reference's 0/0.
 * amp (may be NULL) gets ||tF[s]||_F over channels and PCs, the amplitudes
 * save_to_phy writes (io.py:330), from the same pass over tF.
 */
int ksm_spike_positions(const float *tF, const int64_t *ids, long long n, int nC, int P,
                        const int64_t *iU, const int64_t *iCC, const uint8_t *mask, int C,
                        const float *xc, const float *yc, float *xs, float *ys, float *amp,
                        void *stream);

/*
 * Output formats (the per-spike work of io.save_to_phy, io.py:248-469).
 *
 * ksm_group: stable grouping of spikes by label (cluster id in
 * [0, n_groups), n_groups <= 16384, n < 2^31): offsets [n_groups + 1] and
 * order [n], with order[offsets[g] .. offsets[g+1]) the spikes of label g in
 * input order.  The grouping that postprocessing.remove_duplicates
 * (postprocessing.py:9-30, a dict of labels) and CCG.refract (CCG.py:105-120,
 * one boolean mask per label) make on the host.  Labels out of range are
 * left out.  Workspace from ksm_group_workspace (0 = invalid sizes).
 */
size_t ksm_group_workspace(long long n, int n_groups);
int ksm_group(const int32_t *labels, long long n, int n_groups, int64_t *offsets, int64_t *order,
              void *workspace, size_t ws_bytes, void *stream);

/*
 * Replaces postprocessing.remove_duplicates (postprocessing.py:9-30): with
 * the grouping of ksm_group, keep[i] = 1 when spike i is the first of its
 * label or at least dt samples after the last kept spike of its label,
 * walking each label's spikes in input order.
 */
int ksm_remove_duplicates(const int64_t *times, long long n, int n_groups, const int64_t *offsets,
                          const int64_t *order, long long dt, uint8_t *keep, void *stream);

/*
 * Replaces CCG.compute_CCG (CCG.py:8-36) for the autocorrelograms of
 * CCG.refract (:105-120): for every label g with active[g], counts[g][b]
 * (b < 2 nbins + 1, uint32) = pairs (j, k) of its spikes with
 * t_j - dt < t_k <= t_j + dt, dt = nbins * tbin, in bin
 * rint((t_j - t_k) / tbin) + nbins.  times float64 seconds, sorted within
 * each label (the reference sorts each label's times); grouping from
 * ksm_group.  nbins <= 1023.  Exact integer counts.
 */
int ksm_acg(const double *times, long long n, int n_groups, const int64_t *offsets,
            const int64_t *order, const uint8_t *active, int nbins, double tbin, uint32_t *counts,
            void *stream);

/*
 * Replaces CCG.compute_CCG (CCG.py:8-36) for two trains, as called by
 * check_CCG from merging_function (template_matching.py:300-304): counts
 * [2 nbins + 1] (uint32) of pairs (j of t2, k of t1) with
 * t2_j - dt < t1_k <= t2_j + dt, dt = nbins * tbin, in bin
 * rint((t2_j - t1_k) / tbin) + nbins.  t1, t2 float64 seconds, each sorted
 * (the reference sorts both).  nbins <= 1023.
 */
int ksm_ccg(const double *t1, long long n1, const double *t2, long long n2, int nbins,
            double tbin, uint32_t *counts, void *stream);

/*
 * The feature shift of template_matching.roll_features (:361-379) in PC
 * space: X [n][R][P] float32 on the device; for every spike rows[i],
 * X[rows[i]][r][:] <- X[rows[i]][r][:] @ M, M [P][P] float32 (the caller's
 * wPCA . S_dt . wPCA^T, S_dt the clamped time shift).  In place.  P <= 16.
 */
int ksm_pc_map(float *X, const int64_t *rows, long long nrows, int R, int P, const float *M,
               void *stream);

/*
 * The spike-level part of postprocessing.make_pc_features
 * (postprocessing.py:65-130, with clustering_qr.get_data_cpu,
 * clustering_qr.py:560-600): per detection template t < K, sums[t][j][p]
 * (float64) = the sum of tF[s][j][p] over its spikes and counts[t] their
 * number; presence (may be NULL) [n_groups][K] = 1 where label g holds a
 * spike of template t.  tF float32 [n][nC][P]; template ids int64 in
 * [0, K) and labels in [0, n_groups), kept in range by the caller;
 * n < 2^31, K <= 16384.  workspace: ksm_template_sums_workspace(n, K) bytes of device
 * memory (a counting sort of the spikes by template: each feature row is
 * read once and summed by its template's block, no floating-point atomics).
 */
int ksm_template_sums(const float *tF, const int64_t *tmpl, const int32_t *labels, long long n,
                      int nC, int P, int K, int n_groups, double *sums, uint64_t *counts,
                      uint8_t *presence, void *workspace, size_t ws_bytes, void *stream);

/* Device scratch bytes ksm_template_sums needs for n spikes of K templates
 * (its counting sort of the spikes by template). */
size_t ksm_template_sums_workspace(long long n, int K);

/*
 * The per-cluster channel ranking of make_pc_features (postprocessing.py:
 * 104-127), replayed in ascending cluster order on the output of
 * ksm_template_sums: for each label g holding spikes, the channels of its
 * templates (iC [nC][K] int64: channel of feature row j of template t, < C),
 * the norm of the mean current feature row per channel, the nC channels of
 * largest norm (ties: lowest channel) into feature_ind[g][0..nC) (uint32,
 * [n_groups][nC]; rows of labels without spikes untouched), and sigma
 * [K][nC] (int32; row k of template t's spikes <- original row sigma[t][k],
 * or zeros for -1) composed with each choice, as the reference's in-place
 * rewrites of tF do.  sigma_tmp [K][nC] is scratch.  One block; limits
 * nC <= 32, C <= 2048 and C (8 P + 12) + 2 K (nC + 1) <= 57344 bytes of LDS
 * (e.g. 383 channels, 6 PCs and up to 1560 templates of 10 channels).
 */
int ksm_pc_channel_rank(const double *sums, const uint64_t *counts, const uint8_t *presence,
                        const int64_t *iC, int K, int nC, int P, int C, int n_groups,
                        int32_t *sigma, int32_t *sigma_tmp, uint32_t *feature_ind,
                        void *stream);

/*
 * The feature rewrite of make_pc_features (postprocessing.py:120-127):
 * out[s][k][p] (layout 0) or out[s][p][k] (layout 1, the Phy pc_features
 * layout) = tF[s][sigma[t][k]][p], or 0 where sigma[t][k] < 0, t = the
 * spike's template; sigma int32 [K][nC].  out must not alias tF.
 */
int ksm_gather_features(const float *tF, const int64_t *tmpl, long long n, int nC, int P,
                        const int32_t *sigma, float *out, int layout, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* KSM_H */
