/*
 * gpumlib_nmf.h — C-ABI of the MI355X (gfx950) NMF library (multiplicative and
 * additive updates).
 *
 * Plain C types only: device pointers, sizes, a hipStream_t and hipError_t
 * return codes.  Matrices are column-major, as in GPUMLib
 * (src/algorithms/nmf/nmf_base.h:109-111):
 *   V is n x m, V(i,u) = V[u*n + i];  W is n x r, W(i,a) = W[a*n + i];
 *   H is r x m, H(a,u) = H[u*r + a].
 * No entry point allocates device memory: scratch space is a caller-provided
 * workspace whose size the *_workspace_bytes queries return.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the GPUMLib source tree).
 */
#ifndef GPUMLIB_NMF_H
#define GPUMLIB_NMF_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* 1. The reference's own extern "C" launchers (src/cuda/nmf/nmf_kernels.h).  */
/*    Same names and arguments; cudaError_t becomes hipError_t.  The grid and */
/*    block arguments are accepted for drop-in compatibility and ignored: the */
/*    launch geometry is chosen for gfx950.  Work goes to the null stream, as */
/*    in the reference.  The ratios use eps = 1e-8f, the value of             */
/*    SMALL_VALUE_TO_ADD_DENOMINATOR (nmf_kernels.h:34).                      */
/* ------------------------------------------------------------------------ */

/* replaces gpumlib_cuda_nmf_multiplicative_euclidean_update_matrix, nmf_kernels.h:87 */
hipError_t gpumlib_cuda_nmf_multiplicative_euclidean_update_matrix(unsigned int gridDim, unsigned int blockDim,
                                                                   float *nm, float *dm, float *m, int elements);
/* replaces gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_w, nmf_kernels.h:100 */
hipError_t gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_w(dim3 gridDim, dim3 blockDim, float *W, float *H,
                                                                      float *V, float *WH, float *sumH, int n, int m,
                                                                      int r);
/* replaces gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_h, nmf_kernels.h:113 */
hipError_t gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_h(dim3 gridDim, dim3 blockDim, float *H, float *W,
                                                                      float *V, float *WH, float *sumW, int n, int m,
                                                                      int r);

/* replaces gpumlib_cuda_nmf_additive_euclidean_update_matrix, nmf_kernels.h:116 (kernel nmf_kernels.cu:31-39):
 *   X[k] = X[k] + (X[k] / deltaX2[k]) * (deltaX1[k] - deltaX2[k]), 0 if negative (no eps) */
hipError_t gpumlib_cuda_nmf_additive_euclidean_update_matrix(unsigned int gridDim, unsigned int blockDim, float *X,
                                                             float *deltaX1, float *deltaX2, int elements);
/* replaces gpumlib_cuda_nmf_additive_divergence_update_matrix_h, nmf_kernels.h:129
 * (kernel nmf_kernels_additive_divergence.cu:126-193) */
hipError_t gpumlib_cuda_nmf_additive_divergence_update_matrix_h(dim3 gridDim, dim3 blockDim, float *H, float *W,
                                                                float *V, float *WH, float *sumW, int n, int m, int r);
/* replaces gpumlib_cuda_nmf_additive_divergence_update_matrix_w, nmf_kernels.h:142
 * (kernel nmf_kernels_additive_divergence.cu:57-124) */
hipError_t gpumlib_cuda_nmf_additive_divergence_update_matrix_w(dim3 gridDim, dim3 blockDim, float *W, float *H,
                                                                float *V, float *WH, float *sumH, int n, int m, int r);

/* ------------------------------------------------------------------------ */
/* 2. Kernel-level entry points with an explicit stream and eps.              */
/* ------------------------------------------------------------------------ */

/* m[k] *= nm[k] / (dm[k] + eps), k < elements.  Kernel of nmf_kernels.cu:43-47. */
hipError_t gpumlib_nmf_mu_update(const float *nm, const float *dm, float *m, long elements, float eps,
                                 hipStream_t stream);

/* sumW[a] = max(sum_i W(i,a), eps).  Replaces KernelSumW (nmf_kernels.h:64; nmf_kernels.cu:63-147). */
hipError_t gpumlib_nmf_sum_w(const float *W, int n, int r, float eps, float *sumW, hipStream_t stream);

/* sumH[a] = max(sum_u H(a,u), eps).  Replaces KernelSumH (nmf_kernels.h:73; nmf_kernels.cu:149-233). */
hipError_t gpumlib_nmf_sum_h(const float *H, int r, int m, float eps, float *sumH, hipStream_t stream);

/* KL updates over a materialised WH = W H (the reference kernel contract):
 *   H(a,u) *= [sum_i W(i,a) V(i,u) / (WH(i,u) + eps)] / sumW[a]   nmf_kernels.cu:295-341
 *   W(i,a) *= [sum_u H(a,u) V(i,u) / (WH(i,u) + eps)] / sumH[a]   nmf_kernels.cu:247-293 */
hipError_t gpumlib_nmf_kl_update_h(float *H, const float *W, const float *V, const float *WH, const float *sumW,
                                   int n, int m, int r, float eps, hipStream_t stream);
hipError_t gpumlib_nmf_kl_update_w(float *W, const float *H, const float *V, const float *WH, const float *sumH,
                                   int n, int m, int r, float eps, hipStream_t stream);

/* Additive-update element rule (kernel of nmf_kernels.cu:31-39), explicit stream:
 *   X[k] = X[k] + (X[k] / d2[k]) * (d1[k] - d2[k]), set to 0 when negative. */
hipError_t gpumlib_nmf_additive_update(float *X, const float *d1, const float *d2, long elements, hipStream_t stream);

/* Additive KL updates over a materialised WH (nmf_kernels_additive_divergence.cu):
 *   H(a,u) = max(0, H + (H / sumW[a]) sum_i W(i,a) (V(i,u) / (WH(i,u) + eps) - 1))   :126-193
 *   W(i,a) = max(0, W + (W / sumH[a]) sum_u H(a,u) (V(i,u) / (WH(i,u) + eps) - 1))   :57-124 */
hipError_t gpumlib_nmf_additive_kl_update_h(float *H, const float *W, const float *V, const float *WH,
                                            const float *sumW, int n, int m, int r, float eps, hipStream_t stream);
hipError_t gpumlib_nmf_additive_kl_update_w(float *W, const float *H, const float *V, const float *WH,
                                            const float *sumH, int n, int m, int r, float eps, hipStream_t stream);

/* Reference quality  sum_k V[k] log10(WH[k] + eps) - WH[k]  accumulated in FP64 into d_quality[0].
 * Replaces KernelNMFquality (nmf_kernels.h:44; nmf_kernels_quality.cu:29-109), which reduced in
 * FP32 on one workgroup. */
size_t gpumlib_nmf_quality_workspace_bytes(long count);
hipError_t gpumlib_nmf_quality(const float *V, const float *WH, long count, float eps, void *workspace,
                               size_t workspace_bytes, double *d_quality, hipStream_t stream);
/* The same quality, written as a float to d_quality[0] (the type KernelNMFquality writes, nmf_kernels.h:44);
 * stream-ordered, nothing read back to the host. */
hipError_t gpumlib_nmf_quality_f32(const float *V, const float *WH, long count, float eps, void *workspace,
                                   size_t workspace_bytes, float *d_quality, hipStream_t stream);

/* C = alpha op(A) op(B) + beta C, column-major, op 'N' or 'T'.  Replaces the cublasSgemm calls of
 * DeviceMatrix<float>::Multiply / MultiplyBySelfTranspose (src/memory/DeviceMatrix.h:195-227). */
hipError_t gpumlib_nmf_sgemm(char transa, char transb, int M, int N, int K, float alpha, const float *A, int lda,
                             const float *B, int ldb, float beta, float *C, int ldc, hipStream_t stream);

/* Narrow-operand variants of gpumlib_nmf_sgemm (no reference counterpart; for callers that keep the
 * DeviceMatrix::Multiply operands in BF16 or FP8 to halve or quarter their bytes): the same contract with
 * A and B in BF16 (uint16 bit patterns) or in OCP E4M3 bytes with one E8M0 scale per operand (element value
 * e4m3(q) * 2^(scale - 127), scale in [0, 254]); C and the accumulation stay FP32.  Both run on the
 * matrix cores of their operand format.  Every product of two such values is exact in FP32: the BF16 form
 * is the FP32-accumulated GEMM of the stored values; the FP8 matrix cores sum each 64-term block of K at
 * less than FP32 accuracy (measured: up to ~6e-5 relative on full-range random operands). */
hipError_t gpumlib_nmf_gemm_bf16(char transa, char transb, int M, int N, int K, float alpha, const uint16_t *A,
                                 int lda, const uint16_t *B, int ldb, float beta, float *C, int ldc, hipStream_t stream);
hipError_t gpumlib_nmf_gemm_fp8(char transa, char transb, int M, int N, int K, float alpha, const uint8_t *A, int lda,
                                int scale_a, const uint8_t *B, int ldb, int scale_b, float beta, float *C, int ldc,
                                hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* 3. Fused iterations.  One call runs whole multiplicative-update           */
/*    iterations of NMF_MultiplicativeEuclidianDistance::DoIteration          */
/*    (nmf_multiplicative_euclidian.cpp:29-67) or                             */
/*    NMF_MultiplicativeDivergence::DoIteration                               */
/*    (nmf_multiplicative_divergence.cpp:29-50) with the same operation order */
/*    (H from the old W, then W from the new H; KL recomputes W H between).   */
/* ------------------------------------------------------------------------ */

/* Update rules.  The additive ones follow NMF_AdditiveEuclidian::DoIteration
 * (nmf_additive_euclidian.cpp:31-56) and NMF_AdditiveDivergence::DoIteration
 * (nmf_additive_divergence.cpp:30-51); they run in FP32 (precision GPUMLIB_NMF_FP32)
 * through gpumlib_nmf_iterate / the object API, not through the batch or the server.
 *   GPUMLIB_NMF_ADDITIVE_EUCLIDEAN: as the reference, H is left unchanged (its H step
 *     hands H to the multiplicative launcher as the numerator, :42, so only scratch is
 *     written); W <- max(0, W + (W / WHHt) (VHt - WHHt)).
 *   GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_WITH_H: also H <- max(0, H + (H / WtWH) (WtV - WtWH))
 *     (no reference counterpart: the H step the reference's scratch buffers were for).
 *   GPUMLIB_NMF_ADDITIVE_KL: H and W by the additive KL kernels, sums clamped at eps. */
enum {
    GPUMLIB_NMF_EUCLIDEAN = 0,
    GPUMLIB_NMF_KL = 1,
    GPUMLIB_NMF_ADDITIVE_EUCLIDEAN = 2,
    GPUMLIB_NMF_ADDITIVE_KL = 3,
    GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_WITH_H = 4
};
/* Storage precision of V (W, H and every update stay FP32).  GPUMLIB_NMF_FP32X: V in FP32, the pass
 * products formed on the BF16 matrix cores from three-way BF16 splits of both operands (six piece
 * products, FP32 accumulation): FP32 results, at FP32 storage traffic. */
enum { GPUMLIB_NMF_FP32 = 0, GPUMLIB_NMF_BF16 = 1, GPUMLIB_NMF_FP8 = 2, GPUMLIB_NMF_FP32X = 3 };
enum {
    GPUMLIB_NMF_OBJ_NONE = 0,
    GPUMLIB_NMF_OBJ_FROBENIUS = 1, /* sum (V - WH)^2 */
    GPUMLIB_NMF_OBJ_KL = 2,        /* sum V ln(V/WH) - V + WH, 0 ln 0 = 0 */
    GPUMLIB_NMF_OBJ_QUALITY = 3    /* reference quality, nmf_kernels_quality.cu:35: each term in FP32 as there,
                                      summed in FP64 */
};

/* V store: the device form of V that gpumlib_nmf_iterate reads for a precision (no reference
 * counterpart).  GPUMLIB_NMF_FP32: V itself (float, n x m column-major).  GPUMLIB_NMF_BF16: V (contraction
 * over i, m columns u) then V^T (contraction over u, n columns i, from a 256-byte boundary) in BF16 (round to
 * nearest even), each tiled for the pass that streams it: blocks of 64 columns x 256 contraction elements
 * (32 KB, column-major inside the block), element (k, col) at ((k/256 * T + col/64) * 64 + col%64) * 256 +
 * k%256 with T = ceil(columns/64), the contraction padded to a multiple of 256 and the columns to 64 T with
 * zeros; every chunk a pass workgroup reads is one contiguous 32 KB block.  GPUMLIB_NMF_FP8: V and V^T in OCP
 * E4M3, tiled the same way in blocks of 64 columns x 512 contraction bytes (element (k, col) at ((k/512 * T +
 * col/64) * 64 + col%64) * 512 + k%512, the contraction padded to a multiple of 512 and the columns to 64 T with
 * zeros), 256-byte aligned parts, followed by one E8M0 scale per 32 x 32 tile of the padded V (ldv = 512 *
 * ceil(n/512), ldt = 512 * ceil(m/512); scale of tile (i/32, u/32) at byte (u/32) * (ldv/32) + i/32; padding
 * tiles hold 127 = 2^0), shared by both orientations.  GPUMLIB_NMF_FP32X: V then V^T (from a 256-byte
 * boundary) in FP32, tiled the same way in blocks of 64 columns x 128 contraction elements (element (k, col) at
 * ((k/128 * T + col/64) * 64 + col%64) * 128 + k%128, the contraction padded to a multiple of 128 and the
 * columns to 64 T with zeros).  _pack builds the store from an FP32 device V on `stream`.  BF16, FP8 and
 * FP32X storage run the Euclidean rule; FP8 supports r <= 64, BF16 and FP32X r <= 128.  BF16 storage also
 * runs the multiplicative KL rule (GPUMLIB_NMF_KL, r <= 128, eps > 0: W H formed per block from BF16 copies
 * of W and H, the FP32 ratio rounded to BF16 for the second product, FP32 sums and updates). */
size_t gpumlib_nmf_vstore_bytes(int precision, int n, int m);
hipError_t gpumlib_nmf_vstore_pack(int precision, int n, int m, const float *V, void *vstore, size_t vstore_bytes,
                                   hipStream_t stream);

/* Cache policy of the BF16 / FP8 passes' reads of the V store (process-wide, read when an iteration or
 * pass starts; no reference counterpart).  GPUMLIB_NMF_STREAM_AUTO (default): for a BF16 store above
 * 384 MiB (16384 x 8192: 537 MB), the H pass's K slices within the first 192 MiB of V keep the default
 * policy and stay in the 256 MB last-level cache between iterations, and the rest of V and all of V^T
 * are read non-temporally so as not to evict them; smaller stores, and FP8 stores, are read with the
 * default policy (they are re-read from that cache).  _NONTEMPORAL / _CACHED force one policy for all of
 * V.  Values are unaffected: only the cache policy of the loads differs. */
#define GPUMLIB_NMF_STREAM_AUTO 0
#define GPUMLIB_NMF_STREAM_NONTEMPORAL 1
#define GPUMLIB_NMF_STREAM_CACHED 2
hipError_t gpumlib_nmf_set_stream_policy(int policy);

/* Device workspace needed by gpumlib_nmf_iterate / gpumlib_nmf_objective for one problem. */
size_t gpumlib_nmf_workspace_bytes(int rule, int precision, int n, int m, int r);

/* Runs `iterations` iterations in place on W and H (device pointers, FP32 for every precision).  V is
 * the V store of `precision` (GPUMLIB_NMF_FP32: the float n x m matrix itself).  When objective != OBJ_NONE the
 * objective of the state on entry is written, in FP64, to d_objective[0] (device memory); it is
 * fused into the first iteration's pass over V.  update_w = 0 updates H only (DoIteration(false)).
 * Ranks up to 128 (up to 64 with FP8 storage) run the fused passes.  Larger ranks run, with FP32 storage
 * only, the reference's own sequence of GEMMs over materialised operands (W^T V, W^T W, W^T W H, V H^T,
 * H H^T, W H H^T; KL: W H and the ratio V / (W H + eps) in an n x m buffer) on the library's MFMA GEMM;
 * their workspace then holds that n x m buffer. */
hipError_t gpumlib_nmf_iterate(int rule, int precision, int n, int m, int r, const void *V, float *W, float *H,
                               float eps, int iterations, int update_w, int objective, double *d_objective,
                               void *workspace, size_t workspace_bytes, hipStream_t stream);

/* No reference counterpart: for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  gpumlib_nmf_iterate with the
 * workspace taken from a static device array of GPUMLIB_NMF_NOALLOC_SCRATCH_BYTES (32 MiB) that the
 * library holds on every GPU it runs on until the process ends; the call allocates nothing.  Same
 * arguments, results and errors as gpumlib_nmf_iterate, plus hipErrorOutOfMemory when the problem's
 * workspace (gpumlib_nmf_workspace_bytes) exceeds the array: e.g. FP32 16384 x 8192 needs 16.2 MiB
 * at r = 64 and 32.6 MiB at r = 128 (rejected).  Every call on a GPU shares that GPU's array and
 * leaves its contents behind: calls that overlap on one GPU (other streams, other threads) race. */
#define GPUMLIB_NMF_NOALLOC_SCRATCH_BYTES (32u << 20)
hipError_t gpumlib_nmf_iterate_noalloc(int rule, int precision, int n, int m, int r, const void *V, float *W,
                                       float *H, float eps, int iterations, int update_w, int objective,
                                       double *d_objective, hipStream_t stream);

/* Objective of the current state, FP64, written to d_objective[0] (device). */
hipError_t gpumlib_nmf_objective(int rule, int precision, int n, int m, int r, const void *V, const float *W,
                                 const float *H, float eps, int objective, double *d_objective, void *workspace,
                                 size_t workspace_bytes, hipStream_t stream);

/* Runs one contraction pass of the fused iteration alone (which = 0: the pass over the rows of V that
 * feeds the H update; which = 1: the pass over the columns of V that feeds the W update), writing only
 * workspace partials; W and H are unchanged.  For per-kernel timing and profiling: exactly one kernel is
 * launched, configured as gpumlib_nmf_iterate launches it.  The workspace must have been used by a gpumlib_nmf_iterate call for the same problem (it
 * zeroes the arrival counters of the passes, which every pass leaves at zero; with BF16 / FP8 storage the
 * pass also reads the narrow copies of W and H that call left in `workspace`). */
hipError_t gpumlib_nmf_pass(int rule, int precision, int which, int n, int m, int r, const void *V, const float *W,
                            const float *H, float eps, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* 4. Handle over the C++ class API (NMF_MultiplicativeEuclidianDistance,     */
/*    NMF_MultiplicativeDivergence, NMF_AdditiveEuclidian,                    */
/*    NMF_AdditiveDivergence; constructed from host V, W, H as in the         */
/*    (v, w, h) constructors of nmf_multiplicative_euclidian.h:62,            */
/*    nmf_multiplicative_divergence.h:57, nmf_additive_euclidian.h:57 and     */
/*    nmf_additive_divergence.h:56).  rule: any GPUMLIB_NMF_* rule; precision */
/*    must be GPUMLIB_NMF_FP32.  For foreign-language callers.                */
/* ------------------------------------------------------------------------ */
typedef struct gpumlib_nmf_object *gpumlib_nmf_handle;

/* host_V (n x m), host_W (n x r), host_H (r x m) are column-major host arrays. */
hipError_t gpumlib_nmf_object_create(int rule, int precision, int n, int m, int r, const float *host_V,
                                     const float *host_W, const float *host_H, gpumlib_nmf_handle *out);
hipError_t gpumlib_nmf_object_set_epsilon(gpumlib_nmf_handle h, float eps);
hipError_t gpumlib_nmf_object_set_quality_interval(gpumlib_nmf_handle h, int interval);
hipError_t gpumlib_nmf_object_do_iteration(gpumlib_nmf_handle h, int update_w);        /* NMF::DoIteration, nmf_base.h:145 */
hipError_t gpumlib_nmf_object_do_iterations(gpumlib_nmf_handle h, int count, int update_w);
float gpumlib_nmf_object_quality_improvement(gpumlib_nmf_handle h);                    /* nmf_base.h:148-150 */
hipError_t gpumlib_nmf_object_get_w(gpumlib_nmf_handle h, float *host_W);              /* nmf_base.h:125-127 */
hipError_t gpumlib_nmf_object_get_h(gpumlib_nmf_handle h, float *host_H);              /* nmf_base.h:131-133 */
hipError_t gpumlib_nmf_object_get_wh(gpumlib_nmf_handle h, float *host_WH);            /* nmf_base.h:137-141 */
hipError_t gpumlib_nmf_object_objective(gpumlib_nmf_handle h, int objective, double *host_value);
void gpumlib_nmf_object_destroy(gpumlib_nmf_handle h);

/* ------------------------------------------------------------------------ */
/* 5. Batches of independent problems (no reference counterpart: the          */
/*    reference's only multi-problem workload, TrainUsingGroups in            */
/*    src/examples/FacesNMF/FacesNMFDlg.cpp:582-656, runs its problems one    */
/*    after another).  All registered problems (Euclidean or KL, FP32, r <=   */
/*    64, each with its own V, W, H and workspace) advance together with six  */
/*    grouped launches per iteration.  A problem's arithmetic depends only on */
/*    its own shape, so its results are bitwise the same alone and in any     */
/*    mix.  Objectives: the FP64 objective of the state on entry to           */
/*    iteration 0 (and 1) is written to output slot 2*slot (and 2*slot+1).    */
/* ------------------------------------------------------------------------ */
typedef struct gpumlib_nmf_batch *gpumlib_nmf_batch_handle;

hipError_t gpumlib_nmf_batch_create(int max_jobs, gpumlib_nmf_batch_handle *out); /* max_jobs <= 1024 */
void gpumlib_nmf_batch_destroy(gpumlib_nmf_batch_handle b);
/* Work items per problem and pass that the K split aims for (default 32); applies to problems added later.
 * Workspaces (gpumlib_nmf_batch_workspace_bytes) are sized for max(target, 128) by default -- room for
 * every candidate of the autotuner -- and for exactly this target after the call. */
hipError_t gpumlib_nmf_batch_set_target(gpumlib_nmf_batch_handle b, int workgroups_per_pass);
/* Engines: GPUMLIB_NMF_BATCH_FUSED (default): per problem and pass one work item per 128-column (H) or
 * 128-row (W) tile over the whole contraction, which applies its tile's update itself (the Euclidean
 * updates take W^T W and H H^T from a gram launch before each pass); four grouped launches per
 * iteration plus one per rule.  GPUMLIB_NMF_BATCH_SPLIT_K: the contraction split into K slices (the
 * target above), partials summed by separate update launches.  A switch re-plans every problem. */
enum { GPUMLIB_NMF_BATCH_SPLIT_K = 0, GPUMLIB_NMF_BATCH_FUSED = 1 };
hipError_t gpumlib_nmf_batch_set_engine(gpumlib_nmf_batch_handle b, int engine);
int gpumlib_nmf_batch_engine(gpumlib_nmf_batch_handle b);
/* Rank blocking of the grouped pass kernels, for every shape bucket: 0 (default): every problem on the
 * kernels built for two rank blocks of 32 (r <= 64; a problem with r <= 32 skips the second block); 1:
 * problems with r <= 32 on kernels built for one block (half the accumulator registers and operand LDS),
 * launched after the two-block kernels of each pass.  Results are the same either way (same arithmetic,
 * same order).  gpumlib_nmf_batch_rank_split returns 1 when any bucket splits (the autotuner chooses
 * per bucket). */
hipError_t gpumlib_nmf_batch_set_rank_split(gpumlib_nmf_batch_handle b, int on);
int gpumlib_nmf_batch_rank_split(gpumlib_nmf_batch_handle b);
/* Tile width of the grouped pass kernels, for every shape bucket: 128 (default) or 64 columns of V per
 * H-pass work item and rows per W-pass work item.  A 64-wide work item has two waves on each 32-column
 * block, each over half of every K chunk, their partials added after the loop: twice the work items
 * for problems with few columns or rows (results within FP32 rounding of the 128-wide ones, not
 * bit-identical).  Re-plans every problem; not under graph capture. */
hipError_t gpumlib_nmf_batch_set_tile_cols(gpumlib_nmf_batch_handle b, int cols);
size_t gpumlib_nmf_batch_workspace_bytes(gpumlib_nmf_batch_handle b, int rule, int n, int m, int r);
hipError_t gpumlib_nmf_batch_add(gpumlib_nmf_batch_handle b, int rule, int n, int m, int r, const float *V, float *W,
                                 float *H, float eps, void *workspace, size_t workspace_bytes, int *slot);
hipError_t gpumlib_nmf_batch_remove(gpumlib_nmf_batch_handle b, int slot);
int gpumlib_nmf_batch_active(gpumlib_nmf_batch_handle b);
/* objectives in {0, 1, 2}, at most iterations. */
hipError_t gpumlib_nmf_batch_iterate(gpumlib_nmf_batch_handle b, int iterations, int objectives, hipStream_t stream);
/* The same work replayed from a hipGraph captured on the first call for (iterations, objectives);
 * `stream` must not be the null stream.  The graph stays valid when problems are added or removed. */
hipError_t gpumlib_nmf_batch_graph_launch(gpumlib_nmf_batch_handle b, int iterations, int objectives,
                                          hipStream_t stream);
/* Copies the 2 * max_jobs objective slots to the host (synchronises `stream`). */
hipError_t gpumlib_nmf_batch_objectives(gpumlib_nmf_batch_handle b, double *host_out, hipStream_t stream);
const double *gpumlib_nmf_batch_objective_device_ptr(gpumlib_nmf_batch_handle b);

/* Autotuner: times `trial_iterations` iterations of the active mix (after one untimed iteration) for
 * each candidate configuration -- the tile shape along K: the fused engine (a tile's whole contraction
 * in one work item) or the split-K engine with a K-split target per problem (4, 8, 16, 32, 64 or 128
 * workgroups per problem and pass, those that fit every problem's workspace), first one target for
 * all problems, then, split-K, one step up and down from the best for each shape bucket of problems
 * on its own (buckets of fewer than 2^18 elements of V, fewer than 2^21, and the rest); then the tile
 * width (gpumlib_nmf_batch_set_tile_cols) per bucket; then the persistent grid sizes of the grouped
 * passes (512, 1024, 2048) and of the updates or, fused, the gram launches (256, 512, 1024); then the
 * rank blocking (gpumlib_nmf_batch_set_rank_split) per bucket with problems of r <= 32 -- and keeps
 * the fastest.  W and H of every problem are saved before and restored
 * after each trial (a temporary device buffer of their size is allocated for that), so no problem's
 * state changes; a re-planned problem sums its K slices in the new grouping afterwards (results within
 * the FP32 tolerance of the untuned ones, not bit-identical to them).  The winner is cached per mix
 * signature (the multiset of rule, n, m, r): a later call for the same mix applies it without trials.
 * Synchronises the stream; not under graph capture.  No reference counterpart. */
typedef struct {
    int engine;                     /* GPUMLIB_NMF_BATCH_FUSED or GPUMLIB_NMF_BATCH_SPLIT_K */
    int target_workgroups, pass_grid, update_grid;
    double ms_per_iteration;        /* of the chosen configuration */
    double ms_per_iteration_before; /* of the configuration in effect before the call */
    int trials;                     /* configurations timed (0 when a cached choice was applied) */
    int cached;
    int split_ranks;                /* 1: some bucket present splits ranks (gpumlib_nmf_batch_set_rank_split) */
    /* per shape bucket (0 in all three for a bucket with no problem); target_workgroups is that of
       the largest bucket present */
    int bucket_targets[3];          /* K-split target */
    int bucket_tile_cols[3];        /* tile width, 128 or 64 */
    int bucket_split_ranks[3];      /* rank blocking */
} gpumlib_nmf_batch_tuning;
hipError_t gpumlib_nmf_batch_autotune(gpumlib_nmf_batch_handle b, int trial_iterations, hipStream_t stream,
                                      gpumlib_nmf_batch_tuning *out);

/* ------------------------------------------------------------------------ */
/* 6. Continuous-batching server (no reference counterpart).  Any host thread */
/*    submits problems (Euclidean or KL, FP32, r <= 64; inputs are copied).   */
/*    An engine thread admits a queued problem when a batch slot is free and  */
/*    its HBM footprint (V, W, H, workspace) fits in the device pool, advances */
/*    all admitted problems round_iterations iterations per round with the    */
/*    grouped launches (optionally replayed from a hipGraph), and retires a   */
/*    problem when its relative objective change over one iteration, measured */
/*    each round, is below its tolerance or it reached max_iterations.        */
/*    Problems advance in whole rounds, so a cap that is not a multiple of    */
/*    round_iterations is reached at the end of the round that crosses it     */
/*    (cap 15, rounds of 10: 20 iterations); the result reports the count run. */
/*    Times are seconds since the server was created.                          */
/* ------------------------------------------------------------------------ */
typedef struct gpumlib_nmf_server *gpumlib_nmf_server_handle;
typedef struct {
    size_t hbm_budget_bytes; /* device pool for resident problems */
    int max_active;          /* batch slots, <= 1024 */
    int round_iterations;    /* iterations per scheduling round */
    int use_graph;           /* 1: replay each round from a captured hipGraph */
    int target_workgroups;   /* K-split target per problem and pass, 0 = default */
    int autotune;            /* > 0: autotune the batch (this many timed iterations per candidate) whenever
                                the number of resident problems has doubled or halved; 0: off */
    int split_k;             /* 0: the fused batch engine (default); 1: the split-K engine (both as
                                gpumlib_nmf_batch_set_engine describes; the autotuner may switch) */
    int sequential;          /* 1: one problem at a time (max_active is taken as 1), each run to its end
                                through gpumlib_nmf_iterate on its pool memory with the same rounds and
                                stopping rule -- the unbatched comparison point; 0: continuous batching */
} gpumlib_nmf_server_config;
typedef struct {
    int iterations;
    int converged;
    double objective; /* FP64 objective after the last measured iteration */
    double submit_s, admit_s, finish_s;
} gpumlib_nmf_job_result;
typedef struct {
    long long submitted, completed, rounds, iterations_batched;
    int peak_active;
    size_t pool_bytes, pool_high_water;
    long long autotunes;                                    /* autotuner calls made */
    int tuned_target, tuned_pass_grid, tuned_update_grid;  /* configuration chosen last */
    int tuned_engine;                                       /* GPUMLIB_NMF_BATCH_FUSED or _SPLIT_K */
    int tuned_split_ranks;                                  /* rank blocking chosen last */
    int tuned_bucket_targets[3];                            /* per shape bucket, chosen last */
    int tuned_bucket_tile_cols[3];
    int tuned_bucket_split_ranks[3];
} gpumlib_nmf_server_stats;

hipError_t gpumlib_nmf_server_create(const gpumlib_nmf_server_config *cfg, gpumlib_nmf_server_handle *out);
hipError_t gpumlib_nmf_server_submit(gpumlib_nmf_server_handle s, int rule, int n, int m, int r, const float *V,
                                     const float *W, const float *H, float eps, double tol, int max_iterations,
                                     long long *job_id);
/* Blocks until the problem is retired; copies W (n x r) and H (r x m) out when the pointers are non-null. */
hipError_t gpumlib_nmf_server_wait(gpumlib_nmf_server_handle s, long long job_id, float *W_out, float *H_out,
                                   gpumlib_nmf_job_result *result);
hipError_t gpumlib_nmf_server_stats_get(gpumlib_nmf_server_handle s, gpumlib_nmf_server_stats *out);
/* Finishes every submitted problem, then releases the server. */
void gpumlib_nmf_server_destroy(gpumlib_nmf_server_handle s);

/* ------------------------------------------------------------------------ */
/* 7. Seeded host-side input generators (no reference counterpart: the        */
/*    reference draws W and H with rand(), nmf_base.h:56-61).  Counter-based  */
/*    SplitMix64: value k of stream s under seed is                           */
/*      mix(mix(seed ^ (s * 0xD1B54A32D192ED03)) + k)                         */
/*    uniform float = (x >> 40) * 2^-24,  uniform double = (x >> 11) * 2^-53. */
/* ------------------------------------------------------------------------ */
uint64_t gpumlib_nmf_rng_u64(uint64_t seed, uint64_t stream, uint64_t k);
void gpumlib_nmf_gen_uniform_f32(uint64_t seed, uint64_t stream, size_t count, float *out);
/* V (n x m) = W0 (n x r0) H0 (r0 x m) + noise * U[0,1),  W0, H0 ~ U[0,1)  (streams 1, 2, 3). */
void gpumlib_nmf_gen_euclidean_v(uint64_t seed, int n, int m, int r0, float noise, float *V);
/* V (n x m) ~ Poisson(W0 H0), W0, H0 ~ Exp(1)  (streams 1, 2, 6). */
void gpumlib_nmf_gen_poisson_v(uint64_t seed, int n, int m, int r0, float *V);
/* Initial W (n x r) and H (r x m) ~ U[0,1)  (streams 4, 5). */
void gpumlib_nmf_gen_init(uint64_t seed, int n, int m, int r, float *W, float *H);

/* Host form of the FP8 storage conversion (the device code uses the same routines): OCP E4M3 of
 * x / 2^(scale-127), round to nearest even, saturating at 448; the E8M0 scale for a block maximum;
 * and the decoded value q * 2^(scale-127). */
void gpumlib_nmf_fp8_quantize(const float *x, size_t count, int scale, uint8_t *out);
int gpumlib_nmf_fp8_scale_for(float amax);
float gpumlib_nmf_fp8_decode(uint8_t q, int scale);

#ifdef __cplusplus
}
#endif

#endif /* GPUMLIB_NMF_H */
