// Workspace planning for the fused iteration: slice counts of the two passes and
// the byte layout of the caller-provided workspace.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "nmf_common.hpp"

namespace gpumlib_nmf {

// eps of the reference: SMALL_VALUE_TO_ADD_DENOMINATOR = CUDA_VALUE(10e-9) = 1e-8f
// (src/cuda/nmf/nmf_kernels.h:34).
constexpr float kReferenceEpsilon = 10e-9f;
// Fused kernels hold up to 4 MFMA row blocks of 32 rank rows.  Larger ranks (FP32 storage)
// run the unfused iteration: the reference's GEMM sequence over materialised operands.
constexpr int kMaxFusedRank = 128;
// Workgroups per pass aimed for: two resident workgroups on each of 256 CUs.
constexpr int kTargetWorkgroups = 512;
// V stores (both orientations) larger than this are read partly with the non-temporal cache
// policy by the BF16 passes: 1.5 x the MI355X's 256 MB last-level cache.  The H pass's K slices
// that start within the first kStreamCacheBytes of V keep the default policy (that part stays in
// the cache from one iteration to the next); the other slices and all of V^T are streamed
// non-temporally, so that they do not evict it.  Smaller stores are read with the default policy
// throughout (they are re-read from the cache).  At 16384 x 8192 (V 268 MB per orientation, H
// pass in 4 slices) the iteration measured 121.6 us with everything non-temporal, 113.0 with the
// first 2 slices cached, 112.2 with 3 (this budget), 116.1 with all of V.
constexpr size_t kStreamNtBytes = 384ull << 20;
constexpr size_t kStreamCacheBytes = 192ull << 20;
// Work items aimed for by the objective pass on split factors.
constexpr int kObjectiveWorkgroups = 512;
// Padded rank of the split factors: whole MFMA steps of 16, in the kernel's four register
// shapes (16, 32, 64, 128).
__host__ __device__ inline int split_rank(int r) { return r <= 16 ? 16 : (r <= 32 ? 32 : (r <= 64 ? 64 : 128)); }

// Update rules (GPUMLIB_NMF_* rule constants of the C-ABI): which passes a rule runs
// (the KL ones or the Euclidean ones) and whether its update is the additive form.
inline bool rule_is_kl(int rule) { return rule == 1 /* KL */ || rule == 3 /* ADDITIVE_KL */; }
inline bool rule_is_additive(int rule) { return rule >= 2; }

struct Plan {
    int rule, precision, n, m, r;
    int unfused;          // 1: r > kMaxFusedRank, the unfused iteration (offsets off_u*)
    int npad, rpad, ldr, ldp;
    int ldv8, ldt8;       // BF16 storage: padded contraction lengths (bf16_ld of n and of m): the K extent
                          // of the tiled V and V^T and the leading dimensions of the W / H^T shadows
    int ldx_n, ldx_m;     // FP32X storage: padded contraction lengths (x3_ld of n and of m), as ldv8 / ldt8
    int lp_nt;            // BF16 / FP32X passes: 64-column blocks per workgroup (1 or 2)
    int nt_k0[2];         // BF16 / FP8 passes (H pass over n, W pass over m): V chunks from this
                          // contraction index on are read with the non-temporal policy (n or m: none)
    int tiles1, S1, ks1;  // H pass: column tiles (+ gram tile), K slices over n
    int tiles3, S3, ks3;  // W pass: row tiles (+ gram tile), K slices over m
    long pstride1, pstride3;
    int nobjp;
    int ld16n, ld16m, nbi, nbu;  // FP8 storage: leading dimensions (fp8_ld of n, m) and their 32-blocks
    // objective pass on split factors (nmf_objective.hip): padded rank, rows and columns of the
    // pieces, 128-row blocks, 32-column blocks and column ranges per row block
    int ob_rp, ob_nrows, ob_ncols, ob_rblocks, ob_cblocks, ob_csplit;
    size_t off_P, off_Psum, off_objp, off_G, off_cnt, off_Wb, off_Hb, off_Wsc, off_Hsc, off_Ws, off_Hs, bytes;
    // KL rule with BF16 storage: rank-contiguous BF16 shadows of W (ldv8 rows) and of H (ldt8 rows),
    // rows of rps = r rounded up to 32 elements (off_Wb / off_Hb hold the K-contiguous ones)
    int rps;
    size_t off_Wr, off_Hr;
    // unfused iteration: WH (n x m; the KL ratio in place), numerator and denominator products
    // (max(r m, n r) floats each), the r x r gram, the clamped sums (r), FP64 objective partials
    size_t off_uWH, off_uA, off_uB, off_uG, off_usum, off_uobj;
};

// Padded contraction lengths of the narrow V stores and leading dimensions of the operand
// shadows: whole pass chunks (256 BF16 or 512 FP8 elements along the contraction), the
// padding zero-filled, so the streaming passes load every chunk without bounds checks.
inline int bf16_ld(int k) { return (k + 255) / 256 * 256; }
inline int fp8_ld(int k) { return (k + 511) / 512 * 512; }
inline int x3_ld(int k) { return (k + 127) / 128 * 128; }

// Byte layout of the FP8 V store: Vq (contraction i, m columns) | Vtq (contraction u, n columns),
// each tiled in blocks of 64 columns x 512 contraction bytes (fp8_tiled_index; contraction padded
// to fp8_ld, columns to 64 ceil(cols/64), zeros) | tile scales (fp8_ld(m)/32 x fp8_ld(n)/32,
// padding tiles 127 = 2^0).
struct Fp8Store {
    size_t off_t, off_sc, bytes;
};
Fp8Store fp8_store_layout(int n, int m);

// BF16 V store: V (contraction over n, m columns), then V^T (contraction over m, n columns)
// from a 256-byte boundary.  Each orientation is tiled for the pass that streams it: blocks
// of 64 columns x 256 contraction elements (32 KB, column-major inside the block: 512
// contiguous bytes per column), block (chunk c = k / 256, column tile t = col / 64) at block
// index c * ntiles + t, ntiles = ceil(columns / 64); padding (k up to bf16_ld(K), columns
// up to 64 ntiles) zero.  One pass chunk of a column tile is one contiguous 32 KB read.
// (bf16_tiled_index, the element offset, is in nmf_common.hpp for the kernels.)
inline int lp_tiles(int cols) { return (cols + 63) / 64; }
inline size_t bf16_tiled_elems(int K, int cols) { return (size_t)bf16_ld(K) * 64 * lp_tiles(cols); }
size_t vstore_bf16_offset_t(int n, int m);
// FP32X V store: V then V^T (from a 256-byte boundary) in FP32, each tiled the same way in blocks
// of 64 columns x 128 contraction elements (x3_tiled_index; contraction padded to x3_ld).
inline size_t x3_tiled_elems(int K, int cols) { return (size_t)x3_ld(K) * 64 * lp_tiles(cols); }
size_t vstore_x3_offset_t(int n, int m);
size_t vstore_bytes(int precision, int n, int m);

// target: workgroups aimed for per pass (K slices are chosen to reach it).  tile: columns (H
// pass) and rows (W pass) per work item of the FP32 passes, 128 or 64 (the grouped launches
// have both widths; the single-problem launches take 128).
bool make_plan(int rule, int precision, int n, int m, int r, Plan* pl, int target = kTargetWorkgroups,
               int tile = TILE);
void pass_args(const Plan& pl, const float* V, const float* W, const float* H, void* ws, float eps, int which,
               PassArgs* a);
hipError_t run_iterations(const Plan& pl, const float* V, float* W, float* H, float eps, int iterations, int update_w,
                          int objective, double* d_objective, void* ws, hipStream_t st);
hipError_t run_objective(const Plan& pl, const void* V, const float* W, const float* H, float eps, int objective,
                         double* d_objective, void* ws, hipStream_t st);
hipError_t run_iterations_unfused(const Plan& pl, const float* V, float* W, float* H, float eps, int iterations,
                                  int update_w, int objective, double* d_objective, void* ws, hipStream_t st);
hipError_t run_objective_unfused(const Plan& pl, const float* V, const float* W, const float* H, float eps,
                                 int objective, double* d_objective, void* ws, hipStream_t st);
hipError_t run_iterations_fp8(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                              int update_w, int objective, double* d_objective, void* ws, hipStream_t st);
hipError_t run_iterations_bf16(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                               int update_w, int objective, double* d_objective, void* ws, hipStream_t st);
hipError_t run_iterations_kl_bf16(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                                  int update_w, int objective, double* d_objective, void* ws, hipStream_t st);
void kl_bf16_args(const Plan& pl, const void* Vstore, void* ws, float eps, int which, KlLpArgs* a);
hipError_t run_iterations_x3(const Plan& pl, const void* Vstore, float* W, float* H, float eps, int iterations,
                             int update_w, int objective, double* d_objective, void* ws, hipStream_t st);

}  // namespace gpumlib_nmf
