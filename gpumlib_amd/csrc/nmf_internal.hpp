// Internal launcher declarations shared by the kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "nmf_common.hpp"

namespace gpumlib_nmf {

// Slice grams a row-form update sums itself while staging the gram (register-staged: r <= 64).
constexpr int kMaxGramSlices = 8;

// Pass kernels (nmf_pass_f32.hip).
// H pass (KL: with any objective of the entry state; Euclidean: without)
hipError_t launch_pass_h_f32(const PassArgs& a, int kl, int obj, hipStream_t st);
hipError_t launch_pass_w_f32(const PassArgs& a, int kl, hipStream_t st);

// BF16 storage path (nmf_pass_bf16.hip).
hipError_t launch_pass_tn_bf16(const LpPassArgs& a, hipStream_t st);
hipError_t launch_pack_v_bf16(const float* V, int n, int m, uint16_t* Vb, int ldv, uint16_t* Vt, int ldt,
                              hipStream_t st);
hipError_t launch_shadow_bf16(const float* W, const float* H, int n, int m, int r, uint16_t* Wb, int ldw, uint16_t* Ht,
                              int ldh, hipStream_t st);
// KL rule with BF16 storage (nmf_kl_bf16.hip; r <= 128, eps > 0, rps = kl_rank_pitch(r), kslice
// a multiple of 128, nt = 2).
hipError_t launch_pass_kl_bf16(const KlLpArgs& a, hipStream_t st);
// Xr[j*rps + a] = bf16(X[j*xs_j + a*xs_a]) for j < count, a < r; zeros up to rows x rps.
hipError_t launch_shadow_rank_bf16(const float* X, long xs_j, long xs_a, int count, int r, int rows, uint16_t* Xr,
                                   int rps, hipStream_t st);
// X(j,a) *= sum_s P[s][j*ldr + a] / max(sum_s Psum[s*rpad + a], eps), X(j,a) at X[j*xs_j + a*xs_a];
// then Xk[a*ldk + j] and Xr[j*rps + a] = bf16 of the new X.
hipError_t launch_update_kl_rows_bf16(float* X, long xs_j, long xs_a, const float* P, const float* Psum, int count,
                                      int r, int ldr, int rpad, int S, long pstride, float eps, uint16_t* Xk, int ldk,
                                      uint16_t* Xr, int rps, hipStream_t st);
// FP32X storage path (nmf_pass_x3.hip): FP32 V and V^T tiled by x3_tiled_index (contraction
// padded to ldv / ldt, multiples of 128), the A operand as three BF16 piece planes.
hipError_t launch_pass_tn_x3(const LpPassArgs& a, hipStream_t st);
hipError_t launch_pack_v_x3(const float* V, int n, int m, float* Vx, int ldv, float* Vtx, int ldt, hipStream_t st);
hipError_t launch_shadow_x3(const float* W, const float* H, int n, int m, int r, uint16_t* Wp, int ldw, uint16_t* Hp,
                            int ldh, hipStream_t st);
// FP8 storage path (nmf_pass_fp8.hip).
hipError_t launch_pass_tn_fp8(const Fp8PassArgs& a, hipStream_t st);
hipError_t launch_pack_v_fp8(const float* V, int n, int m, uint8_t* Vq, int ldv, uint8_t* Vtq, int ldt, uint8_t* Vsc,
                             hipStream_t st);
hipError_t launch_shadow_fp8(const float* X, int n, int m, int r, int which, uint8_t* Xq, int ldq, uint8_t* Xsc,
                             hipStream_t st);
// Objective pass on split factors (nmf_objective.hip): W, H -> their three BF16 pieces
// (rp = split_rank(r), nrows a multiple of 128 >= n, ncols a multiple of 32 >= m),
// then the per-work-item FP64 partials of the objective over V (vtype 0 FP32, 1 BF16, 2 FP8,
// 3 the tiled FP32 store of the FP32X path).
hipError_t launch_split_wh(const float* W, const float* H, int n, int m, int r, int rp, int nrows, int ncols,
                           uint16_t* Ws, uint16_t* Hs, hipStream_t st);
hipError_t launch_objective_split(const SplitObjArgs& a, int vtype, int obj, hipStream_t st);

// Update kernels (nmf_update.hip).
// Euclidean H:  H(a,u) *= WtV(a,u) / (sum_b WtW(a,b) H(b,u) + eps)
// Hs (nullable): BF16 copy of the new H written as H^T column-major (ld ldhs); with shplane > 0
// (FP32X), its three BF16 pieces in planes shplane elements apart (ldhs a multiple of 128).
// G: the reduced gram W^T W written by the H pass, WtW(a,b) = G[b*ldr + a];  with
// gsl in [1, kMaxGramSlices] (r <= 64), the first of the gsl slice grams of the pass (stride pstride),
// which the update sums in slice order.
// add != 0 (every Euclidean / KL update below): the additive-update rule of
// NMF_AdditiveEuclidian / NMF_AdditiveDivergence in place of the multiplicative one.
hipError_t launch_update_h_euc(float* H, const float* P, const float* G, int m, int r, int ldr, int S, long pstride,
                               float eps, uint16_t* Hs, int ldhs, hipStream_t st, int add = 0, int gsl = 0,
                               long shplane = 0);
// Euclidean W:  W(i,a) *= VHt(i,a) / (sum_b W(i,b) HHt(b,a) + eps)
// G: the reduced gram H H^T written by the W pass, HHt(a,b) = G[b*rpad + a].
hipError_t launch_update_w_euc(float* W, const float* P, const float* G, int n, int r, int ldp, int S, long pstride,
                               float eps, hipStream_t st, int add = 0);
// Euclidean W from partials laid out P[s][i*ldr + a] (gram rows at i >= n), as the
// BF16 pass over V^T writes them; Wb (nullable): BF16 copy of the new W (ld ldwb).
hipError_t launch_update_w_euc_t(float* W, const float* P, const float* G, int n, int r, int ldr, int S, long pstride,
                                 float eps, uint16_t* Wb, int ldwb, hipStream_t st, int add = 0, int gsl = 0,
                                 long shplane = 0);
// FP8 storage path: the same updates, also writing the E4M3 shadow of the new factor
// (one E8M0 scale per 32 rows along the contraction).
hipError_t launch_update_h_euc_q8(float* H, const float* P, const float* G, int m, int r, int ldr, int S,
                                  long pstride, float eps, uint8_t* Hq, int ldq, uint8_t* Hsc, hipStream_t st,
                                  int gsl = 0);
hipError_t launch_update_w_euc_t_q8(float* W, const float* P, const float* G, int n, int r, int ldr, int S,
                                    long pstride, float eps, uint8_t* Wq, int ldq, uint8_t* Wsc, hipStream_t st,
                                    int gsl = 0);
// Zeroes n ints (arrival counters) with a kernel, so that it is safe inside captured graphs.
hipError_t launch_zero_i32(int* p, int n, hipStream_t st);
// KL H:  H(a,u) *= num(a,u) / max(sum_s Psum[s][a], eps)
hipError_t launch_update_h_kl(float* H, const float* P, const float* Psum, int m, int r, int ldr,
                              int S, long pstride, float eps, hipStream_t st, int add = 0);
// KL W:  W(i,a) *= num(i,a) / max(sum_s Psum[s][a], eps)
hipError_t launch_update_w_kl(float* W, const float* P, const float* Psum, int n, int r, int ldp,
                              int S, long pstride, float eps, hipStream_t st, int add = 0);
// Deterministic sum of count doubles into out[0].
hipError_t launch_reduce_f64(const double* in, int count, double* out, hipStream_t st);

// Grouped launches over many problems (nmf_pass_f32.hip, nmf_update.hip).  grid
// is the persistent workgroup count; each workgroup strides over the table's items.
// mb: rank blocks of 32 the grouped pass kernels are built for (1: the table's problems have
// r <= 32; 2: r <= 64)
// Grouped passes: mb rank blocks of 32 (1 or 2), tw columns (H pass) / rows (W pass) per work
// item (128 or 64), as the table's problems were planned.
hipError_t launch_group_pass_h(const JobArgs* jobs, const int* table, int kl, int obj, int mb, int tw, int grid,
                               hipStream_t st);
hipError_t launch_group_pass_w(const JobArgs* jobs, const int* table, int kl, int mb, int tw, int grid,
                               hipStream_t st);
hipError_t launch_group_update_h(const JobArgs* jobs, const int* table, int grid, hipStream_t st);
hipError_t launch_group_update_w(const JobArgs* jobs, const int* table, int grid, hipStream_t st);
// Fused full-K grouped launches (one slice per problem; each item updates its own tile):
// the gram of W (which 0, into h.G) or of H (which 1, into w.G) per Euclidean problem, and
// the fused H / W passes.
hipError_t launch_group_gram(const JobArgs* jobs, const int* table, int which, int grid, hipStream_t st);
hipError_t launch_group_fused_h(const JobArgs* jobs, const int* table, int kl, int obj, int mb, int tw, int grid,
                                hipStream_t st);
hipError_t launch_group_fused_w(const JobArgs* jobs, const int* table, int kl, int mb, int tw, int grid,
                                hipStream_t st);
hipError_t launch_group_reduce_obj(const JobArgs* jobs, const int* table, int slot, int grid, hipStream_t st);

// Reference-shaped kernels over materialised operands.
hipError_t launch_mu_update(const float* nm, const float* dm, float* mm, long elements, float eps,
                            hipStream_t st);
hipError_t launch_add_update(float* X, const float* d1, const float* d2, long elements, hipStream_t st);
hipError_t launch_sum_w(const float* W, int n, int r, float eps, float* out, hipStream_t st);
hipError_t launch_sum_h(const float* H, int r, int m, float eps, float* out, hipStream_t st);
// out_f32 non-null: the total is written there as a float instead of to out.
hipError_t launch_quality(const float* V, const float* WH, long count, float eps, double* partials,
                          int npartials, double* out, hipStream_t st, float* out_f32 = nullptr);
hipError_t launch_kl_update_h_wh(float* H, const float* W, const float* V, const float* WH,
                                 const float* sumW, int n, int m, int r, float eps, hipStream_t st, int add = 0);
hipError_t launch_kl_update_w_wh(float* W, const float* H, const float* V, const float* WH,
                                 const float* sumH, int n, int m, int r, float eps, hipStream_t st, int add = 0);

// Unfused iteration (ranks above kMaxFusedRank; nmf_unfused.hip).
// Q = V / (WH + eps) (minus 1 when add != 0), in place over WH.
hipError_t launch_kl_ratio(const float* V, float* WH, long count, float eps, int add, hipStream_t st);
// X[k] = X[k] num[k] / sum[a(k)] (add: the additive KL rule), a(k) = k % rows when rank_is_row
// (H, r x m), else k / rows (W, n x r).
hipError_t launch_kl_apply(float* X, const float* num, const float* sum, long count, int rows, int rank_is_row,
                           int add, hipStream_t st);
// FP64 objective (OBJ_FROB / OBJ_KL / OBJ_QUALITY) of V against a materialised WH into out[0];
// partials holds objective_wh_partials(count) doubles.
int objective_wh_partials(long count);
hipError_t launch_objective_wh(const float* V, const float* WH, long count, float eps, int obj, double* partials,
                               double* out, hipStream_t st);

// General FP32 GEMM, column-major, cuBLAS-style operation flags (nmf_gemm.hip).
hipError_t launch_sgemm(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda,
                        const float* B, int ldb, float beta, float* C, int ldc, hipStream_t st);
// The same with operands in format fmt (0 FP32, 1 BF16 bit patterns, 2 OCP E4M3 times the
// operand's E8M0 scale 2^(sa-127) / 2^(sb-127), sa, sb in [0, 254]; ignored for 0 and 1).
hipError_t launch_gemm(int fmt, char ta, char tb, int M, int N, int K, float alpha, const void* A, int lda, int sa,
                       const void* B, int ldb, int sb, float beta, float* C, int ldc, hipStream_t st);

}  // namespace gpumlib_nmf
