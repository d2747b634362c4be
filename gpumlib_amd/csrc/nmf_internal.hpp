// Internal launcher declarations shared by the kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "nmf_common.hpp"

namespace gpumlib_nmf {

// Pass kernels (nmf_pass_f32.hip).
hipError_t launch_pass_h_f32(const PassArgs& a, int kl, int obj, int gemm2, hipStream_t st);
hipError_t launch_pass_w_f32(const PassArgs& a, int kl, hipStream_t st);

// Update kernels (nmf_update.hip).
// Euclidean H:  H(a,u) *= WtV(a,u) / (sum_b WtW(a,b) H(b,u) + eps)
hipError_t launch_update_h_euc(float* H, const float* P, int m, int r, int ldr, int S, long pstride,
                               float eps, hipStream_t st);
// Euclidean W:  W(i,a) *= VHt(i,a) / (sum_b W(i,b) HHt(b,a) + eps)
hipError_t launch_update_w_euc(float* W, const float* P, int n, int r, int ldp, int npad, int S,
                               long pstride, float eps, hipStream_t st);
// KL H:  H(a,u) *= num(a,u) / max(sum_s Psum[s][a], eps)
hipError_t launch_update_h_kl(float* H, const float* P, const float* Psum, int m, int r, int ldr,
                              int S, long pstride, float eps, hipStream_t st);
// KL W:  W(i,a) *= num(i,a) / max(sum_s Psum[s][a], eps)
hipError_t launch_update_w_kl(float* W, const float* P, const float* Psum, int n, int r, int ldp,
                              int S, long pstride, float eps, hipStream_t st);
// Deterministic sum of count doubles into out[0].
hipError_t launch_reduce_f64(const double* in, int count, double* out, hipStream_t st);

// Reference-shaped kernels over materialised operands.
hipError_t launch_mu_update(const float* nm, const float* dm, float* mm, long elements, float eps,
                            hipStream_t st);
hipError_t launch_sum_w(const float* W, int n, int r, float eps, float* out, hipStream_t st);
hipError_t launch_sum_h(const float* H, int r, int m, float eps, float* out, hipStream_t st);
hipError_t launch_quality(const float* V, const float* WH, long count, float eps, double* partials,
                          int npartials, double* out, hipStream_t st);
hipError_t launch_kl_update_h_wh(float* H, const float* W, const float* V, const float* WH,
                                 const float* sumW, int n, int m, int r, float eps, hipStream_t st);
hipError_t launch_kl_update_w_wh(float* W, const float* H, const float* V, const float* WH,
                                 const float* sumH, int n, int m, int r, float eps, hipStream_t st);

// General FP32 GEMM, column-major, cuBLAS-style operation flags (nmf_gemm.hip).
hipError_t launch_sgemm(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda,
                        const float* B, int ldb, float beta, float* C, int ldc, hipStream_t st);

}  // namespace gpumlib_nmf
