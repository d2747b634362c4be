// Kernels of the unfused iteration (ranks above what the fused passes hold): the
// reference's own sequence of GEMMs over materialised operands, with the GEMMs on
// the library's MFMA GEMM and these element-wise / reduction kernels around them.
//
//   KL (nmf_multiplicative_divergence.cpp:29-50, nmf_additive_divergence.cpp:30-51):
//     Q = V / (WH + eps)  [- 1 for the additive rule]   (nmf_kernels.cu:271,319)
//     H(a,u) = rule(H(a,u), (W^T Q)(a,u), sumW[a])       (nmf_kernels.cu:289-291)
//     W(i,a) = rule(W(i,a), (Q H^T)(i,a), sumH[a])       (nmf_kernels.cu:337-339)
//   Objectives of a materialised WH, FP64, many workgroups, fixed-order combine.
#include "nmf_common.hpp"
#include "nmf_internal.hpp"

namespace gpumlib_nmf {

namespace {

constexpr int kMaxObjPartials = 1024;

// Q[k] = V[k] / (WH[k] + eps) - sub, in place over WH.
__global__ __launch_bounds__(NTHREADS) void k_kl_ratio(const float* __restrict__ V, float* __restrict__ WH,
                                                       long count, float eps, float sub) {
    for (long k = (long)blockIdx.x * NTHREADS + threadIdx.x; k < count; k += (long)gridDim.x * NTHREADS)
        WH[k] = V[k] / (WH[k] + eps) - sub;
}

// X[k] = rule(X[k], num[k], sum[a(k)]) over a column-major factor: a(k) = k % rows for
// H (r x m, rows = r), a(k) = k / rows for W (n x r, rows = n).
__global__ __launch_bounds__(NTHREADS) void k_kl_apply(float* __restrict__ X, const float* __restrict__ num,
                                                       const float* __restrict__ sum, long count, int rows,
                                                       int rank_is_row, int add) {
    for (long k = (long)blockIdx.x * NTHREADS + threadIdx.x; k < count; k += (long)gridDim.x * NTHREADS) {
        const int a = rank_is_row ? (int)(k % rows) : (int)(k / rows);
        if (add)
            X[k] = additive_kl_rule(X[k], sum[a], num[k]);
        else
            X[k] *= (num[k] / sum[a]);
    }
}

// Per-workgroup FP64 partial of sum_k obj_term(V[k], WH[k]).
template <int OBJ>
__global__ __launch_bounds__(NTHREADS) void k_objective_wh(const float* __restrict__ V, const float* __restrict__ WH,
                                                           long count, float eps, double* partials) {
    __shared__ double red[NTHREADS / 64];
    double acc = 0.0;
    for (long k = (long)blockIdx.x * NTHREADS + threadIdx.x; k < count; k += (long)gridDim.x * NTHREADS)
        acc += obj_term<OBJ>(V[k], WH[k], eps);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < NTHREADS / 64; w++) t += red[w];
        partials[blockIdx.x] = t;
    }
}

unsigned elementwise_grid(long count) {
    long g = (count + NTHREADS - 1) / NTHREADS;
    return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

}  // namespace

int objective_wh_partials(long count) {
    long p = (count + (long)NTHREADS * 16 - 1) / ((long)NTHREADS * 16);
    return (int)(p < 1 ? 1 : (p > kMaxObjPartials ? kMaxObjPartials : p));
}

hipError_t launch_kl_ratio(const float* V, float* WH, long count, float eps, int add, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kl_ratio, dim3(elementwise_grid(count)), dim3(NTHREADS), 0, st, V, WH, count, eps,
                       add ? 1.f : 0.f);
    return hipGetLastError();
}

hipError_t launch_kl_apply(float* X, const float* num, const float* sum, long count, int rows, int rank_is_row,
                           int add, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kl_apply, dim3(elementwise_grid(count)), dim3(NTHREADS), 0, st, X, num, sum, count, rows,
                       rank_is_row, add);
    return hipGetLastError();
}

hipError_t launch_objective_wh(const float* V, const float* WH, long count, float eps, int obj, double* partials,
                               double* out, hipStream_t st) {
    const int np = objective_wh_partials(count);
    switch (obj) {
        case OBJ_FROB: hipLaunchKernelGGL(k_objective_wh<OBJ_FROB>, dim3(np), dim3(NTHREADS), 0, st, V, WH, count, eps, partials); break;
        case OBJ_KL: hipLaunchKernelGGL(k_objective_wh<OBJ_KL>, dim3(np), dim3(NTHREADS), 0, st, V, WH, count, eps, partials); break;
        case OBJ_QUALITY: hipLaunchKernelGGL(k_objective_wh<OBJ_QUALITY>, dim3(np), dim3(NTHREADS), 0, st, V, WH, count, eps, partials); break;
        default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_f64(partials, np, out, st);
}

}  // namespace gpumlib_nmf
