// General FP32 GEMM on MFMA for the container API:
//   C (M x N, column-major, ldc) = alpha * op(A) * op(B) + beta * C
// with op = 'N' or 'T', the contract of DeviceMatrix<float>::Multiply /
// MultiplyBySelfTranspose (reference src/memory/DeviceMatrix.h:195-227, which
// call cublasSgemm).  beta == 0 never reads C.
//
// 64x64 output tile per workgroup, one 32x32x2 MFMA block per wave, K staged
// through LDS 16 at a time; global loads follow whichever operand dimension is
// contiguous so that they coalesce for both 'N' and 'T'.
//
// Narrow-operand variants (no reference counterpart): A and B stored in BF16 (bit
// patterns) or in OCP E4M3 with one E8M0 scale per operand, widened to FP32 as they
// are loaded.  A product of two BF16 or two scaled E4M3 values is exact in FP32, so
// the result is the FP32-accumulated GEMM of the stored values, for half or a quarter
// of the operand bytes.
#include "nmf_common.hpp"
#include "nmf_fp8.hpp"
#include "nmf_internal.hpp"

namespace gpumlib_nmf {

constexpr int GT = 64;   // output tile
constexpr int GK = 16;   // K step
constexpr int GP = GT + 32;

// Element idx of an operand in format FMT (0 FP32, 1 BF16 bits, 2 E4M3 times scale) as FP32.
template <int FMT>
__device__ __forceinline__ float ld_operand(const void* __restrict__ p, size_t idx, float scale) {
    if constexpr (FMT == 0) return static_cast<const float*>(p)[idx];
    else if constexpr (FMT == 1) return __uint_as_float((uint32_t) static_cast<const uint16_t*>(p)[idx] << 16);
    else return e4m3_to_f32(static_cast<const uint8_t*>(p)[idx]) * scale;
}

template <int FMT>
__global__ __launch_bounds__(NTHREADS) void k_gemm(int ta, int tb, int M, int N, int K, float alpha,
                                                   const void* __restrict__ A, int lda, float sa,
                                                   const void* __restrict__ B, int ldb, float sb, float beta,
                                                   float* __restrict__ C, int ldc) {
    __shared__ float sA[GK][GP];  // sA[k][i] = op(A)(i0 + i, k0 + k)
    __shared__ float sB[GK][GP];  // sB[k][j] = op(B)(k0 + k, j0 + j)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = lane & 31, h = lane >> 5;
    const int i0 = blockIdx.x * GT, j0 = blockIdx.y * GT;
    const int wm = wave & 1, wn = wave >> 1;
    f32x16 acc;
    for (int v = 0; v < 16; v++) acc[v] = 0.f;

    for (int k0 = 0; k0 < K; k0 += GK) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int e = tid + NTHREADS * j;
            int i, k;
            if (!ta) { i = e % GT; k = e / GT; } else { k = e % GK; i = e / GK; }
            const int gi = i0 + i, gk = k0 + k;
            float v = 0.f;
            if (gi < M && gk < K) v = ld_operand<FMT>(A, ta ? (size_t)gi * lda + gk : (size_t)gk * lda + gi, sa);
            sA[k][i] = v;
            int jj, kb;
            if (!tb) { kb = e % GK; jj = e / GK; } else { jj = e % GT; kb = e / GT; }
            const int gj = j0 + jj, gkb = k0 + kb;
            float w = 0.f;
            if (gj < N && gkb < K) w = ld_operand<FMT>(B, tb ? (size_t)gkb * ldb + gj : (size_t)gj * ldb + gkb, sb);
            sB[kb][jj] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK / 2; kk++)
            acc = mfma32x2(sA[2 * kk + h][wm * 32 + c], sB[2 * kk + h][wn * 32 + c], acc);
        __syncthreads();
    }
    const int j = j0 + wn * 32 + c;
    if (j >= N) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = i0 + wm * 32 + acc_row(v, h);
        if (i < M) {
            float* cp = &C[(size_t)j * ldc + i];
            *cp = (beta == 0.f) ? alpha * acc[v] : alpha * acc[v] + beta * *cp;
        }
    }
}

hipError_t launch_gemm(int fmt, char ta, char tb, int M, int N, int K, float alpha, const void* A, int lda, float sa,
                       const void* B, int ldb, float sb, float beta, float* C, int ldc, hipStream_t st) {
    auto flag = [](char t) { return (t == 'T' || t == 't') ? 1 : ((t == 'N' || t == 'n') ? 0 : -1); };
    const int fa = flag(ta), fb = flag(tb);
    if (fa < 0 || fb < 0 || M < 0 || N < 0 || K < 0 || fmt < 0 || fmt > 2) return hipErrorInvalidValue;
    if (ldc < (M > 1 ? M : 1)) return hipErrorInvalidValue;
    if (M == 0 || N == 0) return hipSuccess;
    dim3 grid((M + GT - 1) / GT, (N + GT - 1) / GT);
    if (fmt == 0)
        hipLaunchKernelGGL(k_gemm<0>, grid, dim3(NTHREADS), 0, st, fa, fb, M, N, K, alpha, A, lda, sa, B, ldb, sb,
                           beta, C, ldc);
    else if (fmt == 1)
        hipLaunchKernelGGL(k_gemm<1>, grid, dim3(NTHREADS), 0, st, fa, fb, M, N, K, alpha, A, lda, sa, B, ldb, sb,
                           beta, C, ldc);
    else
        hipLaunchKernelGGL(k_gemm<2>, grid, dim3(NTHREADS), 0, st, fa, fb, M, N, K, alpha, A, lda, sa, B, ldb, sb,
                           beta, C, ldc);
    return hipGetLastError();
}

hipError_t launch_sgemm(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                        int ldb, float beta, float* C, int ldc, hipStream_t st) {
    return launch_gemm(0, ta, tb, M, N, K, alpha, A, lda, 1.f, B, ldb, 1.f, beta, C, ldc, st);
}

}  // namespace gpumlib_nmf
