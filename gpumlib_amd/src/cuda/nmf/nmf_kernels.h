// NMF kernel interface.  Drop-in for the reference's src/cuda/nmf/nmf_kernels.h:
// the extern "C" launchers (:87, :100, :113) are declared, with hipError_t
// returns, by the library's C-ABI header; the C++ helpers KernelNMFquality (:44),
// KernelSumW (:64) and KernelSumH (:73) are provided here over that C-ABI.
#ifndef GPUMLIB_NMF_KERNELS_H
#define GPUMLIB_NMF_KERNELS_H

#include <hip/hip_runtime.h>

#include "../../../../include/gpumlib_nmf.h"
#include "../definitions.h"

namespace GPUMLib {

// Small value added to the denominator of the update ratios; 10e-9f == 1e-8f, the
// reference value (nmf_kernels.h:34).
#define SMALL_VALUE_TO_ADD_DENOMINATOR (CUDA_VALUE(10e-9))

// sumW[a] = max(sum_i W(i,a), eps); blockSize is accepted and ignored.
inline void KernelSumW(int /*blockSize*/, cudafloat* W, int n, int r, cudafloat* sumW) {
    (void)gpumlib_nmf_sum_w(W, n, r, SMALL_VALUE_TO_ADD_DENOMINATOR, sumW, nullptr);
}

// sumH[a] = max(sum_u H(a,u), eps); blockSize is accepted and ignored.
inline void KernelSumH(int /*blockSize*/, cudafloat* H, int r, int m, cudafloat* sumH) {
    (void)gpumlib_nmf_sum_h(H, r, m, SMALL_VALUE_TO_ADD_DENOMINATOR, sumH, nullptr);
}

// quality[0] = sum_k V[k] log10(WH[k] + eps) - WH[k] over n elements, written as a
// cudafloat like the reference.  Runs on the null stream; the FP64 partial sums
// live in a small device buffer owned by this helper for the life of the process.
inline void KernelNMFquality(int /*blockSize*/, cudafloat* V, cudafloat* WH, int n, cudafloat* quality) {
    static double* scratch = nullptr;
    static size_t scratchBytes = 0;
    const size_t need = gpumlib_nmf_quality_workspace_bytes(n) + sizeof(double);
    if (scratchBytes < need) {
        if (scratch) (void)hipFree(scratch);
        scratch = nullptr;
        scratchBytes = 0;
        if (hipMalloc((void**)&scratch, need) != hipSuccess) return;
        scratchBytes = need;
    }
    double* result = scratch;
    if (gpumlib_nmf_quality(V, WH, n, SMALL_VALUE_TO_ADD_DENOMINATOR, scratch + 1, scratchBytes - sizeof(double),
                            result, nullptr) != hipSuccess)
        return;
    double q = 0.0;
    if (hipMemcpy(&q, result, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess) {
        const cudafloat qf = (cudafloat)q;
        (void)hipMemcpy(quality, &qf, sizeof(cudafloat), hipMemcpyHostToDevice);
    }
}

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_KERNELS_H
