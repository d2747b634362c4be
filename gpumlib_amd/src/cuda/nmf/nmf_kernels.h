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

// sumW[a] = max(sum_i W(i,a), eps), any r; blockSize is accepted and ignored.  Null stream, as
// the reference; a launch error is returned (the reference's helper returned nothing).
inline hipError_t KernelSumW(int /*blockSize*/, cudafloat* W, int n, int r, cudafloat* sumW) {
    return gpumlib_nmf_sum_w(W, n, r, SMALL_VALUE_TO_ADD_DENOMINATOR, sumW, nullptr);
}

// sumH[a] = max(sum_u H(a,u), eps), any r; blockSize is accepted and ignored.
inline hipError_t KernelSumH(int /*blockSize*/, cudafloat* H, int r, int m, cudafloat* sumH) {
    return gpumlib_nmf_sum_h(H, r, m, SMALL_VALUE_TO_ADD_DENOMINATOR, sumH, nullptr);
}

// quality[0] = sum_k V[k] log10(WH[k] + eps) - WH[k] over n elements, written as a cudafloat
// like the reference, stream-ordered on the null stream (nothing is read back to the host).
// The FP64 partial sums live in a device buffer this helper allocates on the first call (and
// again only when a larger n arrives), kept for the life of the process.
inline hipError_t KernelNMFquality(int /*blockSize*/, cudafloat* V, cudafloat* WH, int n, cudafloat* quality) {
    static void* scratch = nullptr;
    static size_t scratchBytes = 0;
    const size_t need = gpumlib_nmf_quality_workspace_bytes(n);
    if (scratchBytes < need) {
        if (scratch) (void)hipFree(scratch);
        scratch = nullptr;
        scratchBytes = 0;
        const hipError_t e = hipMalloc(&scratch, need);
        if (e != hipSuccess) return e;
        scratchBytes = need;
    }
    return gpumlib_nmf_quality_f32(V, WH, n, SMALL_VALUE_TO_ADD_DENOMINATOR, scratch, scratchBytes, quality, nullptr);
}

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_KERNELS_H
