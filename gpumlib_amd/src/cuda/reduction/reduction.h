// Reduction framework.  Drop-in for GPUMLib::Reduction (reference
// src/cuda/reduction/reduction.h:41-157, reduction.cpp:28-128): static Sum,
// Average, Min, MinIndex, Max and MaxIndex over DeviceArray / DeviceMatrix inputs,
// results written to device memory on the given stream (no synchronisation).
//
// Each call is one gfx950 reduction of the library's C-ABI (gpumlib_reduce_*): a
// grid of wave64 workgroups producing FP64 sums or (value, index) candidates, then
// one workgroup combining them in a fixed order.  Semantics beyond the reference's:
// sums accumulate in FP64; among equal extremes the lowest index is reported; NaN
// elements never become the extreme; an empty input gives sum 0, min FLT_MAX /
// max -FLT_MAX and index -1.
//
// temporaryBuffer is the workspace of every call, grown on demand, as in the
// reference.  It is never freed (its memory goes with the process), so no device
// call runs while the HIP runtime is being torn down at exit.
#ifndef GPUMLIB_REDUCTION_H_CLASS
#define GPUMLIB_REDUCTION_H_CLASS

#include <hip/hip_runtime.h>

#include "../../../../include/gpumlib_reduction.h"
#include "../../memory/DeviceArray.h"
#include "../../memory/DeviceMatrix.h"
#include "../definitions.h"

namespace GPUMLib {

class Reduction {
  private:
    // Grows temporaryBuffer to the workspace of count elements; false if it cannot.
    static bool Workspace(long count) {
        const size_t floats = (gpumlib_reduce_workspace_bytes(count) + sizeof(cudafloat) - 1) / sizeof(cudafloat);
        if (temporaryBuffer.Length() < floats) temporaryBuffer.ResizeWithoutPreservingData(floats);
        return temporaryBuffer.Length() >= floats;
    }
    static size_t WorkspaceBytes() { return temporaryBuffer.Length() * sizeof(cudafloat); }

    static void Sum(cudafloat* inputs, cudafloat* output, int numInputs, cudafloat multiplyFactor, hipStream_t stream) {
        if (Workspace(numInputs))
            (void)gpumlib_reduce_sum(inputs, numInputs, multiplyFactor, output, temporaryBuffer.Pointer(),
                                     WorkspaceBytes(), stream);
    }
    static void MinIndex(cudafloat* inputs, cudafloat* output, int* minIndex, int numInputs, hipStream_t stream) {
        if (Workspace(numInputs))
            (void)gpumlib_reduce_min(inputs, numInputs, output, minIndex, temporaryBuffer.Pointer(), WorkspaceBytes(),
                                     stream);
    }
    static void Min(cudafloat* inputs, cudafloat* output, int numInputs, hipStream_t stream) {
        MinIndex(inputs, output, nullptr, numInputs, stream);
    }
    static void MaxIndex(cudafloat* inputs, cudafloat* output, int* maxIndex, int numInputs, hipStream_t stream) {
        if (Workspace(numInputs))
            (void)gpumlib_reduce_max(inputs, numInputs, output, maxIndex, temporaryBuffer.Pointer(), WorkspaceBytes(),
                                     stream);
    }
    static void Max(cudafloat* inputs, cudafloat* output, int numInputs, hipStream_t stream) {
        MaxIndex(inputs, output, nullptr, numInputs, stream);
    }

  public:
    //! Workspace of the reductions; callers may use it between calls (reference: public too).
    static inline DeviceArray<cudafloat>& temporaryBuffer = *new DeviceArray<cudafloat>();

    //! output[0] = multiplyFactor * sum of the inputs.
    static void Sum(DeviceArray<cudafloat>& inputs, cudafloat* output, cudafloat multiplyFactor = CUDA_VALUE(1.0),
                    hipStream_t stream = nullptr) {
        Sum(inputs.Pointer(), output, (int)inputs.Length(), multiplyFactor, stream);
    }
    static void Sum(DeviceArray<cudafloat>& inputs, DeviceArray<cudafloat>& output,
                    cudafloat multiplyFactor = CUDA_VALUE(1.0), hipStream_t stream = nullptr) {
        Sum(inputs.Pointer(), output.Pointer(), (int)inputs.Length(), multiplyFactor, stream);
    }
    //! output[0] = mean of the inputs (factor 1/length computed in double, as in the reference).
    static void Average(DeviceArray<cudafloat>& inputs, DeviceArray<cudafloat>& output, hipStream_t stream = nullptr) {
        const double factor = 1.0 / (double)inputs.Length();
        Sum(inputs.Pointer(), output.Pointer(), (int)inputs.Length(), (cudafloat)factor, stream);
    }

    static void Min(DeviceArray<cudafloat>& inputs, DeviceArray<cudafloat>& output, hipStream_t stream = nullptr) {
        Min(inputs.Pointer(), output.Pointer(), (int)inputs.Length(), stream);
    }
    static void Min(DeviceMatrix<cudafloat>& inputs, DeviceArray<cudafloat>& output, hipStream_t stream = nullptr) {
        Min(inputs.Pointer(), output.Pointer(), (int)inputs.Elements(), stream);
    }
    static void MinIndex(DeviceArray<cudafloat>& inputs, DeviceArray<cudafloat>& min, DeviceArray<int>& minIndex,
                         hipStream_t stream = nullptr) {
        MinIndex(inputs.Pointer(), min.Pointer(), minIndex.Pointer(), (int)inputs.Length(), stream);
    }
    static void MinIndex(DeviceMatrix<cudafloat>& inputs, DeviceArray<cudafloat>& min, DeviceArray<int>& minIndex,
                         hipStream_t stream = nullptr) {
        MinIndex(inputs.Pointer(), min.Pointer(), minIndex.Pointer(), (int)inputs.Elements(), stream);
    }
    static void MinIndex(DeviceMatrix<cudafloat>& inputs, DeviceArray<int>& minIndex, hipStream_t stream = nullptr) {
        MinIndex(inputs.Pointer(), nullptr, minIndex.Pointer(), (int)inputs.Elements(), stream);
    }

    static void Max(DeviceArray<cudafloat>& inputs, DeviceArray<cudafloat>& output, hipStream_t stream = nullptr) {
        Max(inputs.Pointer(), output.Pointer(), (int)inputs.Length(), stream);
    }
    static void Max(DeviceMatrix<cudafloat>& inputs, DeviceArray<cudafloat>& output, hipStream_t stream = nullptr) {
        Max(inputs.Pointer(), output.Pointer(), (int)inputs.Elements(), stream);
    }
    static void MaxIndex(DeviceArray<cudafloat>& inputs, DeviceArray<cudafloat>& max, DeviceArray<int>& maxIndex,
                         hipStream_t stream = nullptr) {
        MaxIndex(inputs.Pointer(), max.Pointer(), maxIndex.Pointer(), (int)inputs.Length(), stream);
    }
    static void MaxIndex(DeviceMatrix<cudafloat>& inputs, DeviceArray<cudafloat>& max, DeviceArray<int>& maxIndex,
                         hipStream_t stream = nullptr) {
        MaxIndex(inputs.Pointer(), max.Pointer(), maxIndex.Pointer(), (int)inputs.Elements(), stream);
    }
};

}  // namespace GPUMLib

#endif
