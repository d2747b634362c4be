/* C-ABI of the reduction library (gfx950): sum, minimum and maximum of a float array
 * in device memory.  Replaces the static functions of GPUMLib's Reduction class
 * (reference src/cuda/reduction/reduction.h:41-157, reduction.cpp:28-128), which
 * chose their own launch geometry and used a static temporary buffer; here the
 * caller passes a workspace of gpumlib_reduce_workspace_bytes(count) bytes and a
 * stream.  Every function returns hipErrorInvalidValue for bad arguments before it
 * touches the device, and enqueues two kernels otherwise (no allocation, no sync).
 */
#ifndef GPUMLIB_REDUCTION_H
#define GPUMLIB_REDUCTION_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bytes of workspace a reduction of count elements needs. */
size_t gpumlib_reduce_workspace_bytes(long count);

/* out[0] = factor * sum_k x[k], accumulated in FP64 (count may exceed 2^31).
 * Replaces Reduction::Sum (reduction.h:60-71, reduction.cpp:28-44) and, with factor
 * 1/count, Reduction::Average (reduction.h:77-80). */
hipError_t gpumlib_reduce_sum(const float *x, long count, float factor, float *out, void *workspace,
                              size_t workspace_bytes, hipStream_t stream);

/* Minimum (out, nullable) and its index (index, nullable; one of them given), count <= INT_MAX.
 * Start value FLT_MAX; an element replaces the candidate only when strictly smaller (NaN never
 * does); the lowest index wins among equal minima; index -1 if nothing replaced the start value.
 * Replaces Reduction::Min / MinIndex (reduction.h:86-123, reduction.cpp:46-86). */
hipError_t gpumlib_reduce_min(const float *x, long count, float *out, int *index, void *workspace,
                              size_t workspace_bytes, hipStream_t stream);

/* Maximum and its index, as gpumlib_reduce_min with start value -FLT_MAX and "strictly larger".
 * Replaces Reduction::Max / MaxIndex (reduction.h:128-157, reduction.cpp:88-128). */
hipError_t gpumlib_reduce_max(const float *x, long count, float *out, int *index, void *workspace,
                              size_t workspace_bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GPUMLIB_REDUCTION_H */
