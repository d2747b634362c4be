// GPUMLib scalar type and build constants for gfx950.
// Drop-in for src/cuda/definitions.h of the reference (cudafloat :68, CUDA_VALUE :69,
// CUDA_MAX_THREADS_PER_BLOCK :40-43, MAX_CUDAFLOAT / MIN_CUDAFLOAT :78, :91); only the
// names the NMF path and the reduction framework use are kept.
#ifndef GPUMLIB_CUDA_DEFINITIONS_H
#define GPUMLIB_CUDA_DEFINITIONS_H

#include <float.h>
#include <math.h>

// 1024 threads per workgroup on gfx950.
#ifndef CUDA_MAX_THREADS_PER_BLOCK
#define CUDA_MAX_THREADS_PER_BLOCK 1024
#endif

// Reduction geometry constants of the reference (definitions.h:48, :52), kept for
// source compatibility; the library's reductions choose their own launch geometry.
#define SIZE_SMALL_CUDA_VECTOR (3 * CUDA_MAX_THREADS_PER_BLOCK)
#define OPTIMAL_BLOCK_SIZE_REDUCTION 128

typedef float cudafloat;
#define CUDA_VALUE(X) (X##f)
#define MAX_CUDAFLOAT (FLT_MAX)
#define MIN_POSITIVE_CUDAFLOAT (FLT_MIN)
#define MIN_CUDAFLOAT (-FLT_MAX)

#endif  // GPUMLIB_CUDA_DEFINITIONS_H
