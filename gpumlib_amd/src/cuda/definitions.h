// GPUMLib scalar type and build constants for gfx950.
// Drop-in for src/cuda/definitions.h of the reference (cudafloat :68, CUDA_VALUE :69,
// CUDA_MAX_THREADS_PER_BLOCK :40-43); only the names the NMF path uses are kept.
#ifndef GPUMLIB_CUDA_DEFINITIONS_H
#define GPUMLIB_CUDA_DEFINITIONS_H

#include <float.h>
#include <math.h>

// 1024 threads per workgroup on gfx950.
#ifndef CUDA_MAX_THREADS_PER_BLOCK
#define CUDA_MAX_THREADS_PER_BLOCK 1024
#endif

typedef float cudafloat;
#define CUDA_VALUE(X) (X##f)
#define MAX_CUDAFLOAT (FLT_MAX)
#define MIN_POSITIVE_CUDAFLOAT (FLT_MIN)
#define MIN_CUDAFLOAT (-FLT_MAX)

#endif  // GPUMLIB_CUDA_DEFINITIONS_H
