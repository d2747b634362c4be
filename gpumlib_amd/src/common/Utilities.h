// Launch-geometry helpers.  Drop-in for NumberThreadsPerBlockThatBestFit and
// NumberBlocks (reference src/common/Utilities.h:38-56).
#ifndef GPUMLIB_UTILITIES_H
#define GPUMLIB_UTILITIES_H

#include "../cuda/definitions.h"

namespace GPUMLib {

// Smallest power of two >= threads, capped at maxThreadsPerBlock.
inline int NumberThreadsPerBlockThatBestFit(int threads, int maxThreadsPerBlock = CUDA_MAX_THREADS_PER_BLOCK) {
    int p = 1;
    while (p < threads && p < maxThreadsPerBlock) p *= 2;
    return p;
}

// ceil(threads / blockSize)
inline int NumberBlocks(int threads, int blockSize) { return (threads + blockSize - 1) / blockSize; }

}  // namespace GPUMLib

#endif
