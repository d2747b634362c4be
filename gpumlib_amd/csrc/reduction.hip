// Reductions over a float array on gfx950: sum (times a factor) and minimum /
// maximum with the index of the extreme.  The MI355X form of GPUMLib's Reduction
// class (reference src/cuda/reduction/reduction.h:41-157, reduction.cpp:28-128,
// kernels reduction_kernels_{sum,min,max}.cu).
//
// Two launches per call:
//   1. partials: G workgroups of 256 threads (4 wave64s).  Each thread walks the
//      array with a grid stride, issuing four 16-byte loads before it uses any (when
//      the array is 16-byte aligned), and keeps an FP64 running sum or a
//      (value, index) candidate; the lanes of a wave combine through cross-lane
//      shuffles, the four waves through LDS; one partial per workgroup.
//   2. final: one workgroup combines the G partials in workgroup order.
// G depends only on the element count and every combining order is fixed, so a
// result is reproducible bit for bit.  The passes are HBM-bound (4 bytes read per
// element, nothing written but the partials).
//
// Semantics (the reference's, with the cases it leaves open fixed):
//   sum:  out = factor * sum_k x[k], accumulated in FP64 and rounded once to float
//         (the reference accumulates in float);
//   min:  start from FLT_MAX (MAX_CUDAFLOAT), an element replaces the candidate only
//         when strictly smaller, so NaN never does; among equal minima the lowest
//         index wins; index -1 when no element replaced the start value (empty array);
//   max:  the same from -FLT_MAX (MIN_CUDAFLOAT) with "strictly larger".
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <stdint.h>

#include "../../include/gpumlib_reduction.h"

namespace gpumlib_reduce {

constexpr int NT = 256;         // threads per workgroup
constexpr int MAX_GROUPS = 1024;
constexpr int NONE = INT_MAX;   // "no index yet": loses every tie

// Workgroups of the partial pass for count elements: about 16 elements per thread
// at least, at most MAX_GROUPS (4 per CU).
static int groups_for(long count) {
    long g = (count + (long)NT * 16 - 1) / ((long)NT * 16);
    if (g < 1) g = 1;
    if (g > MAX_GROUPS) g = MAX_GROUPS;
    return (int)g;
}

// ------------------------------------------------------------------ sum --
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of the workgroup's values; the result is valid in thread 0.
__device__ __forceinline__ double group_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; w++) t += red[w];
    return t;
}

__global__ __launch_bounds__(NT) void k_sum_partials(const float* __restrict__ x, long count, int vec,
                                                     double* __restrict__ part) {
    __shared__ double red[NT / 64];
    const long stride = (long)gridDim.x * NT;
    const long t0 = (long)blockIdx.x * NT + threadIdx.x;
    double acc = 0.0;
    long tail = 0;
    if (vec) {
        const float4* x4 = reinterpret_cast<const float4*>(x);
        const long n4 = count >> 2;
        long k = t0;
        for (; k + 3 * stride < n4; k += 4 * stride) {
            const float4 a = x4[k], b = x4[k + stride], c = x4[k + 2 * stride], d = x4[k + 3 * stride];
            acc += (double)a.x + (double)a.y + (double)a.z + (double)a.w;
            acc += (double)b.x + (double)b.y + (double)b.z + (double)b.w;
            acc += (double)c.x + (double)c.y + (double)c.z + (double)c.w;
            acc += (double)d.x + (double)d.y + (double)d.z + (double)d.w;
        }
        for (; k < n4; k += stride) {
            const float4 a = x4[k];
            acc += (double)a.x + (double)a.y + (double)a.z + (double)a.w;
        }
        tail = n4 << 2;
    }
    for (long k = tail + t0; k < count; k += stride) acc += (double)x[k];
    const double t = group_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(NT) void k_sum_final(const double* __restrict__ part, int groups, float factor,
                                                  float* __restrict__ out) {
    __shared__ double red[NT / 64];
    double acc = 0.0;
    for (int g = threadIdx.x; g < groups; g += NT) acc += part[g];
    const double t = group_sum(acc, red);
    if (threadIdx.x == 0) out[0] = (float)(t * (double)factor);
}

// -------------------------------------------------------------- min / max --
struct Cand {
    float v;
    int i;
};

template <bool MAX>
__device__ __forceinline__ bool better(float a, float b) {  // a strictly better than b (false for NaN)
    return MAX ? (a > b) : (a < b);
}

// b replaces a: strictly better, or equal with a lower index.
template <bool MAX>
__device__ __forceinline__ Cand pick(Cand a, Cand b) {
    return (better<MAX>(b.v, a.v) || (b.v == a.v && b.i < a.i)) ? b : a;
}

template <bool MAX>
__device__ __forceinline__ Cand wave_pick(Cand c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        Cand o;
        o.v = __shfl_xor(c.v, off, 64);
        o.i = __shfl_xor(c.i, off, 64);
        c = pick<MAX>(c, o);
    }
    return c;
}

template <bool MAX>
__device__ __forceinline__ Cand group_pick(Cand c, float* rv, int* ri) {
    c = wave_pick<MAX>(c);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        rv[wave] = c.v;
        ri[wave] = c.i;
    }
    __syncthreads();
    Cand t = {rv[0], ri[0]};
    if (threadIdx.x == 0)
        for (int w = 1; w < NT / 64; w++) t = pick<MAX>(t, Cand{rv[w], ri[w]});
    return t;
}

// One element in index order: replaces the candidate only when strictly better, so
// within a thread the earliest of equal extremes is kept.
template <bool MAX>
__device__ __forceinline__ void visit(Cand& c, float v, long k) {
    if (better<MAX>(v, c.v)) {
        c.v = v;
        c.i = (int)k;
    }
}

template <bool MAX>
__device__ __forceinline__ void visit4(Cand& c, float4 q, long base) {
    visit<MAX>(c, q.x, base);
    visit<MAX>(c, q.y, base + 1);
    visit<MAX>(c, q.z, base + 2);
    visit<MAX>(c, q.w, base + 3);
}

template <bool MAX>
__global__ __launch_bounds__(NT) void k_ext_partials(const float* __restrict__ x, long count, int vec,
                                                     float* __restrict__ pv, int* __restrict__ pi) {
    __shared__ float rv[NT / 64];
    __shared__ int ri[NT / 64];
    const long stride = (long)gridDim.x * NT;
    const long t0 = (long)blockIdx.x * NT + threadIdx.x;
    Cand c = {MAX ? -FLT_MAX : FLT_MAX, NONE};
    long tail = 0;
    if (vec) {
        const float4* x4 = reinterpret_cast<const float4*>(x);
        const long n4 = count >> 2;
        long k = t0;
        for (; k + 3 * stride < n4; k += 4 * stride) {
            const float4 a = x4[k], b = x4[k + stride], e = x4[k + 2 * stride], d = x4[k + 3 * stride];
            visit4<MAX>(c, a, k << 2);
            visit4<MAX>(c, b, (k + stride) << 2);
            visit4<MAX>(c, e, (k + 2 * stride) << 2);
            visit4<MAX>(c, d, (k + 3 * stride) << 2);
        }
        for (; k < n4; k += stride) visit4<MAX>(c, x4[k], k << 2);
        tail = n4 << 2;
    }
    for (long k = tail + t0; k < count; k += stride) visit<MAX>(c, x[k], k);
    const Cand t = group_pick<MAX>(c, rv, ri);
    if (threadIdx.x == 0) {
        pv[blockIdx.x] = t.v;
        pi[blockIdx.x] = t.i;
    }
}

template <bool MAX>
__global__ __launch_bounds__(NT) void k_ext_final(const float* __restrict__ pv, const int* __restrict__ pi,
                                                  int groups, float* __restrict__ out, int* __restrict__ index) {
    __shared__ float rv[NT / 64];
    __shared__ int ri[NT / 64];
    Cand c = {MAX ? -FLT_MAX : FLT_MAX, NONE};
    for (int g = threadIdx.x; g < groups; g += NT) c = pick<MAX>(c, Cand{pv[g], pi[g]});
    const Cand t = group_pick<MAX>(c, rv, ri);
    if (threadIdx.x == 0) {
        if (out) out[0] = t.v;
        if (index) index[0] = (t.i == NONE) ? -1 : t.i;
    }
}

static size_t workspace_bytes(long count) {
    // G doubles (sum) or G floats + G ints (min / max), 256-byte rounded
    const size_t g = (size_t)groups_for(count < 0 ? 0 : count);
    return (g * 8 + 255) / 256 * 256;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

template <bool MAX>
static hipError_t extreme(const float* x, long count, float* out, int* index, void* ws, size_t ws_bytes,
                          hipStream_t st) {
    if (count < 0 || count > INT_MAX || (count > 0 && !x) || (!out && !index) || !ws ||
        ws_bytes < workspace_bytes(count))
        return hipErrorInvalidValue;
    const int g = groups_for(count);
    float* pv = (float*)ws;
    int* pi = (int*)((char*)ws + (size_t)g * 4);
    hipLaunchKernelGGL(k_ext_partials<MAX>, dim3(g), dim3(NT), 0, st, x, count, aligned16(x) ? 1 : 0, pv, pi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ext_final<MAX>, dim3(1), dim3(NT), 0, st, pv, pi, g, out, index);
    return hipGetLastError();
}

}  // namespace gpumlib_reduce

using namespace gpumlib_reduce;

extern "C" {

size_t gpumlib_reduce_workspace_bytes(long count) { return workspace_bytes(count); }

hipError_t gpumlib_reduce_sum(const float* x, long count, float factor, float* out, void* workspace,
                              size_t workspace_bytes_, hipStream_t stream) {
    if (count < 0 || (count > 0 && !x) || !out || !workspace || workspace_bytes_ < workspace_bytes(count))
        return hipErrorInvalidValue;
    const int g = groups_for(count);
    double* part = (double*)workspace;
    hipLaunchKernelGGL(k_sum_partials, dim3(g), dim3(NT), 0, stream, x, count, aligned16(x) ? 1 : 0, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(NT), 0, stream, part, g, factor, out);
    return hipGetLastError();
}

hipError_t gpumlib_reduce_min(const float* x, long count, float* out, int* index, void* workspace,
                              size_t workspace_bytes_, hipStream_t stream) {
    return extreme<false>(x, count, out, index, workspace, workspace_bytes_, stream);
}

hipError_t gpumlib_reduce_max(const float* x, long count, float* out, int* index, void* workspace,
                              size_t workspace_bytes_, hipStream_t stream) {
    return extreme<true>(x, count, out, index, workspace, workspace_bytes_, stream);
}

}  // extern "C"
