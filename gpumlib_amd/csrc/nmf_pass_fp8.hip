// FP8 (OCP E4M3) storage path with block scales.
//
// V is quantised once with one E8M0 scale per 32 x 32 tile of V, shared by both
// stored orientations (V and V^T), so the two passes see the same quantised
// matrix.  The W and H operands are re-quantised after every update with one
// scale per 32 consecutive elements along the contraction dimension.  Products
// run on the scaled MFMA v_mfma_scale_f32_32x32x64_f8f6f4, which applies the
// operand scales per 32-element K block and accumulates in FP32; updates, ratio
// and eps guard stay FP32.  Like the BF16 path, both passes are K-contiguous
// passes P[s](a, u) = sum_k A(k, a) B(k, u) (W half = H half of (V^T, H^T, W^T)).
#include <hip/hip_runtime.h>

#include "nmf_common.hpp"
#include "nmf_fp8.hpp"
#include "nmf_internal.hpp"

#include <type_traits>

namespace gpumlib_nmf {

typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int K8 = 128;        // K per pipeline step: 2 MFMAs of K = 64
constexpr int LD8 = K8 + 16;   // LDS row pitch in bytes (36 words)

template <int MB>
__global__ __launch_bounds__(NTHREADS, 2) void k_pass_tn_fp8(Fp8PassArgs p) {
    constexpr int RP = MB * 32;
    __shared__ __attribute__((aligned(16))) uint8_t sA[2][RP][LD8];
    __shared__ __attribute__((aligned(16))) uint8_t sB[2][TILE][LD8];
    __shared__ uint8_t sAs[2][RP][4];    // E8M0 scale per (row of A^T, 32-block of this step)
    __shared__ uint8_t sBs[2][TILE][4];  // E8M0 scale per (column of B, 32-block of this step)

    const int total = p.tiles * p.S;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: branches on it stay scalar
    const int c = lane & 31, h = lane >> 5;
    // the first column tile of every slice also accumulates the slice's gram A^T A
    // from the A rows staged in LDS (block (mb, nb) on wave mb*MB + nb, modulo 4)
    const bool gram = (t == 0) && p.G != nullptr;
    const int col0 = t * TILE;
    const int ncols = min(TILE, p.N - col0);
    // The stores and shadows are padded to whole chunks with zeros (and scales 127), so
    // every chunk is loaded without bounds checks; lanes past the last column or rank
    // row re-read the last valid one (their products land only in output columns that
    // are not stored and in rank rows >= r, which no update reads).
    const int kbeg = s * p.kslice;
    const int kend = min(round_up(p.K, K8), kbeg + p.kslice);
    constexpr int GB = (MB * MB + 3) / 4;

    // Two register sets of one chunk each (native vectors, selected by a compile-time
    // tag, so they stay in registers).
    u32x4 ra[2][MB], rb[2][4];
    uint8_t rsa[2], rsb[2][2];
    using Set0 = std::integral_constant<int, 0>;
    using Set1 = std::integral_constant<int, 1>;
    auto gload = [&](auto set, int k0) __attribute__((always_inline)) {
        constexpr int R = decltype(set)::value;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = tid + NTHREADS * j, col = q >> 3, kq = (q & 7) * 16;
            const int cc = col < ncols ? col : ncols - 1;
            rb[R][j] = *reinterpret_cast<const u32x4*>(p.B + (size_t)(col0 + cc) * p.ldb + k0 + kq);
        }
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const int q = tid + NTHREADS * j, a = q >> 3, kq = (q & 7) * 16;
            const int aa = a < p.r ? a : p.r - 1;
            ra[R][j] = *reinterpret_cast<const u32x4*>(p.A + (size_t)aa * p.lda + k0 + kq);
        }
        const int kb0 = k0 >> 5;
        {
            const int a = (tid >> 2) < p.r ? (tid >> 2) : p.r - 1, kb = kb0 + (tid & 3);
            rsa[R] = p.Asc[(size_t)a * p.nbk + kb];
        }
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int e = tid + NTHREADS * jj, col = e >> 2, kb = kb0 + (e & 3);
            const int cc = col < ncols ? col : ncols - 1;
            rsb[R][jj] = p.Bsc[(size_t)((col0 + cc) >> p.bcol_shift) * p.bs_col + (size_t)kb * p.bs_k];
        }
    };
    auto sstore = [&](auto set, int buf) __attribute__((always_inline)) {
        constexpr int R = decltype(set)::value;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = tid + NTHREADS * j, col = q >> 3, kq = (q & 7) * 16;
            *reinterpret_cast<u32x4*>(&sB[buf][col][kq]) = rb[R][j];
        }
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const int q = tid + NTHREADS * j, a = q >> 3, kq = (q & 7) * 16;
            *reinterpret_cast<u32x4*>(&sA[buf][a][kq]) = ra[R][j];
        }
        if (tid < RP * 4) sAs[buf][tid >> 2][tid & 3] = rsa[R];
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int e = tid + NTHREADS * jj;
            sBs[buf][e >> 2][e & 3] = rsb[R][jj];
        }
    };

    f32x16 acc[MB], gacc[GB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
        for (int v = 0; v < 16; v++) acc[mb][v] = 0.f;
#pragma unroll
    for (int q = 0; q < GB; q++)
        for (int v = 0; v < 16; v++) gacc[q][v] = 0.f;

    auto compute = [&](int buf) __attribute__((always_inline)) {
        // The instruction takes K block 0 of a row from bytes 0..15 of both lane halves and
        // its scale from lane c, K block 1 from bytes 16..31 and its scale from lane c + 32.
        // So lane (c, h) supplies k = g*64 + 16h + j in bytes j < 16 and k = g*64 + 32 + 16h + j
        // in bytes 16 + j, and the scale of block 2g + h of this step.
#pragma unroll
        for (int g = 0; g < K8 / 64; g++) {
            const uint8_t* pb = &sB[buf][wave * 32 + c][g * 64 + h * 16];
            const uint4 b0 = *reinterpret_cast<const uint4*>(pb), b1 = *reinterpret_cast<const uint4*>(pb + 32);
            const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
            const int sb = sBs[buf][wave * 32 + c][2 * g + h];
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                const uint8_t* pa = &sA[buf][mb * 32 + c][g * 64 + h * 16];
                const uint4 a0 = *reinterpret_cast<const uint4*>(pa), a1 = *reinterpret_cast<const uint4*>(pa + 32);
                const i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                                  (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                const int sa = sAs[buf][mb * 32 + c][2 * g + h];
                acc[mb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc[mb], 0, 0, 0, sa, 0, sb);
            }
        }
        if (gram) {
#pragma unroll
            for (int q = 0; q < GB; q++) {
                const int blk = wave + 4 * q;
                if (blk >= MB * MB) continue;
                const int mb = blk / MB, nb = blk - (blk / MB) * MB;
#pragma unroll
                for (int g = 0; g < K8 / 64; g++) {
                    const uint8_t* pa = &sA[buf][mb * 32 + c][g * 64 + h * 16];
                    const uint8_t* pb = &sA[buf][nb * 32 + c][g * 64 + h * 16];
                    const uint4 a0 = *reinterpret_cast<const uint4*>(pa), a1 = *reinterpret_cast<const uint4*>(pa + 32);
                    const uint4 b0 = *reinterpret_cast<const uint4*>(pb), b1 = *reinterpret_cast<const uint4*>(pb + 32);
                    const i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                                      (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                    const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                                      (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
                    const int sa = sAs[buf][mb * 32 + c][2 * g + h], sb = sAs[buf][nb * 32 + c][2 * g + h];
                    gacc[q] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, gacc[q], 0, 0, 0, sa, 0, sb);
                }
            }
        }
    };

    // Stages of two chunks: the loads of chunks ch+2 and ch+3 are issued, chunks ch and
    // ch+1 are multiplied out of LDS buffers 0 and 1 while those loads are in flight, and
    // after a barrier the new chunks are stored (no loop-carried staging registers).
    const int nchunks = (kend - kbeg + K8 - 1) / K8;
    gload(Set0{}, kbeg);
    if (nchunks > 1) gload(Set1{}, kbeg + K8);
    sstore(Set0{}, 0);
    if (nchunks > 1) sstore(Set1{}, 1);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch += 2) {
        const bool more0 = ch + 2 < nchunks, more1 = ch + 3 < nchunks;
        if (more0) gload(Set0{}, kbeg + (ch + 2) * K8);
        if (more1) gload(Set1{}, kbeg + (ch + 3) * K8);
        compute(0);
        if (ch + 1 < nchunks) compute(1);
        __syncthreads();
        if (more0) sstore(Set0{}, 0);
        if (more1) sstore(Set1{}, 1);
        __syncthreads();
    }

    float* slice = p.P + (size_t)s * p.pstride;
    const int ul = wave * 32 + c;
    if (ul < ncols) {
        float* dst = slice + (size_t)(col0 + ul) * p.ldr;
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int a0 = mb * 32 + g * 8 + h * 4;
                if (a0 < p.ldr) {
                    float4 o = make_float4(acc[mb][g * 4], acc[mb][g * 4 + 1], acc[mb][g * 4 + 2], acc[mb][g * 4 + 3]);
                    *reinterpret_cast<float4*>(dst + a0) = o;
                }
            }
        }
    }
    if (gram) {
        // gacc[q][v] = gram(a = mb*32 + acc_row(v,h), b = nb*32 + c), stored at column N + b
#pragma unroll
        for (int q = 0; q < GB; q++) {
            const int blk = wave + 4 * q;
            if (blk >= MB * MB) continue;
            const int mb = blk / MB, b = (blk - (blk / MB) * MB) * 32 + c;
            if (b >= p.r) continue;
            float* dst = slice + (size_t)(p.N + b) * p.ldr;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int a0 = mb * 32 + g * 8 + h * 4;
                if (a0 < p.ldr)
                    *reinterpret_cast<float4*>(dst + a0) =
                        make_float4(gacc[q][g * 4], gacc[q][g * 4 + 1], gacc[q][g * 4 + 2], gacc[q][g * 4 + 3]);
            }
        }
        gram_finish4(p.P + (size_t)p.N * p.ldr, p.pstride, p.S, p.S, p.r * p.ldr, p.G, p.gcount);
    }
}

// ---------------------------------------------------------------------------
// V (n x m, FP32) -> Vq (column-major, ld ldv), Vtq (V^T column-major, ld ldt) and one
// E8M0 scale per 32 x 32 tile: Vsc[ub * nbi + ib] (nbi = ldv / 32).  One workgroup per
// tile of the padded stores; tiles past V's edges get zeros and the scale 127 (2^0).
__global__ __launch_bounds__(NTHREADS) void k_pack_v_fp8(const float* __restrict__ V, int n, int m,
                                                         uint8_t* __restrict__ Vq, int ldv, uint8_t* __restrict__ Vtq,
                                                         int ldt, uint8_t* __restrict__ Vsc, int nbi) {
    __shared__ float tile[32][33];
    __shared__ float red[NTHREADS / 64];
    __shared__ int sscale;
    const int ib = blockIdx.x, ub = blockIdx.y;
    const int i0 = ib * 32, u0 = ub * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    float amax = 0.f;
    for (int uu = ty; uu < 32; uu += 8) {
        const int i = i0 + tx, u = u0 + uu;
        const float v = (i < n && u < m) ? V[(size_t)u * n + i] : 0.f;
        tile[uu][tx] = v;
        amax = fmaxf(amax, fabsf(v));
    }
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_down(amax, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = amax;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = red[0];
        for (int k = 1; k < NTHREADS / 64; k++) a = fmaxf(a, red[k]);
        const int sc = e8m0_for_amax(a);
        sscale = sc;
        Vsc[(size_t)ub * nbi + ib] = (uint8_t)sc;
    }
    __syncthreads();
    const int sc = sscale;
    // the grid covers the padded stores (ldv x ldt): padding gets zeros (tile values 0)
    for (int uu = ty; uu < 32; uu += 8) {
        const int i = i0 + tx, u = u0 + uu;
        if (u < m) Vq[(size_t)u * ldv + i] = f32_to_e4m3(e8m0_unscale(tile[uu][tx], sc));
    }
    for (int ii = ty; ii < 32; ii += 8) {
        const int u = u0 + tx, i = i0 + ii;
        if (i < n) Vtq[(size_t)i * ldt + u] = f32_to_e4m3(e8m0_unscale(tile[tx][ii], sc));
    }
}

// Operand shadows with one scale per 32 elements along the contraction dimension:
//   which 0: W (n x r) -> Wq column-major (ld ldq), Wsc[a * nblk + ib]   (blocks along i)
//   which 1: H (r x m) -> Hq = H^T column-major (ld ldq), Hsc[a * nblk + ub] (blocks along u)
// nblk = ldq / 32 covers the padded length; entries past n (or m) are zeros.
// One thread per (rank column a, block).
__global__ __launch_bounds__(NTHREADS) void k_shadow_fp8(const float* __restrict__ X, int n, int m, int r, int which,
                                                         uint8_t* __restrict__ Xq, int ldq, uint8_t* __restrict__ Xsc,
                                                         int nblk) {
    const int t = blockIdx.x * NTHREADS + threadIdx.x;
    if (t >= r * nblk) return;
    const int a = t / nblk, blk = t - a * nblk;
    const int len = which == 0 ? n : m;
    const int k0 = blk * 32, k1 = min(len, k0 + 32);
    auto at = [&](int k) -> float { return which == 0 ? X[(size_t)a * n + k] : X[(size_t)k * r + a]; };
    float amax = 0.f;
    for (int k = k0; k < k1; k++) amax = fmaxf(amax, fabsf(at(k)));
    const int sc = e8m0_for_amax(amax);
    Xsc[t] = (uint8_t)sc;
    for (int k = k0; k < k0 + 32; k++) Xq[(size_t)a * ldq + k] = k < k1 ? f32_to_e4m3(e8m0_unscale(at(k), sc)) : 0;
}

// ---------------------------------------------------------------------------
template <int MB>
static void launch_tn8(const Fp8PassArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_pass_tn_fp8<MB>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS), 0, st, a);
}

hipError_t launch_pass_tn_fp8(const Fp8PassArgs& a, hipStream_t st) {
    const int rp = round_up(a.r, 32);
    if (rp <= 32)
        launch_tn8<1>(a, st);
    else if (rp <= 64)
        launch_tn8<2>(a, st);
    else
        return hipErrorInvalidValue;  // two rank blocks at most on this path
    return hipGetLastError();
}

hipError_t launch_pack_v_fp8(const float* V, int n, int m, uint8_t* Vq, int ldv, uint8_t* Vtq, int ldt, uint8_t* Vsc,
                             hipStream_t st) {
    if (ldv % 32 || ldt % 32 || ldv < n || ldt < m) return hipErrorInvalidValue;
    const int nbi = ldv / 32, nbu = ldt / 32;
    hipLaunchKernelGGL(k_pack_v_fp8, dim3(nbi, nbu), dim3(NTHREADS), 0, st, V, n, m, Vq, ldv, Vtq, ldt, Vsc, nbi);
    return hipGetLastError();
}

hipError_t launch_shadow_fp8(const float* X, int n, int m, int r, int which, uint8_t* Xq, int ldq, uint8_t* Xsc,
                             hipStream_t st) {
    if (ldq % 32 || ldq < (which == 0 ? n : m)) return hipErrorInvalidValue;
    const int nblk = ldq / 32;
    const int threads = r * nblk;
    hipLaunchKernelGGL(k_shadow_fp8, dim3((threads + NTHREADS - 1) / NTHREADS), dim3(NTHREADS), 0, st, X, n, m, r,
                       which, Xq, ldq, Xsc, nblk);
    return hipGetLastError();
}

}  // namespace gpumlib_nmf
