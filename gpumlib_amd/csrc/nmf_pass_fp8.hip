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

constexpr int K8 = 512;        // K per chunk: 512 contiguous bytes of every column (8 MFMA steps of 64)
constexpr int LD8 = K8 + 16;   // LDS row pitch in bytes
constexpr int NKB = K8 / 32;   // E8M0 scale blocks per chunk

__device__ __forceinline__ f32x16 mfma_fp8(u32x4 a0, u32x4 a1, u32x4 b0, u32x4 b1, int sa, int sb, f32x16 acc) {
    const i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
    const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, sa, 0, sb);
}

// Same structure as the BF16 pass: one workgroup per (slice s, column tile t) of 64
// columns, chunks 512 elements (bytes) deep so every column is read in 512-byte runs,
// the chunk's B columns and A rows (and their block scales) staged through LDS, the next
// chunk's loads in flight in registers while the current one is multiplied, the 2*MB
// output blocks spread over the 4 waves, the slice's gram spread as in the BF16 pass
// (block t on column tile t, its 8 k steps per chunk split over the waves, the waves'
// parts added in wave order; on tile 0, by wave, when there are fewer than MB * MB tiles).
//
// B scales come from the 32 x 32 tile scales of V: one scale per (32-column block, K
// block), so the 64-column tile needs two rows of 16 scales per chunk; A scales are per
// (rank row, K block).
//
// The instruction takes K block 0 of a row from bytes 0..15 of both lane halves and its
// scale from lane c, K block 1 from bytes 16..31 and its scale from lane c + 32.  So lane
// (c, h) supplies k = g*64 + 16h + j in bytes j < 16 and k = g*64 + 32 + 16h + j in bytes
// 16 + j, and the scale of block 2g + h of the chunk.
//
// The stores and shadows are padded to whole chunks with zeros (scales 127), so no load
// needs a bounds check; lanes past the last column or rank row re-read the last valid one
// (their products land only in output columns that are not stored and in rank rows >= r,
// which no update reads).
// SNT: the V loads take the non-temporal policy (launched when p.snt_k0 < K: every slice; this
// pass takes the policy for all of V or none of it).
template <int MB, bool SNT>
__global__ __launch_bounds__(NTHREADS, 2) void k_pass_tn_fp8(Fp8PassArgs p) {
    constexpr int RP = MB * 32;
    constexpr int NBLK = 2 * MB;
    constexpr int BPW = NBLK >= 4 ? NBLK / 4 : 1;
    constexpr int GB = (MB * MB + 3) / 4;
    constexpr int BL = TILE_LP * K8 / 16 / NTHREADS;  // 16-byte B loads per thread per chunk
    constexpr int AL = RP * K8 / 16 / NTHREADS;       // 16-byte A loads per thread per chunk
    constexpr int LPC = K8 / 16;                      // lanes per column (or row) in a load
    __shared__ __attribute__((aligned(16))) uint8_t sA[RP][LD8];
    __shared__ __attribute__((aligned(16))) uint8_t sB[TILE_LP][LD8];
    // scale of (rank row a, K block kb) at sAs[kb][a]: the 32 lanes of a half-wave read 32
    // consecutive bytes (8 dwords, no bank conflict) where a row-major [a][kb] layout put
    // lanes c and c + 8 on one bank
    __shared__ __attribute__((aligned(16))) uint8_t sAs[NKB][RP];
    __shared__ __attribute__((aligned(16))) uint8_t sBs[2][NKB];   // scale of (32-column block, K block)

    const int total = p.tiles * p.S;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: branches on it stay scalar
    const int c = lane & 31, h = lane >> 5;
    const bool spread = p.tiles >= MB * MB;  // grid-uniform
    const bool gram = p.G != nullptr && (spread ? t < MB * MB : t == 0);
    const int col0 = t * TILE_LP;
    const int ncols = min(TILE_LP, p.N - col0);
    const int kbeg = s * p.kslice;
    const int kend = min(round_up(p.K, K8), kbeg + p.kslice);

    u32x4 rb[BL], ra[AL];
    uint32_t rsa = 0;   // 4 A scales (row tid/4, blocks 4*(tid%4) .. +3)
    uint8_t rsb = 127;  // one B scale (column block tid/16, block tid%16), threads < 32
    auto gload = [&](int k0) __attribute__((always_inline)) {
        // the chunk's B block: 32 KB contiguous, element (kq, col) at col*K8 + kq, so load q
        // (16 bytes) lands at sB[q / LPC][(q % LPC) * 16]
        const uint8_t* bblk = p.B + ((size_t)(k0 / K8) * p.tiles + t) * (TILE_LP * K8);
#pragma unroll
        for (int j = 0; j < BL; j++) {
            const u32x4* src = reinterpret_cast<const u32x4*>(bblk + (size_t)(tid + NTHREADS * j) * 16);
            rb[j] = SNT ? stream_load(src) : *src;
        }
#pragma unroll
        for (int j = 0; j < AL; j++) {
            const int q = tid + NTHREADS * j, a = q / LPC, kq = (q % LPC) * 16;
            const int aa = a < p.r ? a : p.r - 1;
            ra[j] = *reinterpret_cast<const u32x4*>(p.A + (size_t)aa * p.lda + k0 + kq);
        }
        const int kb0 = k0 >> 5;
        if (tid < RP * 4) {
            const int a = tid >> 2, aa = a < p.r ? a : p.r - 1;
            rsa = *reinterpret_cast<const uint32_t*>(p.Asc + (size_t)aa * p.nbk + kb0 + 4 * (tid & 3));
        }
        if (tid < 2 * NKB) {
            const int cb = tid / NKB, kb = kb0 + tid % NKB;
            const int cc = cb * 32 < ncols ? cb * 32 : ncols - 1;
            rsb = p.Bsc[(size_t)((col0 + cc) >> p.bcol_shift) * p.bs_col + (size_t)kb * p.bs_k];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the multiply
    };
    auto sstore = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < BL; j++) {
            const int q = tid + NTHREADS * j, col = q / LPC, kq = (q % LPC) * 16;
            *reinterpret_cast<u32x4*>(&sB[col][kq]) = rb[j];
        }
#pragma unroll
        for (int j = 0; j < AL; j++) {
            const int q = tid + NTHREADS * j, a = q / LPC, kq = (q % LPC) * 16;
            *reinterpret_cast<u32x4*>(&sA[a][kq]) = ra[j];
        }
        if (tid < RP * 4) {
#pragma unroll
            for (int j = 0; j < 4; j++) sAs[4 * (tid & 3) + j][tid >> 2] = (uint8_t)(rsa >> (8 * j));
        }
        if (tid < 2 * NKB) sBs[tid / NKB][tid % NKB] = rsb;
    };

    f32x16 acc[BPW], gacc[GB];
#pragma unroll
    for (int q = 0; q < BPW; q++)
        for (int v = 0; v < 16; v++) acc[q][v] = 0.f;
#pragma unroll
    for (int q = 0; q < GB; q++)
        for (int v = 0; v < 16; v++) gacc[q][v] = 0.f;

    auto compute = [&](auto gram_tag) __attribute__((always_inline)) {
        constexpr int GR = decltype(gram_tag)::value;  // 0 no gram, 1 on tile 0 by wave, 2 spread
#pragma unroll
        for (int g = 0; g < K8 / 64; g++) {
#pragma unroll
            for (int q = 0; q < BPW; q++) {
                const int blk = wave * BPW + q;
                if (NBLK % 4 == 0 || blk < NBLK) {
                    const int mb = blk >> 1, nb = blk & 1;
                    const uint8_t* pa = &sA[mb * 32 + c][g * 64 + h * 16];
                    const uint8_t* pb = &sB[nb * 32 + c][g * 64 + h * 16];
                    acc[q] = mfma_fp8(*reinterpret_cast<const u32x4*>(pa), *reinterpret_cast<const u32x4*>(pa + 32),
                                      *reinterpret_cast<const u32x4*>(pb), *reinterpret_cast<const u32x4*>(pb + 32),
                                      sAs[2 * g + h][mb * 32 + c], sBs[nb][2 * g + h], acc[q]);
                }
            }
            if constexpr (GR == 1) {
                {
#pragma unroll
                    for (int q = 0; q < GB; q++) {
                        const int blk = wave + 4 * q;
                        if ((MB * MB) % 4 == 0 || blk < MB * MB) {
                            const int mb = blk / MB, nb = blk - (blk / MB) * MB;
                            const uint8_t* pa = &sA[mb * 32 + c][g * 64 + h * 16];
                            const uint8_t* pb = &sA[nb * 32 + c][g * 64 + h * 16];
                            gacc[q] = mfma_fp8(*reinterpret_cast<const u32x4*>(pa),
                                               *reinterpret_cast<const u32x4*>(pa + 32),
                                               *reinterpret_cast<const u32x4*>(pb),
                                               *reinterpret_cast<const u32x4*>(pb + 32), sAs[2 * g + h][mb * 32 + c],
                                               sAs[2 * g + h][nb * 32 + c], gacc[q]);
                        }
                    }
                }
            }
        }
        if constexpr (GR == 2) {
            // this wave's k steps g = wave + 4 i of gram block t
            const int mb = t / MB, nb = t - (t / MB) * MB;
#pragma unroll
            for (int i = 0; i < K8 / 64 / 4; i++) {
                const int g = wave + 4 * i;
                const uint8_t* pa = &sA[mb * 32 + c][g * 64 + h * 16];
                const uint8_t* pb = &sA[nb * 32 + c][g * 64 + h * 16];
                gacc[0] = mfma_fp8(*reinterpret_cast<const u32x4*>(pa), *reinterpret_cast<const u32x4*>(pa + 32),
                                   *reinterpret_cast<const u32x4*>(pb), *reinterpret_cast<const u32x4*>(pb + 32),
                                   sAs[2 * g + h][mb * 32 + c], sAs[2 * g + h][nb * 32 + c], gacc[0]);
            }
        }
    };

    const int nchunks = (kend - kbeg) / K8;  // whole chunks (padded operands)
    auto mainloop = [&](auto gram_tag) __attribute__((always_inline)) {
        gload(kbeg);
        sstore();
        __syncthreads();
        for (int ch = 0; ch < nchunks; ch++) {
            const bool more = ch + 1 < nchunks;
            if (more) gload(kbeg + (ch + 1) * K8);
            compute(gram_tag);
            __syncthreads();
            if (more) {
                sstore();
                __syncthreads();
            }
        }
    };
    if (gram && spread)
        mainloop(std::integral_constant<int, 2>{});
    else if (gram)
        mainloop(std::integral_constant<int, 1>{});
    else
        mainloop(std::integral_constant<int, 0>{});

    // acc[q][v] = P(a = mb*32 + acc_row(v,h), u = nb*32 + c) for block wave*BPW + q
    float* slice = p.P + (size_t)s * p.pstride;
#pragma unroll
    for (int q = 0; q < BPW; q++) {
        const int blk = wave * BPW + q;
        if (blk >= NBLK) continue;
        const int mb = blk >> 1, ul = (blk & 1) * 32 + c;
        if (ul >= ncols) continue;
        float* dst = slice + (size_t)(col0 + ul) * p.ldr;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int a0 = mb * 32 + g * 8 + h * 4;
            if (a0 < p.ldr)
                *reinterpret_cast<float4*>(dst + a0) =
                    make_float4(acc[q][g * 4], acc[q][g * 4 + 1], acc[q][g * 4 + 2], acc[q][g * 4 + 3]);
        }
    }
    if (gram && spread) {
        // the 4 waves' parts of gram block t = (mb, nb) meet in the chunks' LDS (free: the
        // main loop ends with a barrier after the last multiply) and are added in wave order
        static_assert(4 * 16 * 64 * sizeof(float) <= sizeof(sB), "gram partials fit in sB");
        float* red = reinterpret_cast<float*>(&sB[0][0]);
#pragma unroll
        for (int v = 0; v < 16; v++) red[(wave * 16 + v) * 64 + lane] = gacc[0][v];
        __syncthreads();
        const int mb = t / MB, nb = t - (t / MB) * MB;
        float* gdst = slice + (size_t)p.N * p.ldr;
        for (int e = tid; e < 16 * 64; e += NTHREADS) {
            const int v = e >> 6, ln = e & 63;
            float sum = red[v * 64 + ln];
#pragma unroll
            for (int w2 = 1; w2 < 4; w2++) sum += red[(w2 * 16 + v) * 64 + ln];
            const int a = mb * 32 + acc_row(v, ln >> 5), b = nb * 32 + (ln & 31);
            if (b < p.r && a < p.ldr) gdst[(size_t)b * p.ldr + a] = sum;
        }
    } else if (gram) {
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
    }
    if (gram && !p.gkeep)  // S slices of MB * MB (spread) or one gram workgroup(s) each
        gram_finish4<(MB >= 2 ? 8 : 2)>(p.P + (size_t)p.N * p.ldr, p.pstride, p.S, spread ? p.S * MB * MB : p.S,
                                        p.r * p.ldr, p.G, p.gcount);
}

// ---------------------------------------------------------------------------
// V (n x m, FP32) -> Vq (contraction i, m columns) and Vtq (contraction u, n columns), both
// tiled (fp8_tiled_index with ntm / ntn column tiles), and one E8M0 scale per 32 x 32 tile:
// Vsc[ub * nbi + ib] (nbi = ldv / 32).  One workgroup per tile of the padded stores (ldv x
// ldt, the padded contraction lengths); tiles past V's edges get zeros and the scale 127 (2^0).
__global__ __launch_bounds__(NTHREADS) void k_pack_v_fp8(const float* __restrict__ V, int n, int m,
                                                         uint8_t* __restrict__ Vq, int ntm, uint8_t* __restrict__ Vtq,
                                                         int ntn, uint8_t* __restrict__ Vsc, int nbi) {
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
        if (u < ntm * 64) Vq[fp8_tiled_index(i, u, ntm)] = f32_to_e4m3(e8m0_unscale(tile[uu][tx], sc));
    }
    for (int ii = ty; ii < 32; ii += 8) {
        const int u = u0 + tx, i = i0 + ii;
        if (i < ntn * 64) Vtq[fp8_tiled_index(u, i, ntn)] = f32_to_e4m3(e8m0_unscale(tile[tx][ii], sc));
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
    if (a.snt_k0 < a.K)
        hipLaunchKernelGGL((k_pass_tn_fp8<MB, true>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS), 0, st, a);
    else
        hipLaunchKernelGGL((k_pass_tn_fp8<MB, false>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS), 0, st, a);
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
    // ldv, ldt: the padded contraction lengths, whole 512-byte chunks (covering the padded
    // column counts 64 ceil(n/64), 64 ceil(m/64) as well)
    if (ldv % 512 || ldt % 512 || ldv < n || ldt < m) return hipErrorInvalidValue;
    const int nbi = ldv / 32, nbu = ldt / 32;
    hipLaunchKernelGGL(k_pack_v_fp8, dim3(nbi, nbu), dim3(NTHREADS), 0, st, V, n, m, Vq, (m + 63) / 64, Vtq,
                       (n + 63) / 64, Vsc, nbi);
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
