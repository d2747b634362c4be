// BF16 storage path: V held in BF16 in both orientations (V and V^T, each tiled in
// blocks of 64 columns x 256 contraction elements, zero padded: bf16_tiled_index), BF16
// shadows of W and H as MFMA operands, FP32 accumulation and FP32 updates.
//
// With V^T stored, the W half of an iteration is the H half of the transposed
// problem (V^T, H^T, W^T): U_W(V, W, H) = U_H(V^T, H^T, W^T)^T.  So both
// contractions use one kernel, a K-contiguous ("TN") BF16 MFMA pass:
//   P[s](a, u) = sum_{k in slice s} A(k, a) B(k, u),
//   H pass:  A = W (n x r), B = V   -> W^T V, and the gram W^T W
//   W pass:  A = H^T (m x r), B = V^T -> H V^T = (V H^T)^T, and the gram H H^T
// Both passes stream V once at 2 bytes per element.
#include <hip/hip_runtime.h>

#include "nmf_common.hpp"
#include "nmf_internal.hpp"

#include <type_traits>

namespace gpumlib_nmf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // a native vector: staging stays in registers

constexpr int KB = 256;         // K per chunk: 512 contiguous bytes of every column
constexpr int LDB = KB + 8;     // LDS row pitch in bf16 elements (528 bytes)

__device__ __forceinline__ f32x16 mfma_bf16(u32x4 a, u32x4 b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
}

// One workgroup per (slice s, column tile t) of NT * 64 columns, 4 NT waves.  Every
// chunk is 256 elements deep; a tile's chunk is NT adjacent blocks of the tiled store,
// one contiguous read.  NT = 2 (r <= 64): every A chunk (the W or H^T shadow, re-read
// from L2 by each workgroup) serves 128 columns of V instead of 64, halving the A loads
// that otherwise take half of each CU's load traffic; one such workgroup holds the same
// 8 waves per CU as two of 4.  The chunk's B columns and A rows are staged through LDS;
// the next chunk's loads are in flight in registers while the current one is
// multiplied.  The MB * 2 NT output blocks of 32 x 32 (rank block mb, column block nb)
// are spread over the waves.
//
// The slice's gram A^T A is accumulated beside the product.  With at least MB * MB column
// tiles, gram block t = (mb, nb) rides on column tile t and its 16 k steps per chunk are
// split over the waves (step g on wave g % NW), so a gram tile does 1/NW more MFMA work
// than a plain one and no workgroup holds the one-round grid back; the waves' partial
// blocks are then added in wave order through LDS.  With fewer tiles every block rides on
// tile 0 (block mb*MB + nb on wave (mb*MB + nb) % NW, all k steps).  The last of the gram
// workgroups reduces the S slice grams into G, unless gkeep is set: then the update kernel
// sums the S partials as it stages the gram.  That reduction (a fence, an atomic and a
// round trip on one workgroup) otherwise runs after every other workgroup of a
// one-round grid has finished, some 3-4 us of pure tail per pass at 16384 x 8192.
//
// V, V^T and the shadows are padded to whole chunks with zeros, so no load needs a
// bounds check; lanes past the last column or rank row re-read the last valid one:
// those products land only in output columns that are not stored and in rank rows
// >= r of P and of the gram, which no update reads.
// SNT: the K slices that start at or after contraction index p.snt_k0 load their V chunks with
// the non-temporal policy (the part of a large V that the last-level cache cannot keep from one
// pass to the next); each workgroup runs the loop copy of its slice's policy.
template <int MB, int NT, bool SNT>
__global__ __launch_bounds__(NTHREADS * NT, 2 / NT) void k_pass_tn_bf16(LpPassArgs p) {
    constexpr int THR = NTHREADS * NT;                // threads: 4 waves per 64 columns
    constexpr int NW = THR / 64;                      // waves
    constexpr int RP = MB * 32;
    constexpr int TW = NT * TILE_LP;                  // columns per workgroup
    constexpr int NCB = TW / 32;                      // column blocks of 32
    constexpr int NBLK = MB * NCB;                    // output blocks of 32 x 32
    constexpr int BPW = NBLK >= NW ? NBLK / NW : 1;   // blocks per wave
    constexpr int GB = (MB * MB + NW - 1) / NW;       // gram blocks per wave
    constexpr int BLB = TILE_LP * KB / 8 / THR;       // 16-byte loads per thread per 64-column store block
    constexpr int BL = NT * BLB;                      // 16-byte B loads per thread per chunk
    constexpr int AL = RP * KB / 8 / THR;             // 16-byte A loads per thread per chunk
    constexpr int LPC = KB / 8;                       // lanes per column (or row) in a load
    static_assert(AL >= 1 && BLB >= 1, "load split");
    __shared__ __attribute__((aligned(16))) uint16_t sA[RP][LDB];
    __shared__ __attribute__((aligned(16))) uint16_t sB[TW][LDB];

    const int total = p.tiles * p.S;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: branches on it stay scalar
    const int c = lane & 31, h = lane >> 5;
    const bool spread = p.tiles >= MB * MB;  // grid-uniform
    const bool gram = p.G != nullptr && (spread ? t < MB * MB : t == 0);
    const int col0 = t * TW;
    const int ncols = min(TW, p.N - col0);
    const int kbeg = s * p.kslice;
    const int kend = min(round_up(p.K, KB), kbeg + p.kslice);

    u32x4 rb[BL], ra[AL];
    auto gload = [&](auto nt_tag, int k0) __attribute__((always_inline)) {
        // the chunk's NT store blocks of 64 columns (32 KB each, adjacent in the store: one
        // contiguous read): element (kq, col) of block b at b*TILE_LP*KB + col*KB + kq, so load q
        // (16 bytes) lands at sB[q / LPC][(q % LPC) * 8].  A block past the store's last (the
        // second half of a last, odd tile) re-reads the last one: those columns are not stored.
        const int nt64 = (p.N + TILE_LP - 1) / TILE_LP;
        const uint16_t* crow = p.B + (size_t)(k0 / KB) * nt64 * (TILE_LP * KB);
        auto src = [&](int j) __attribute__((always_inline)) {
            const int q = tid + THR * j, tb = j / BLB;  // store block of this load
            const int b = min(t * NT + tb, nt64 - 1);
            return reinterpret_cast<const u32x4*>(crow + (size_t)b * (TILE_LP * KB) + (size_t)(q - tb * BLB * THR) * 8);
        };
#pragma unroll
        for (int j = 0; j < BL; j++) rb[j] = decltype(nt_tag)::value ? stream_load(src(j)) : *src(j);
#pragma unroll
        for (int j = 0; j < AL; j++) {
            const int q = tid + THR * j, a = q / LPC, kq = (q % LPC) * 8;
            const int aa = a < p.r ? a : p.r - 1;
            ra[j] = *reinterpret_cast<const u32x4*>(p.A + (size_t)aa * p.lda + k0 + kq);
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the multiply
    };
    auto sstore = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < BL; j++) {
            const int q = tid + THR * j, col = q / LPC, kq = (q % LPC) * 8;
            *reinterpret_cast<u32x4*>(&sB[col][kq]) = rb[j];
        }
#pragma unroll
        for (int j = 0; j < AL; j++) {
            const int q = tid + THR * j, a = q / LPC, kq = (q % LPC) * 8;
            *reinterpret_cast<u32x4*>(&sA[a][kq]) = ra[j];
        }
    };

    f32x16 acc[BPW], gacc[GB];
#pragma unroll
    for (int q = 0; q < BPW; q++)
        for (int v = 0; v < 16; v++) acc[q][v] = 0.f;
#pragma unroll
    for (int q = 0; q < GB; q++)
        for (int v = 0; v < 16; v++) gacc[q][v] = 0.f;

    // lane (c, h) supplies k = g*16 + h*8 + j (j < 8) for both operands; the gram-carrying
    // and the plain workgroups run separate copies of the loop (no branch between MFMAs)
    auto compute = [&](auto gram_tag) __attribute__((always_inline)) {
        constexpr int GR = decltype(gram_tag)::value;  // 0 no gram, 1 on tile 0 by wave, 2 spread
#pragma unroll
        for (int g = 0; g < KB / 16; g++) {
#pragma unroll
            for (int q = 0; q < BPW; q++) {
                const int blk = wave * BPW + q;
                if (NBLK % NW == 0 || blk < NBLK) {
                    const int mb = blk / NCB, nb = blk % NCB;
                    const u32x4 av = *reinterpret_cast<const u32x4*>(&sA[mb * 32 + c][g * 16 + h * 8]);
                    const u32x4 bv = *reinterpret_cast<const u32x4*>(&sB[nb * 32 + c][g * 16 + h * 8]);
                    acc[q] = mfma_bf16(av, bv, acc[q]);
                }
            }
            if constexpr (GR == 1) {
                {
#pragma unroll
                    for (int q = 0; q < GB; q++) {
                        const int blk = wave + NW * q;
                        if ((MB * MB) % NW == 0 || blk < MB * MB) {
                            const int mb = blk / MB, nb = blk - (blk / MB) * MB;
                            const u32x4 av = *reinterpret_cast<const u32x4*>(&sA[mb * 32 + c][g * 16 + h * 8]);
                            const u32x4 bv = *reinterpret_cast<const u32x4*>(&sA[nb * 32 + c][g * 16 + h * 8]);
                            gacc[q] = mfma_bf16(av, bv, gacc[q]);
                        }
                    }
                }
            }
        }
        if constexpr (GR == 2) {
            // this wave's k steps g = wave + NW i of gram block t
            const int mb = t / MB, nb = t - (t / MB) * MB;
#pragma unroll
            for (int i = 0; i < KB / 16 / NW; i++) {
                const int g = wave + NW * i;
                const u32x4 av = *reinterpret_cast<const u32x4*>(&sA[mb * 32 + c][g * 16 + h * 8]);
                const u32x4 bv = *reinterpret_cast<const u32x4*>(&sA[nb * 32 + c][g * 16 + h * 8]);
                gacc[0] = mfma_bf16(av, bv, gacc[0]);
            }
        }
    };

    const int nchunks = (kend - kbeg) / KB;  // whole chunks (padded operands)
    auto mainloop = [&](auto gram_tag, auto nt_tag) __attribute__((always_inline)) {
        gload(nt_tag, kbeg);
        sstore();
        __syncthreads();
        for (int ch = 0; ch < nchunks; ch++) {
            const bool more = ch + 1 < nchunks;
            if (more) gload(nt_tag, kbeg + (ch + 1) * KB);
            compute(gram_tag);
            __syncthreads();
            if (more) {
                sstore();
                __syncthreads();
            }
        }
    };
    auto run = [&](auto nt_tag) __attribute__((always_inline)) {
        if (gram && spread)
            mainloop(std::integral_constant<int, 2>{}, nt_tag);
        else if (gram)
            mainloop(std::integral_constant<int, 1>{}, nt_tag);
        else
            mainloop(std::integral_constant<int, 0>{}, nt_tag);
    };
    if (SNT && kbeg >= p.snt_k0)
        run(std::bool_constant<SNT>{});  // (true here; SNT keeps the copy out of the plain kernel)
    else
        run(std::false_type{});

    // acc[q][v] = P(a = mb*32 + acc_row(v,h), u = nb*32 + c) for block wave*BPW + q
    float* slice = p.P + (size_t)s * p.pstride;
#pragma unroll
    for (int q = 0; q < BPW; q++) {
        const int blk = wave * BPW + q;
        if (blk >= NBLK) continue;
        const int mb = blk / NCB, ul = (blk % NCB) * 32 + c;
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
        // gacc[0][v] of wave w = its k steps' part of gram(a = mb*32 + acc_row(v,h), b = nb*32 + c)
        // for block t = (mb, nb).  The LDS of the chunks is free (the main loop ends with a
        // barrier after the last multiply): the NW parts meet there and are added in wave order.
        static_assert(NW * 16 * 64 * sizeof(float) <= sizeof(sB), "gram partials fit in sB");
        float* red = reinterpret_cast<float*>(&sB[0][0]);
#pragma unroll
        for (int v = 0; v < 16; v++) red[(wave * 16 + v) * 64 + lane] = gacc[0][v];
        __syncthreads();
        const int mb = t / MB, nb = t - (t / MB) * MB;
        float* gdst = slice + (size_t)p.N * p.ldr;
        for (int e = tid; e < 16 * 64; e += THR) {
            const int v = e >> 6, ln = e & 63;
            float sum = red[v * 64 + ln];
#pragma unroll
            for (int w2 = 1; w2 < NW; w2++) sum += red[(w2 * 16 + v) * 64 + ln];
            const int a = mb * 32 + acc_row(v, ln >> 5), b = nb * 32 + (ln & 31);
            if (b < p.r && a < p.ldr) gdst[(size_t)b * p.ldr + a] = sum;
        }
    } else if (gram) {
        // gacc[q][v] = gram(a = mb*32 + acc_row(v,h), b = nb*32 + c), stored at column N + b
#pragma unroll
        for (int q = 0; q < GB; q++) {
            const int blk = wave + NW * q;
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
// V (n x m, FP32 column-major) -> Vb (contraction i, columns u) and Vt (contraction u,
// columns i), both BF16 (round to nearest even) in the tiled layout of bf16_tiled_index.
// 64 x 64 tiles through LDS; the grid covers i < ldv and u < ldt (the padded contraction
// lengths, multiples of 256, which also cover the padded column counts): padding gets zeros.
__global__ __launch_bounds__(NTHREADS) void k_pack_v_bf16(const float* __restrict__ V, int n, int m,
                                                          uint16_t* __restrict__ Vb, int ntm,
                                                          uint16_t* __restrict__ Vt, int ntn) {
    __shared__ uint16_t tile[64][66];
    const int i0 = blockIdx.x * 64, u0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int uu = ty; uu < 64; uu += 4) {
        const int i = i0 + tx, u = u0 + uu;
        const uint16_t b = (i < n && u < m) ? __builtin_bit_cast(uint16_t, (__bf16)V[(size_t)u * n + i]) : 0;
        if (u < 64 * ntm) Vb[bf16_tiled_index(i, u, ntm)] = b;
        tile[uu][tx] = b;
    }
    __syncthreads();
    for (int ii = ty; ii < 64; ii += 4) {
        const int u = u0 + tx, i = i0 + ii;
        if (i < 64 * ntn) Vt[bf16_tiled_index(u, i, ntn)] = tile[tx][ii];
    }
}

// W (n x r, FP32) -> Wb (column-major, ld ldw);  H (r x m, FP32) -> Ht (H^T column-major, ld ldh).
// Padding (i in [n, ldw), u in [m, ldh)) gets zeros.
__global__ __launch_bounds__(NTHREADS) void k_shadow_bf16(const float* __restrict__ W, const float* __restrict__ H,
                                                          int n, int m, int r, uint16_t* __restrict__ Wb, int ldw,
                                                          uint16_t* __restrict__ Ht, int ldh) {
    const long nw = (long)ldw * r, nh = (long)r * ldh;
    for (long e = (long)blockIdx.x * NTHREADS + threadIdx.x; e < nw + nh; e += (long)gridDim.x * NTHREADS) {
        if (e < nw) {
            const long a = e / ldw, i = e - a * ldw;
            Wb[e] = i < n ? __builtin_bit_cast(uint16_t, (__bf16)W[a * n + i]) : (uint16_t)0;
        } else {
            const long f = e - nw, a = f / ldh, u = f - a * ldh;
            Ht[f] = u < m ? __builtin_bit_cast(uint16_t, (__bf16)H[u * r + a]) : (uint16_t)0;
        }
    }
}

// ---------------------------------------------------------------------------
template <int MB, int NT>
static void launch_tn(const LpPassArgs& a, hipStream_t st) {
    if (a.snt_k0 < a.K)
        hipLaunchKernelGGL((k_pass_tn_bf16<MB, NT, true>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS * NT), 0, st, a);
    else
        hipLaunchKernelGGL((k_pass_tn_bf16<MB, NT, false>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS * NT), 0, st,
                           a);
}

hipError_t launch_pass_tn_bf16(const LpPassArgs& a, hipStream_t st) {
    // 128-column tiles for r <= 64 (as the plan chooses them), 64-column tiles above
    const int rp = round_up(a.r, 32), nt = a.nt == 2 ? 2 : 1;
    if (rp <= 32 && nt == 2)
        launch_tn<1, 2>(a, st);
    else if (rp <= 64 && nt == 2)
        launch_tn<2, 2>(a, st);
    else if (rp > 64 && rp <= 128 && nt == 1)
        launch_tn<4, 1>(a, st);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pack_v_bf16(const float* V, int n, int m, uint16_t* Vb, int ldv, uint16_t* Vt, int ldt,
                              hipStream_t st) {
    // ldv, ldt: the padded contraction lengths, whole 256-element chunks (covering the
    // padded column counts 64 ceil(n/64), 64 ceil(m/64) as well)
    if (ldv % 256 || ldt % 256 || ldv < n || ldt < m) return hipErrorInvalidValue;
    dim3 grid(ldv / 64, ldt / 64);
    hipLaunchKernelGGL(k_pack_v_bf16, grid, dim3(NTHREADS), 0, st, V, n, m, Vb, (m + 63) / 64, Vt, (n + 63) / 64);
    return hipGetLastError();
}

hipError_t launch_shadow_bf16(const float* W, const float* H, int n, int m, int r, uint16_t* Wb, int ldw, uint16_t* Ht,
                              int ldh, hipStream_t st) {
    const long total = (long)ldw * r + (long)r * ldh;
    long blocks = (total + NTHREADS - 1) / NTHREADS;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_shadow_bf16, dim3((unsigned)blocks), dim3(NTHREADS), 0, st, W, H, n, m, r, Wb, ldw, Ht, ldh);
    return hipGetLastError();
}

}  // namespace gpumlib_nmf
