// FP32X storage path: FP32-accurate products on the BF16 matrix cores.
//
// V is held in FP32 in both orientations, each tiled for the pass that streams it
// (blocks of 64 columns x 128 contraction elements, 32 KB: x3_tiled_index), and every
// MFMA operand is a three-way BF16 split, x = x1 + x2 + x3 (bf16x2_split3).  A product
// a b is the sum of the six piece products a_p b_q with p + q <= 4, each exact in FP32;
// what is dropped (a2 b3, a3 b2, a3 b3 and the split remainders) stays below 2^-25 |a b|.
// a1 b1 is accumulated in one FP32 accumulator and the five smaller products in a second,
// added at the end: the sums round as an FP32 accumulation of exact products would, at
// six BF16 MFMAs per 16-deep K step, where FP32 MFMAs would take the time of twelve.
//
// As in the BF16 path, the W half of an iteration is the H half of the transposed
// problem, so one K-contiguous ("TN") pass serves both:
//   P[s](a, u) = sum_{k in slice s} A(k, a) B(k, u)
//   H pass:  A = W (split shadow, n x r), B = V   -> W^T V, and the gram W^T W
//   W pass:  A = H^T (split shadow, m x r), B = V^T -> H V^T, and the gram H H^T
// The A operand comes pre-split (three BF16 planes written by the update kernels); the
// chunk of B is split as it is moved from registers to LDS, once per element.
#include <hip/hip_runtime.h>

#include "nmf_common.hpp"
#include "nmf_internal.hpp"

#include <type_traits>

namespace gpumlib_nmf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int KX = 128;      // K per chunk: 512 contiguous bytes of every FP32 column
constexpr int LDH = KX + 8;  // LDS pitch of a row of one piece plane, bf16 (272 bytes)
constexpr int X3_THREADS = 512;  // 8 waves: the workgroup's LDS (up to 153 KB) fills a CU

__device__ __forceinline__ f32x16 mfma_bf16(u32x4 a, u32x4 b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
}

// Buffer resource over `bytes` bytes from base (raw buffer, 32-bit offsets): a load's
// address is then the wave-uniform base plus one 32-bit lane offset, so a chunk's loads
// keep one offset register per lane instead of a 64-bit address each.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float4 as_float4(u32x4 x) {
    return make_float4(__uint_as_float(x[0]), __uint_as_float(x[1]), __uint_as_float(x[2]), __uint_as_float(x[3]));
}

// The six piece products of one 16-deep K step into (big, small): the smaller ones first.
__device__ __forceinline__ void mma6(const u32x4 (&a)[3], const u32x4 (&b)[3], f32x16& big, f32x16& small) {
    small = mfma_bf16(a[0], b[1], small);
    small = mfma_bf16(a[1], b[0], small);
    small = mfma_bf16(a[0], b[2], small);
    small = mfma_bf16(a[1], b[1], small);
    small = mfma_bf16(a[2], b[0], small);
    big = mfma_bf16(a[0], b[0], big);
}

// One workgroup of 8 waves per (slice s, column tile t) of NT * 64 columns: per 128-deep
// chunk, the tile's B columns (FP32, split on the way into three BF16 planes in LDS) and
// the A rows (three BF16 planes) are staged through LDS, the next chunk's loads in flight
// in registers while the current one is multiplied; the MB * 2 NT output blocks of 32 x 32
// go one per wave (MB = 1: four waves stage and split only); gram block t = (t / MB,
// t % MB) rides on column tile t < MB * MB (the plan gives every pass that many tiles),
// its k steps split over the waves, and the last gram workgroup reduces the S slice grams
// unless gkeep leaves the slice partials to the update.  The store and
// the shadows are zero padded to whole chunks and 64-column tiles, so no load needs a
// bounds check; lanes past the last rank row re-read the last valid one (their products
// land only in rank rows >= r of P and of the gram, which nothing reads).
template <int MB, int NT>
__global__ __launch_bounds__(X3_THREADS, 1) void k_pass_tn_x3(LpPassArgs p) {
    constexpr int THR = X3_THREADS;
    constexpr int NW = THR / 64;
    constexpr int RP = MB * 32;
    constexpr int TW = NT * TILE_LP;
    constexpr int NCB = TW / 32;
    constexpr int NBLK = MB * NCB;
    constexpr int BPW = NBLK >= NW ? NBLK / NW : 1;
    constexpr int BLB = TILE_LP * KX / 4 / THR;      // 16-byte B loads per thread per 64-column store block
    constexpr int BL = NT * BLB;                     // per chunk
    constexpr int LPC = KX / 4;                      // B: lanes per column in a load (4 floats each)
    constexpr int LPR = KX / 8;                      // A: lanes per row of one plane in a load (8 bf16 each)
    constexpr int AL = 3 * RP * LPR / THR;           // 16-byte A loads per thread per chunk (three planes)
    static_assert(BLB >= 1 && AL >= 1 && (3 * RP * LPR) % THR == 0, "load split");
    __shared__ __attribute__((aligned(16))) uint16_t sB[3][TW][LDH];
    __shared__ __attribute__((aligned(16))) uint16_t sA[3][RP][LDH];

    const int total = p.tiles * p.S;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, h = lane >> 5;
    const bool gram = p.G != nullptr && t < MB * MB;  // tiles >= MB * MB (the launcher checks)
    const int col0 = t * TW;
    const int ncols = min(TW, p.N - col0);
    const int kbeg = s * p.kslice;
    const int kend = min(round_up(p.K, KX), kbeg + p.kslice);
    const float* Bf = reinterpret_cast<const float*>(p.B);

    // A: the three planes from p.A (their byte size fits 32 bits: the launcher checks); lane
    // offsets of this thread's AL loads (row clamped to r - 1), the chunk start in soffset
    const __amdgpu_buffer_rsrc_t arsrc = buffer_rsrc(p.A, (uint32_t)(3 * p.aplane * 2));
    uint32_t aoff[AL];
#pragma unroll
    for (int j = 0; j < AL; j++) {
        const int q = tid + THR * j, pl = q / (RP * LPR), rem = q - pl * (RP * LPR);
        const int a = rem / LPR, kq = (rem - a * LPR) * 8;
        const int aa = a < p.r ? a : p.r - 1;
        aoff[j] = (uint32_t)(((size_t)pl * p.aplane + (size_t)aa * p.lda + kq) * 2);
    }
    const int nt64 = (p.N + TILE_LP - 1) / TILE_LP;

    u32x4 rb[BL], ra[AL];
    auto gload = [&](int k0) __attribute__((always_inline)) {
        // the chunk's NT store blocks of 64 columns (32 KB each, adjacent): element (kq, col)
        // of block b at b*64*128 + col*128 + kq, so load q (16 bytes) of a block holds column
        // q / LPC, k = (q % LPC) * 4 ...+3.  A block past the store's last (the second half of
        // a last, odd tile) re-reads the last one: those columns are not stored.
        const float* crow = Bf + (size_t)(k0 / KX) * nt64 * (TILE_LP * KX);
#pragma unroll
        for (int tb = 0; tb < NT; tb++) {
            const int b = min(t * NT + tb, nt64 - 1);
            const __amdgpu_buffer_rsrc_t brsrc = buffer_rsrc(crow + (size_t)b * (TILE_LP * KX), TILE_LP * KX * 4);
#pragma unroll
            for (int j = 0; j < BLB; j++)
                rb[tb * BLB + j] = __builtin_amdgcn_raw_buffer_load_b128(brsrc, tid * 16, THR * 16 * j, 0);
        }
#pragma unroll
        for (int j = 0; j < AL; j++) ra[j] = __builtin_amdgcn_raw_buffer_load_b128(arsrc, aoff[j], k0 * 2, 0);
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the multiply
    };
    auto sstore = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < BL; j++) {
            const int q = tid + THR * j, col = q / LPC, kq = (q % LPC) * 4;
            const float4 x = as_float4(rb[j]);
            uint32_t p1[2], p2[2], p3[2];
            bf16x2_split3(x.x, x.y, p1[0], p2[0], p3[0]);
            bf16x2_split3(x.z, x.w, p1[1], p2[1], p3[1]);
            *reinterpret_cast<uint2*>(&sB[0][col][kq]) = make_uint2(p1[0], p1[1]);
            *reinterpret_cast<uint2*>(&sB[1][col][kq]) = make_uint2(p2[0], p2[1]);
            *reinterpret_cast<uint2*>(&sB[2][col][kq]) = make_uint2(p3[0], p3[1]);
        }
#pragma unroll
        for (int j = 0; j < AL; j++) {
            const int q = tid + THR * j, pl = q / (RP * LPR), rem = q - pl * (RP * LPR);
            const int a = rem / LPR, kq = (rem - a * LPR) * 8;
            *reinterpret_cast<u32x4*>(&sA[pl][a][kq]) = ra[j];
        }
    };

    f32x16 big[BPW], small[BPW], gbig, gsmall;
#pragma unroll
    for (int q = 0; q < BPW; q++)
        for (int v = 0; v < 16; v++) big[q][v] = small[q][v] = 0.f;
    for (int v = 0; v < 16; v++) gbig[v] = gsmall[v] = 0.f;

    // lane (c, h) supplies k = g*16 + h*8 + j (j < 8) of row / column c of its block
    auto apieces = [&](int row, int k, u32x4 (&x)[3]) __attribute__((always_inline)) {
#pragma unroll
        for (int pl = 0; pl < 3; pl++) x[pl] = *reinterpret_cast<const u32x4*>(&sA[pl][row][k]);
    };
    auto bpieces = [&](int col, int k, u32x4 (&x)[3]) __attribute__((always_inline)) {
#pragma unroll
        for (int pl = 0; pl < 3; pl++) x[pl] = *reinterpret_cast<const u32x4*>(&sB[pl][col][k]);
    };
    // the operands of k step g + 1 are read while step g multiplies; the scheduling
    // barrier after each step keeps the reads of later steps (and their registers) out
    auto compute = [&](auto gram_tag) __attribute__((always_inline)) {
        constexpr bool GR = decltype(gram_tag)::value;
        u32x4 av[2][BPW][3], bv[2][BPW][3];
        auto fetch = [&](int g, u32x4 (&a)[BPW][3], u32x4 (&b)[BPW][3]) __attribute__((always_inline)) {
            const int k = g * 16 + h * 8;
#pragma unroll
            for (int q = 0; q < BPW; q++) {
                const int blk = (wave * BPW + q) % NBLK;  // waves past the last block repeat one (not stored)
                apieces((blk / NCB) * 32 + c, k, a[q]);
                bpieces((blk % NCB) * 32 + c, k, b[q]);
            }
        };
        const bool active = NBLK % NW == 0 || wave * BPW < NBLK;
        if (active) {
            fetch(0, av[0], bv[0]);
#pragma unroll
            for (int g = 0; g < KX / 16; g++) {
                if (g + 1 < KX / 16) fetch(g + 1, av[(g + 1) & 1], bv[(g + 1) & 1]);
#pragma unroll
                for (int q = 0; q < BPW; q++) mma6(av[g & 1][q], bv[g & 1][q], big[q], small[q]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if constexpr (GR) {
            // this wave's k steps g = wave + NW i of gram block t
            const int mb = t / MB, nb = t - (t / MB) * MB;
#pragma unroll
            for (int i = 0; i < (KX / 16 + NW - 1) / NW; i++) {
                const int g = wave + NW * i;
                if ((KX / 16) % NW == 0 || g < KX / 16) {
                    const int k = g * 16 + h * 8;
                    u32x4 ga[3], gb[3];
                    apieces(mb * 32 + c, k, ga);
                    apieces(nb * 32 + c, k, gb);
                    mma6(ga, gb, gbig, gsmall);
                }
            }
        }
    };

    const int nchunks = (kend - kbeg) / KX;  // whole chunks (padded operands)
    auto mainloop = [&](auto gram_tag) __attribute__((always_inline)) {
        gload(kbeg);
        sstore();
        __syncthreads();
        for (int ch = 0; ch < nchunks; ch++) {
            const bool more = ch + 1 < nchunks;
            if (more) gload(kbeg + (ch + 1) * KX);
            compute(gram_tag);
            __syncthreads();
            if (more) {
                sstore();
                __syncthreads();
            }
        }
    };
    if (gram)
        mainloop(std::true_type{});
    else
        mainloop(std::false_type{});

    // big[q][v] + small[q][v] = P(a = mb*32 + acc_row(v,h), u = nb*32 + c) for block wave*BPW + q
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
                    make_float4(big[q][g * 4] + small[q][g * 4], big[q][g * 4 + 1] + small[q][g * 4 + 1],
                                big[q][g * 4 + 2] + small[q][g * 4 + 2], big[q][g * 4 + 3] + small[q][g * 4 + 3]);
        }
    }
    if (gram) {
        // the waves' parts of gram block t = (mb, nb) meet in the (free) chunk LDS and are
        // added in wave order
        static_assert(NW * 16 * 64 * sizeof(float) <= sizeof(sB), "gram partials fit in sB");
        float* red = reinterpret_cast<float*>(&sB[0][0][0]);
#pragma unroll
        for (int v = 0; v < 16; v++) red[(wave * 16 + v) * 64 + lane] = gbig[v] + gsmall[v];
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
    }
    if (gram && !p.gkeep)
        gram_finish4<(MB >= 2 ? 8 : 2)>(p.P + (size_t)p.N * p.ldr, p.pstride, p.S, p.S * MB * MB, p.r * p.ldr, p.G,
                                        p.gcount);
}

// V (n x m, FP32 column-major) -> Vx (contraction i, columns u) and Vtx (contraction u,
// columns i) in the tiled layout of x3_tiled_index; 64 x 64 tiles through LDS, the grid
// covering the padded extents (multiples of 128, which also cover the padded column
// counts): padding gets zeros.
__global__ __launch_bounds__(NTHREADS) void k_pack_v_x3(const float* __restrict__ V, int n, int m,
                                                        float* __restrict__ Vx, int ntm, float* __restrict__ Vtx,
                                                        int ntn) {
    __shared__ float tile[64][65];
    const int i0 = blockIdx.x * 64, u0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int uu = ty; uu < 64; uu += 4) {
        const int i = i0 + tx, u = u0 + uu;
        const float x = (i < n && u < m) ? V[(size_t)u * n + i] : 0.f;
        if (u < 64 * ntm) Vx[x3_tiled_index(i, u, ntm)] = x;
        tile[uu][tx] = x;
    }
    __syncthreads();
    for (int ii = ty; ii < 64; ii += 4) {
        const int u = u0 + tx, i = i0 + ii;
        if (i < 64 * ntn) Vtx[x3_tiled_index(u, i, ntn)] = tile[tx][ii];
    }
}

// W (n x r) -> Wp: three BF16 planes of W, column-major (W(i,a) at [p*r*ldw + a*ldw + i]);
// H (r x m) -> Hp: three planes of H^T (H(a,u) at [p*r*ldh + a*ldh + u]); padding zeros.
__global__ __launch_bounds__(NTHREADS) void k_shadow_x3(const float* __restrict__ W, const float* __restrict__ H,
                                                        int n, int m, int r, uint16_t* __restrict__ Wp, int ldw,
                                                        uint16_t* __restrict__ Hp, int ldh) {
    // two elements per thread (ldw, ldh are even)
    const long nw = (long)ldw * r / 2, nh = (long)r * ldh / 2;
    for (long e = (long)blockIdx.x * NTHREADS + threadIdx.x; e < nw + nh; e += (long)gridDim.x * NTHREADS) {
        float x0, x1;
        uint32_t* base;
        long plane, off;
        if (e < nw) {
            const long a = (2 * e) / ldw, i = 2 * e - a * ldw;
            x0 = i < n ? W[a * n + i] : 0.f;
            x1 = i + 1 < n ? W[a * n + i + 1] : 0.f;
            base = reinterpret_cast<uint32_t*>(Wp), plane = nw, off = e;
        } else {
            const long f = e - nw, a = (2 * f) / ldh, u = 2 * f - a * ldh;
            x0 = u < m ? H[u * r + a] : 0.f;
            x1 = u + 1 < m ? H[(u + 1) * r + a] : 0.f;
            base = reinterpret_cast<uint32_t*>(Hp), plane = nh, off = f;
        }
        uint32_t w1, w2, w3;
        bf16x2_split3(x0, x1, w1, w2, w3);
        base[off] = w1;
        base[plane + off] = w2;
        base[2 * plane + off] = w3;
    }
}

template <int MB, int NT>
void launch_tn(const LpPassArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_pass_tn_x3<MB, NT>), dim3(xcd_grid(a.tiles * a.S)), dim3(X3_THREADS), 0, st, a);
}

}  // namespace

hipError_t launch_pass_tn_x3(const LpPassArgs& a, hipStream_t st) {
    // 128-column tiles for r <= 64, 64-column tiles above (as the plan chooses them); the
    // shadows' leading dimension and the K slices are whole chunks, the planes whole shadows
    const int rp = round_up(a.r, 32), nt = a.nt == 2 ? 2 : 1;
    if (a.lda % KX || a.lda < a.K || a.aplane < (long)a.r * a.lda || a.kslice % KX) return hipErrorInvalidValue;
    if (a.G && a.tiles < (rp / 32) * (rp / 32)) return hipErrorInvalidValue;  // one gram block per tile
    if (3 * a.aplane * 2 > 0xffffffffL) return hipErrorInvalidValue;           // 32-bit buffer offsets
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

hipError_t launch_pack_v_x3(const float* V, int n, int m, float* Vx, int ldv, float* Vtx, int ldt, hipStream_t st) {
    // ldv, ldt: the padded contraction lengths, whole 128-element chunks
    if (ldv % 128 || ldt % 128 || ldv < n || ldt < m) return hipErrorInvalidValue;
    dim3 grid(ldv / 64, ldt / 64);
    hipLaunchKernelGGL(k_pack_v_x3, grid, dim3(NTHREADS), 0, st, V, n, m, Vx, (m + 63) / 64, Vtx, (n + 63) / 64);
    return hipGetLastError();
}

hipError_t launch_shadow_x3(const float* W, const float* H, int n, int m, int r, uint16_t* Wp, int ldw, uint16_t* Hp,
                            int ldh, hipStream_t st) {
    if (ldw % 2 || ldh % 2 || ldw < n || ldh < m) return hipErrorInvalidValue;
    const long total = ((long)ldw * r + (long)r * ldh) / 2;
    long blocks = (total + NTHREADS - 1) / NTHREADS;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_shadow_x3, dim3((unsigned)blocks), dim3(NTHREADS), 0, st, W, H, n, m, r, Wp, ldw, Hp, ldh);
    return hipGetLastError();
}

}  // namespace gpumlib_nmf
