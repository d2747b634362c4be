// Objective pass: sum over V's elements of a per-element term of (V(i,u), (W H)(i,u)),
// accumulated in FP64 (Frobenius (v - wh)^2, KL v ln(v / wh) - v + wh, and the
// reference quality v log10(wh + eps) - wh of nmf_kernels_quality.cu:35), for V
// stored in FP32 (plain, or tiled as the split-BF16 passes read it), BF16 or FP8.
//
// W H is formed on BF16 MFMAs from a three-way BF16 split of the FP32 factors:
// x = x1 + x2 + x3 with x1 = bf16(x), x2 = bf16(x - x1), x3 = bf16(x - x1 - x2)
// (round to nearest even; each difference is exact in FP32, |x2| <= 2^-9 |x| and
// |x3| <= 2^-18 |x|).  The six products x_p y_q with p + q <= 4 (one-based) are
// exact in FP32; the dropped ones and the split's remainders stay below 2^-25 of each
// product.  The five small products of all K steps are accumulated first and the
// x1 y1 products on top of their sum, so that every rounding of the x1 y1 chain is
// relative to the running W H as in an FP32 accumulation: W H to FP32 accuracy, at
// 6 x 16-deep K steps per 32 x 32 block on the BF16 matrix cores, a third of the
// matrix-core time of 32 FP32 MFMA K steps.

// Layout of the split factors (one launch of k_split_wh before the pass):
//   Ws: piece p of W^T, row i at Ws + (p * nrows + i) * rp, a contiguous
//   Hs: piece p of H,   column u at Hs + (p * ncols + u) * rp, a contiguous
// with rp = split_rank(r) (16, 32, 64 or 128), nrows = round_up(n, 128), ncols =
// round_up(m, 32) and zeros in the padding, so the MFMA operands (8 consecutive a of one row or
// column) are single 16-byte loads.
//
// One workgroup = 4 waves over a block of 128 rows (a 32-row strip per wave, its W
// pieces held in registers) and a range of 32-column blocks: per column block the
// H pieces (3 x 32 columns x rp) are staged through double-buffered LDS and shared
// by the waves, each wave's V block (32 x 32, 16 elements per lane in the MFMA
// result layout) is loaded one block ahead in registers.
#include <hip/hip_runtime.h>

#include "nmf_common.hpp"
#include "nmf_fp8.hpp"
#include "nmf_internal.hpp"
#include "nmf_plan.hpp"

namespace gpumlib_nmf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x16 mfma_bf16(u32x4 a, u32x4 b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
}

__device__ __forceinline__ uint16_t bf16_bits(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__device__ __forceinline__ float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Three-way split of 8 floats into 8 BF16 pieces each (packed pairs, low half first).
__device__ __forceinline__ void split8(const float (&x)[8], u32x4& p1, u32x4& p2, u32x4& p3) {
    uint32_t w1[4], w2[4], w3[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint16_t b1[2], b2[2], b3[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const float v = x[2 * j + e];
            b1[e] = bf16_bits(v);
            const float r1 = v - bf16_value(b1[e]);
            b2[e] = bf16_bits(r1);
            const float r2 = r1 - bf16_value(b2[e]);
            b3[e] = bf16_bits(r2);
        }
        w1[j] = (uint32_t)b1[0] | ((uint32_t)b1[1] << 16);
        w2[j] = (uint32_t)b2[0] | ((uint32_t)b2[1] << 16);
        w3[j] = (uint32_t)b3[0] | ((uint32_t)b3[1] << 16);
    }
    p1 = u32x4{w1[0], w1[1], w1[2], w1[3]};
    p2 = u32x4{w2[0], w2[1], w2[2], w2[3]};
    p3 = u32x4{w3[0], w3[1], w3[2], w3[3]};
}

// W (n x r, column-major) and H (r x m, column-major) -> their split pieces (rp = RP).
// Workgroups [0, nrows / 64) take 64 rows of W each (transposed through LDS), the rest 64
// columns of H each.
template <int RP>
__global__ __launch_bounds__(NTHREADS) void k_split_wh(const float* __restrict__ W, const float* __restrict__ H,
                                                       int n, int m, int r, int nrows, int ncols,
                                                       uint16_t* __restrict__ Ws, uint16_t* __restrict__ Hs) {
    constexpr int rp = RP, upr = RP / 8;  // 8-element groups per row / column
    __shared__ float t[RP][65];
    const int tid = threadIdx.x;
    const int wblocks = nrows / 64;
    if ((int)blockIdx.x < wblocks) {
        const int i0 = blockIdx.x * 64;
        float x[RP / 4];
#pragma unroll
        for (int q = 0; q < RP / 4; q++) {
            const int e = tid + q * NTHREADS, a = e >> 6, i = i0 + (e & 63);
            x[q] = (a < r && i < n) ? W[(size_t)a * n + i] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < RP / 4; q++) {
            const int e = tid + q * NTHREADS;
            t[e >> 6][e & 63] = x[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < (64 * upr + NTHREADS - 1) / NTHREADS; q++) {
            const int e = tid + q * NTHREADS;
            if (e >= 64 * upr) break;
            const int ii = e / upr, g = e - ii * upr;
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = t[g * 8 + j][ii];
            u32x4 p1, p2, p3;
            split8(x, p1, p2, p3);
            const size_t o = (size_t)(i0 + ii) * rp + g * 8;
            *reinterpret_cast<u32x4*>(Ws + o) = p1;
            *reinterpret_cast<u32x4*>(Ws + (size_t)nrows * rp + o) = p2;
            *reinterpret_cast<u32x4*>(Ws + 2 * (size_t)nrows * rp + o) = p3;
        }
    } else {
        const int u0 = (blockIdx.x - wblocks) * 64;
#pragma unroll
        for (int q = 0; q < (64 * upr + NTHREADS - 1) / NTHREADS; q++) {
            const int e = tid + q * NTHREADS;
            if (e >= 64 * upr) break;
            const int uu = e / upr, g = e - uu * upr, u = u0 + uu;
            if (u >= ncols) continue;
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int a = g * 8 + j;
                x[j] = (a < r && u < m) ? H[(size_t)u * r + a] : 0.f;
            }
            u32x4 p1, p2, p3;
            split8(x, p1, p2, p3);
            const size_t o = (size_t)u * rp + g * 8;
            *reinterpret_cast<u32x4*>(Hs + o) = p1;
            *reinterpret_cast<u32x4*>(Hs + (size_t)ncols * rp + o) = p2;
            *reinterpret_cast<u32x4*>(Hs + 2 * (size_t)ncols * rp + o) = p3;
        }
    }
}

// V block registers of one lane: rows is + 8g + 4h + j (g, j < 4) of column u.
template <int TV>
struct VBlock;
template <>
struct VBlock<0> {  // FP32, n % 4 == 0 and V 16-byte aligned
    float4 x[4];
};
template <>
struct VBlock<3> {  // FP32, any n and alignment
    float4 x[4];
};
template <>
struct VBlock<1> {  // BF16 (tiled store)
    uint2 x[4];
};
template <>
struct VBlock<4> {  // FP32 (tiled store of the split-BF16 path)
    float4 x[4];
};
template <>
struct VBlock<2> {  // FP8 E4M3 (tiled store) and the E8M0 scale of its 32 x 32 tile
    uint32_t x[4];
    float sc;
};

template <int TV>
__device__ __forceinline__ void load_vblock(const SplitObjArgs& p, int i4, int u, VBlock<TV>& vb) {
    // i4: first row of the lane's g = 0 group (is + 4h); group g starts at i4 + 8g
    if constexpr (TV == 0) {
        // n % 4 == 0 and V 16-byte aligned: a group of 4 rows is all in or all out;
        // out-of-range groups load a clamped in-range address and are zeroed (no load
        // under a branch)
        const float* V = reinterpret_cast<const float*>(p.V);
        const bool uok = u < p.m;
        const int uc = uok ? u : p.m - 1;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int i = i4 + 8 * g, ic = i < p.n ? i : p.n - 4;
            const uint4 x = *reinterpret_cast<const uint4*>(V + (size_t)uc * p.ldv + ic);
            const uint32_t keep = (uok && i < p.n) ? 0xffffffffu : 0u;  // a mask, not a branch around the load
            vb.x[g] = make_float4(__uint_as_float(x.x & keep), __uint_as_float(x.y & keep), __uint_as_float(x.z & keep),
                                  __uint_as_float(x.w & keep));
        }
    } else if constexpr (TV == 3) {
        // any other FP32 V: element-wise guarded loads
        const float* V = reinterpret_cast<const float*>(p.V);
#pragma unroll
        for (int g = 0; g < 4; g++) vb.x[g] = load4(V + (size_t)u * p.ldv + i4 + 8 * g, u < p.m, i4 + 8 * g, p.n, false);
    } else if constexpr (TV == 4) {
        // zero padded to whole 128-row chunks and 64-column tiles: no bounds; four rows of a
        // group are adjacent in a chunk column
        const float* V = reinterpret_cast<const float*>(p.V);
#pragma unroll
        for (int g = 0; g < 4; g++) vb.x[g] = *reinterpret_cast<const float4*>(V + x3_tiled_index(i4 + 8 * g, u, p.vtiles));
    } else if constexpr (TV == 1) {
        // the store is zero padded to whole 256-row chunks and 64-column tiles: no bounds
        const uint16_t* V = reinterpret_cast<const uint16_t*>(p.V);
#pragma unroll
        for (int g = 0; g < 4; g++) vb.x[g] = *reinterpret_cast<const uint2*>(V + bf16_tiled_index(i4 + 8 * g, u, p.vtiles));
    } else {
        // zero padded to whole 512-row chunks and 64-column tiles; one scale per 32 x 32 tile
        const uint8_t* V = reinterpret_cast<const uint8_t*>(p.V);
#pragma unroll
        for (int g = 0; g < 4; g++) vb.x[g] = *reinterpret_cast<const uint32_t*>(V + fp8_tiled_index(i4 + 8 * g, u, p.vtiles));
        vb.sc = e8m0_value(p.vsc[(size_t)(u >> 5) * p.nbi + (i4 >> 5)]);
    }
}

template <int TV>
__device__ __forceinline__ float vblock_elem(const VBlock<TV>& vb, int g, int j) {
    if constexpr (TV == 0 || TV == 3 || TV == 4) {
        const float4 q = vb.x[g];
        return j == 0 ? q.x : (j == 1 ? q.y : (j == 2 ? q.z : q.w));
    } else if constexpr (TV == 1) {
        const uint32_t w = (j < 2) ? vb.x[g].x : vb.x[g].y;
        return __uint_as_float((j & 1) ? (w & 0xffff0000u) : (w << 16));
    } else {
        return e4m3_to_f32((uint8_t)(vb.x[g] >> (8 * j))) * vb.sc;
    }
}

__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NTHREADS / 64; i++) t += red[i];
    return t;
}

// NK: K steps of 16 of the pieces (rp = 16 NK); TV: V element type (0 FP32 with n % 4 == 0
// and 16-byte alignment, 3 any FP32, 1 BF16 store, 2 FP8 store, 4 tiled FP32 store); OBJ: the term.
template <int NK, int TV, int OBJ>
__global__ __launch_bounds__(NTHREADS) void k_objective_split(SplitObjArgs p) {
    constexpr int RP = NK * 16;                    // the pieces' padded rank (p.rp)
    constexpr int RPL = RP + 8;                    // LDS pitch of a column's pieces (16 bytes of padding)
    constexpr int UPC = RP / 8;                    // 16-byte units per column of one piece
    constexpr int UNITS = 96 * UPC;                // 16-byte units of one column block's pieces (3 x 32 columns)
    constexpr int NQ = (UNITS + NTHREADS - 1) / NTHREADS;
    // VALU instructions of the previous block's terms placed after each MFMA (16 terms over
    // 6 NK MFMAs; the log terms are longer)
    constexpr int kValuPerMfma = (OBJ == OBJ_FROB ? 64 : 192) / (6 * NK) + 1;
    __shared__ __attribute__((aligned(16))) uint16_t sB[2][3][32][RPL];
    __shared__ double red[NTHREADS / 64];

    const int total = p.rblocks * p.csplit;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    // consecutive work items (which the XCD mapping keeps on one XCD) take consecutive row
    // blocks of one column range: its H pieces are read from HBM once per XCD and then hit
    // that XCD's L2
    const int cs = w / p.rblocks, rb = w - cs * p.rblocks;
    const int cb0 = (int)((long)cs * p.cblocks / p.csplit), cb1 = (int)((long)(cs + 1) * p.cblocks / p.csplit);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, h = lane >> 5;
    const int is = rb * 128 + wave * 32;

    // this lane's A operands: row is + c of the three W pieces, a = 16 ks + 8 h .. + 7
    u32x4 ar[3][NK];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int ks = 0; ks < NK; ks++)
            ar[q][ks] = *reinterpret_cast<const u32x4*>(p.Ws + ((size_t)q * p.nrows + is + c) * RP + ks * 16 + h * 8);

    // Unit e of a column block's pieces (e < UNITS; threads past UNITS repeat units from the
    // start, writing the same values to the same LDS addresses): piece q, column col, a = 8 k8.
    // two register sets for the H pieces: a block's pieces are loaded two steps before
    // they are stored to LDS
    u32x4 rb0[NQ], rb1[NQ];
    auto gload_b = [&](u32x4 (&rb)[NQ], int cb) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int e = (tid + j * NTHREADS) % UNITS;
            const int q = e / (32 * UPC), rem = e % (32 * UPC), col = rem / UPC, k8 = rem % UPC;
            rb[j] = *reinterpret_cast<const u32x4*>(p.Hs + ((size_t)q * p.ncols + cb * 32 + col) * RP + k8 * 8);
        }
    };
    auto sstore_b = [&](const u32x4 (&rb)[NQ], int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            const int e = (tid + j * NTHREADS) % UNITS;
            const int q = e / (32 * UPC), rem = e % (32 * UPC), col = rem / UPC, k8 = rem % UPC;
            *reinterpret_cast<u32x4*>(&sB[buf][q][col][k8 * 8]) = rb[j];
        }
    };

    // Padding rows and columns hold V = 0 and W H = 0 exactly (zero pieces, zero-padded or
    // masked V): their Frobenius and KL terms are 0; the quality's log is taken of 1 there,
    // so that its term is 0 for any eps.
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // W H of the 32 x 32 block whose H pieces are in LDS buffer buf.  One accumulator: the
    // five small products of every K step first, then the x1 y1 products on top of their
    // sum, so that every rounding of the x1 y1 chain is relative to the running W H, as in
    // an FP32 accumulation of the full products.
    auto mma = [&](int buf, f32x16& wh) __attribute__((always_inline)) {
#pragma unroll
        for (int v = 0; v < 16; v++) wh[v] = 0.f;
#pragma unroll
        for (int ks = 0; ks < NK; ks++) {
            const u32x4 b1 = *reinterpret_cast<const u32x4*>(&sB[buf][0][c][ks * 16 + h * 8]);
            const u32x4 b2 = *reinterpret_cast<const u32x4*>(&sB[buf][1][c][ks * 16 + h * 8]);
            const u32x4 b3 = *reinterpret_cast<const u32x4*>(&sB[buf][2][c][ks * 16 + h * 8]);
            wh = mfma_bf16(ar[0][ks], b2, wh);
            wh = mfma_bf16(ar[1][ks], b1, wh);
            wh = mfma_bf16(ar[0][ks], b3, wh);
            wh = mfma_bf16(ar[1][ks], b2, wh);
            wh = mfma_bf16(ar[2][ks], b1, wh);
        }
#pragma unroll
        for (int ks = 0; ks < NK; ks++) {
            const u32x4 b1 = *reinterpret_cast<const u32x4*>(&sB[buf][0][c][ks * 16 + h * 8]);
            wh = mfma_bf16(ar[0][ks], b1, wh);
        }
    };
    // the terms of block cb: wh[v] = (W H)(is + acc_row(v, h), cb * 32 + c).  Frobenius: the
    // lane's 16 squared residuals summed in FP32 (four chains of 4), each block's sums added
    // in FP64; KL: every term in FP64; quality: terms in FP32, added in FP64.
    auto epi = [&](int cb, const f32x16& wh, const VBlock<TV>& vb) __attribute__((always_inline)) {
        const bool uok = cb * 32 + c < p.m;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            float sq = 0.f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float w = wh[g * 4 + j];
                const float x = vblock_elem<TV>(vb, g, j);
                if constexpr (OBJ == OBJ_QUALITY) {
                    const bool ok = uok && is + 8 * g + 4 * h + j < p.n;
                    acc[g] += (double)fmaf(x, log10f(ok ? w + p.eps : 1.f), -w);
                } else if constexpr (OBJ == OBJ_FROB) {
                    const float d = x - w;
                    sq = fmaf(d, d, sq);
                } else {
                    acc[g] += obj_term<OBJ>(x, w, p.eps);
                }
            }
            if constexpr (OBJ == OBJ_FROB) acc[g] += (double)sq;
        }
    };
    // Issue order of one pipelined step: per K step the three LDS reads of the H pieces, then
    // its six MFMAs, each followed by a few of the previous block's VALU terms, so that the
    // wave's VALU work issues while its MFMAs run.
    auto interleave = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int ks = 0; ks < NK; ks++) {
            __builtin_amdgcn_sched_group_barrier(0x100, 3, 0);  // DS read
#pragma unroll
            for (int q = 0; q < 5; q++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, kValuPerMfma, 0);  // VALU
            }
        }
#pragma unroll
        for (int ks = 0; ks < NK; ks++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, kValuPerMfma, 0);
        }
    };
    if (cb0 < cb1) {
        // Software pipeline over the column blocks k = 0 .. nsteps - 1.  Step k runs the
        // MFMAs of block k + 1 (its H pieces in LDS buffer (k + 1) & 1) interleaved with the
        // terms of block k (its product in the other accumulator, its V in registers); the V
        // of block k + 1 and the H pieces of block k + 3 are loaded at the step's start, and
        // the H pieces of block k + 2 (loaded one step earlier) are stored to LDS buffer k & 1
        // at its end.  Every step issues its loads unconditionally (past the last block they
        // re-read the last one and are not used), so that no load sits under a branch.
        const int nsteps = cb1 - cb0;
        auto blk = [&](int k) { return cb0 + (k < nsteps ? k : nsteps - 1); };
        VBlock<TV> v0, v1;
        f32x16 whA, whB;
        load_vblock<TV>(p, is + 4 * h, blk(0) * 32 + c, v0);
        gload_b(rb0, blk(0));
        sstore_b(rb0, 0);
        __syncthreads();
        gload_b(rb1, blk(1));
        gload_b(rb0, blk(2));
        __builtin_amdgcn_sched_barrier(0);
        mma(0, whA);
        __builtin_amdgcn_sched_barrier(0);
        sstore_b(rb1, 1);
        __syncthreads();
        // here (k even): block k's product in whA, its V in v0, block k + 1's H pieces in
        // buffer 1, block k + 2's in flight in rb0
        for (int k = 0; k < nsteps; k += 2) {
            gload_b(rb1, blk(k + 3));
            load_vblock<TV>(p, is + 4 * h, blk(k + 1) * 32 + c, v1);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 < nsteps) {
                mma(1, whB);
                epi(blk(k), whA, v0);
                interleave();
            } else {
                epi(blk(k), whA, v0);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 >= nsteps) break;
            sstore_b(rb0, 0);
            __syncthreads();
            gload_b(rb0, blk(k + 4));
            load_vblock<TV>(p, is + 4 * h, blk(k + 2) * 32 + c, v0);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 2 < nsteps) {
                mma(0, whA);
                epi(blk(k + 1), whB, v1);
                interleave();
            } else {
                epi(blk(k + 1), whB, v1);
            }
            __builtin_amdgcn_sched_barrier(0);
            sstore_b(rb1, 1);
            __syncthreads();
        }
    }
    const double tot = block_sum((acc[0] + acc[1]) + (acc[2] + acc[3]), red);
    if (tid == 0) p.objp[w] = tot;
}

template <int NK, int TV>
hipError_t launch_nk(const SplitObjArgs& a, int obj, hipStream_t st) {
    const dim3 grid(xcd_grid(a.rblocks * a.csplit)), block(NTHREADS);
    switch (obj) {
        case OBJ_FROB: hipLaunchKernelGGL((k_objective_split<NK, TV, OBJ_FROB>), grid, block, 0, st, a); break;
        case OBJ_KL: hipLaunchKernelGGL((k_objective_split<NK, TV, OBJ_KL>), grid, block, 0, st, a); break;
        case OBJ_QUALITY: hipLaunchKernelGGL((k_objective_split<NK, TV, OBJ_QUALITY>), grid, block, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int TV>
hipError_t launch_tv(const SplitObjArgs& a, int obj, hipStream_t st) {
    switch (a.rp) {
        case 16: return launch_nk<1, TV>(a, obj, st);
        case 32: return launch_nk<2, TV>(a, obj, st);
        case 64: return launch_nk<4, TV>(a, obj, st);
        case 128: return launch_nk<8, TV>(a, obj, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_split_wh(const float* W, const float* H, int n, int m, int r, int rp, int nrows, int ncols,
                           uint16_t* Ws, uint16_t* Hs, hipStream_t st) {
    if (r < 1 || r > kMaxFusedRank || rp != split_rank(r) || nrows % 128 || nrows < n || ncols % 32 || ncols < m)
        return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(nrows / 64 + (ncols + 63) / 64);
    switch (rp) {
        case 16: hipLaunchKernelGGL(k_split_wh<16>, dim3(grid), dim3(NTHREADS), 0, st, W, H, n, m, r, nrows, ncols, Ws, Hs); break;
        case 32: hipLaunchKernelGGL(k_split_wh<32>, dim3(grid), dim3(NTHREADS), 0, st, W, H, n, m, r, nrows, ncols, Ws, Hs); break;
        case 64: hipLaunchKernelGGL(k_split_wh<64>, dim3(grid), dim3(NTHREADS), 0, st, W, H, n, m, r, nrows, ncols, Ws, Hs); break;
        default: hipLaunchKernelGGL(k_split_wh<128>, dim3(grid), dim3(NTHREADS), 0, st, W, H, n, m, r, nrows, ncols, Ws, Hs); break;
    }
    return hipGetLastError();
}

hipError_t launch_objective_split(const SplitObjArgs& a, int vtype, int obj, hipStream_t st) {
    if (a.rp != split_rank(a.r) || a.rblocks * 128 > a.nrows || a.cblocks * 32 > a.ncols || a.csplit < 1 ||
        a.csplit > a.cblocks || (vtype == 0 && a.vec && (a.n % 4 || ((uintptr_t)a.V & 15))))
        return hipErrorInvalidValue;
    switch (vtype) {
        case 0: return a.vec ? launch_tv<0>(a, obj, st) : launch_tv<3>(a, obj, st);
        case 1: return launch_tv<1>(a, obj, st);
        case 2: return launch_tv<2>(a, obj, st);
        case 3: return launch_tv<4>(a, obj, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace gpumlib_nmf
