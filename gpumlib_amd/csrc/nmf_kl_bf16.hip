// KL-divergence rule with BF16 storage: V held in BF16 in both orientations (the tiled
// store of the Euclidean BF16 path), BF16 shadows of W and H as MFMA operands, the ratio
// V / (W H + eps) in FP32, FP32 accumulation, FP32 updates and clamped sums.
//
// The reference iteration (nmf_multiplicative_divergence.cpp:29-50, kernels
// nmf_kernels.cu:63-341) materialises W H twice per iteration; here each pass forms the
// W H block it needs on the matrix cores from the shadows and feeds the ratio straight into
// the second product, so V is the only n x m operand streamed.  As in the Euclidean BF16
// path, the W half is the H half of the transposed problem, U_W(V, W, H) = U_H(V^T, H^T, W^T)^T,
// so one kernel serves both passes:
//   H pass:  A = W (K = n), C = H,   B = V    ->  P(a, u) = sum_i W(i,a) V(i,u) / ((W H)(i,u) + eps)
//   W pass:  A = H^T (K = m), C = W^T, B = V^T ->  P(a, i) = sum_u H(a,u) V(i,u) / ((W H)(i,u) + eps)
// The W pass takes the new H and the old W, as the reference recomputes W H between the
// halves (nmf_multiplicative_divergence.cpp:44).
#include <hip/hip_runtime.h>

#include "nmf_common.hpp"
#include "nmf_internal.hpp"

#include <type_traits>

namespace gpumlib_nmf {

typedef __bf16 kbf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 kbf16x2 __attribute__((ext_vector_type(2)));
typedef float kf32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int ku32x4 __attribute__((ext_vector_type(4)));

constexpr int KKL = 128;        // K per chunk: half of a 256-deep block of the tiled store
constexpr int LDK2 = KKL + 8;   // LDS pitch (bf16) of the K-contiguous tiles: 272 bytes

// The K-contiguous tiles are read 8 bytes per lane (4 k of rows c = 0..31): with a 68-dword
// pitch, rows c and c + 8 fall on the same LDS banks.  Rows with bit 3 set therefore hold
// each 8-byte unit at k ^ 4 (the halves of every 16-byte unit swapped), so those two rows
// differ by 2 banks; a 16-byte unit keeps its place.
__device__ __forceinline__ int kswz(int row, int k) { return k ^ (((row >> 3) & 1) << 2); }
__device__ __forceinline__ ku32x4 kswz16(int row, ku32x4 v) {
    return ((row >> 3) & 1) ? ku32x4{v[2], v[3], v[0], v[1]} : v;
}

__device__ __forceinline__ f32x16 mfma_kl(ku32x4 a, ku32x4 b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kbf16x8, a), __builtin_bit_cast(kbf16x8, b),
                                                   acc, 0, 0, 0);
}

__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
    const kf32x2 x = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, kbf16x2));  // round to nearest even
}

// One workgroup per (slice s, column tile t of NT * 64 columns), 4 NT waves.  Per 128-deep
// chunk, the V columns, the A rows in both layouts and (once per workgroup) the tile's C
// columns are staged through LDS; the next chunk's loads are in flight in registers while
// the current one is multiplied.  Wave w owns column block ub = w % NCB and k blocks
// w / NCB + 2 j of every chunk; per 32 x 32 block (k block, ub) it forms A C (K = the rank,
// RP / 16 MFMAs), turns it into the ratio in the accumulator layout (lane (c, h) holds rows
// acc_row(v, h) of column c), rounds it to BF16 and uses it directly as the B operand of the
// second product over that block's 32 k (two MFMAs per rank block), whose A operand reads
// the same rows of A (the contraction order is free, so the k a lane supplies follows the
// accumulator layout).  The two waves of a column block add their partial products in wave
// order at the end.  V, the shadows and C are padded with zeros to whole chunks and tiles:
// padded rows contribute V = 0 (ratio 0, eps > 0); padded columns are not stored; rank rows
// >= r of P (A rows clamped to r - 1) are not read by the update.
// SNT: the K slices that start at or after contraction index p.snt_k0 load their V chunks with
// the non-temporal policy (each workgroup runs the loop copy of its slice's policy).
template <int MB, int NT, bool SNT>
__global__ __launch_bounds__(NTHREADS * NT) void k_pass_kl_bf16(KlLpArgs p) {
    constexpr int THR = NTHREADS * NT;
    constexpr int NW = THR / 64;
    constexpr int RP = MB * 32;
    constexpr int TW = NT * TILE_LP;
    constexpr int NCB = TW / 32;           // column blocks of 32
    constexpr int WPC = NW / NCB;          // waves per column block
    constexpr int NKB = KKL / 32;          // k blocks per chunk
    constexpr int KPW = NKB / WPC;         // k blocks per wave per chunk
    constexpr int LDR2 = RP + 8;           // LDS pitch (bf16) of the rank-contiguous tiles
    constexpr int BLB = TILE_LP * KKL / 8 / THR;  // 16-byte B loads per thread per 64-column store block
    constexpr int BL = NT * BLB;
    constexpr int AKL = RP * KKL / 8 / THR;       // 16-byte loads of the K-contiguous A chunk
    constexpr int ARL = KKL * RP / 8 / THR;       // 16-byte loads of the rank-contiguous A chunk
    constexpr int CLD = TW * RP / 8 / THR;        // 16-byte loads of the tile's C
    constexpr int RQ = RP / 8;                    // 16-byte groups per rank-contiguous row
    static_assert(WPC * NCB == NW && KPW * WPC == NKB, "wave split");
    static_assert(BLB >= 1 && AKL >= 1 && ARL >= 1 && CLD >= 1, "load split");
    // the V chunk [col][k], the A chunk [a][k] and [k][a], the tile's C [col][a]; after the
    // last chunk the same LDS holds the waves' partial products
    struct Stage {
        uint16_t sB[TW][LDK2];
        uint16_t sAk[RP][LDK2];
        uint16_t sAr[KKL][LDR2];
        uint16_t sC[TW][LDR2];
    };
    union Lds {
        Stage st;
        float red[NCB * MB * 16 * 64];
    };
    __shared__ __attribute__((aligned(16))) Lds lds;
    auto& sB = lds.st.sB;
    auto& sAk = lds.st.sAk;
    auto& sAr = lds.st.sAr;
    auto& sC = lds.st.sC;

    const int total = p.tiles * p.S;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, h = lane >> 5;
    const int ub = wave % NCB, kb0 = wave / NCB;
    const int col0 = t * TW;
    const int ncols = min(TW, p.N - col0);
    const int kbeg = s * p.kslice;
    const int kend = min(round_up(p.K, KKL), kbeg + p.kslice);
    const int nt64 = (p.N + TILE_LP - 1) / TILE_LP;

    // the tile's C columns (rank-contiguous rows; rows past N are zero in the shadow)
    {
        ku32x4 rc[CLD];
#pragma unroll
        for (int j = 0; j < CLD; j++) {
            const int q = tid + THR * j, col = q / RQ, aq = (q % RQ) * 8;
            rc[j] = *reinterpret_cast<const ku32x4*>(p.Cr + (size_t)(col0 + col) * p.rps + aq);
        }
#pragma unroll
        for (int j = 0; j < CLD; j++) {
            const int q = tid + THR * j, col = q / RQ, aq = (q % RQ) * 8;
            *reinterpret_cast<ku32x4*>(&sC[col][aq]) = rc[j];
        }
    }

    ku32x4 rb[BL], rk[AKL], rr[ARL];
    auto gload = [&](auto nt_tag, int k0) __attribute__((always_inline)) {
        // the chunk is half of a 256-deep block: 256 contiguous bytes of each of its 64 columns
        const uint16_t* crow = p.B + (size_t)(k0 >> 8) * nt64 * (TILE_LP * 256) + (k0 & 255);
        auto src = [&](int j) __attribute__((always_inline)) {
            const int q = tid + THR * j, tb = j / BLB, ql = q - tb * BLB * THR;
            const int b = min(t * NT + tb, nt64 - 1);
            const int col = ql / (KKL / 8), kq = (ql % (KKL / 8)) * 8;
            return reinterpret_cast<const ku32x4*>(crow + (size_t)b * (TILE_LP * 256) + col * 256 + kq);
        };
#pragma unroll
        for (int j = 0; j < BL; j++) rb[j] = decltype(nt_tag)::value ? stream_load(src(j)) : *src(j);
#pragma unroll
        for (int j = 0; j < AKL; j++) {
            const int q = tid + THR * j, a = q / (KKL / 8), kq = (q % (KKL / 8)) * 8;
            const int aa = a < p.r ? a : p.r - 1;
            rk[j] = *reinterpret_cast<const ku32x4*>(p.Ak + (size_t)aa * p.lda + k0 + kq);
        }
#pragma unroll
        for (int j = 0; j < ARL; j++) {
            const int q = tid + THR * j, k = q / RQ, aq = (q % RQ) * 8;
            rr[j] = *reinterpret_cast<const ku32x4*>(p.Ar + (size_t)(k0 + k) * p.rps + aq);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto sstore = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < BL; j++) {
            const int q = tid + THR * j, tb = j / BLB, ql = q - tb * BLB * THR;
            const int col = tb * TILE_LP + ql / (KKL / 8), kq = (ql % (KKL / 8)) * 8;
            *reinterpret_cast<ku32x4*>(&sB[col][kq]) = kswz16(col, rb[j]);
        }
#pragma unroll
        for (int j = 0; j < AKL; j++) {
            const int q = tid + THR * j, a = q / (KKL / 8), kq = (q % (KKL / 8)) * 8;
            *reinterpret_cast<ku32x4*>(&sAk[a][kq]) = kswz16(a, rk[j]);
        }
#pragma unroll
        for (int j = 0; j < ARL; j++) {
            const int q = tid + THR * j, k = q / RQ, aq = (q % RQ) * 8;
            *reinterpret_cast<ku32x4*>(&sAr[k][aq]) = rr[j];
        }
    };

    f32x16 acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
        for (int v = 0; v < 16; v++) acc[mb][v] = 0.f;
    // column sums of A over the slice (tile 0 only): thread -> rank a = tid / 4, a quarter of
    // every chunk's k range, added in k order; the quarters are combined at the end
    float csum = 0.f;

    const float eps = p.eps;
    // one 32 x 32 block (k block kb, column block ub) of the chunk
    auto block = [&](int kb) __attribute__((always_inline)) {
        {
            f32x16 wh;
#pragma unroll
            for (int v = 0; v < 16; v++) wh[v] = 0.f;
#pragma unroll(MB <= 2 ? RP / 16 : 2)
            for (int g = 0; g < RP / 16; g++) {
                const ku32x4 av = *reinterpret_cast<const ku32x4*>(&sAr[kb * 32 + c][g * 16 + h * 8]);
                const ku32x4 bv = *reinterpret_cast<const ku32x4*>(&sC[ub * 32 + c][g * 16 + h * 8]);
                wh = mfma_kl(av, bv, wh);
            }
            // ratio of rows kb*32 + 8q + 4h + (0..3), column ub*32 + c, packed as the B operand of
            // the k steps of 16: step st takes q = 2 st (its first 4 k) and q = 2 st + 1
            ku32x4 rv[2];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint2 vv =
                    *reinterpret_cast<const uint2*>(&sB[ub * 32 + c][kswz(c, kb * 32 + 8 * q + 4 * h)]);
                const float v0 = __uint_as_float(vv.x << 16), v1 = __uint_as_float(vv.x & 0xffff0000u);
                const float v2 = __uint_as_float(vv.y << 16), v3 = __uint_as_float(vv.y & 0xffff0000u);
                const uint32_t p01 = pack_bf16x2(kl_ratio(v0, wh[4 * q], eps), kl_ratio(v1, wh[4 * q + 1], eps));
                const uint32_t p23 = pack_bf16x2(kl_ratio(v2, wh[4 * q + 2], eps), kl_ratio(v3, wh[4 * q + 3], eps));
                rv[q >> 1][(q & 1) * 2] = p01;
                rv[q >> 1][(q & 1) * 2 + 1] = p23;
            }
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
#pragma unroll
                for (int st = 0; st < 2; st++) {
                    if constexpr (MB > 2) __builtin_amdgcn_sched_barrier(0);  // no hoisting across rank blocks
                    const uint2 a0 =
                        *reinterpret_cast<const uint2*>(&sAk[mb * 32 + c][kswz(c, kb * 32 + 16 * st + 4 * h)]);
                    const uint2 a1 =
                        *reinterpret_cast<const uint2*>(&sAk[mb * 32 + c][kswz(c, kb * 32 + 16 * st + 8 + 4 * h)]);
                    const ku32x4 av = {a0.x, a0.y, a1.x, a1.y};
                    acc[mb] = mfma_kl(av, rv[st], acc[mb]);
                }
            }
        }
    };
    auto compute = [&](auto sums_tag) __attribute__((always_inline)) {
        if constexpr (MB <= 2) {
#pragma unroll
            for (int j = 0; j < KPW; j++) block(kb0 + WPC * j);
        } else {
            // four rank blocks: one k block at a time keeps the registers within a wave's share
#pragma unroll 1
            for (int j = 0; j < KPW; j++) block(kb0 + WPC * j);
        }
        if constexpr (decltype(sums_tag)::value) {
            if (tid < 4 * RP) {
                const int a = tid >> 2, part = tid & 3;
#pragma unroll
                for (int e = 0; e < KKL / 4 / 8; e++) {
                    const ku32x4 x = *reinterpret_cast<const ku32x4*>(&sAk[a][part * (KKL / 4) + e * 8]);
#pragma unroll
                    for (int d = 0; d < 4; d++) {
                        csum += __uint_as_float(x[d] << 16);
                        csum += __uint_as_float(x[d] & 0xffff0000u);
                    }
                }
            }
        }
    };

    const int nchunks = (kend - kbeg) / KKL;
    auto mainloop = [&](auto sums_tag, auto nt_tag) __attribute__((always_inline)) {
        gload(nt_tag, kbeg);
        sstore();
        __syncthreads();
        for (int ch = 0; ch < nchunks; ch++) {
            const bool more = ch + 1 < nchunks;
            if (more) gload(nt_tag, kbeg + (ch + 1) * KKL);
            compute(sums_tag);
            __syncthreads();
            if (more) {
                sstore();
                __syncthreads();
            }
        }
    };
    auto run = [&](auto nt_tag) __attribute__((always_inline)) {
        if (t == 0)
            mainloop(std::integral_constant<bool, true>{}, nt_tag);
        else
            mainloop(std::integral_constant<bool, false>{}, nt_tag);
    };
    if (SNT && kbeg >= p.snt_k0)
        run(std::bool_constant<SNT>{});
    else
        run(std::false_type{});

    if (t == 0 && tid < 4 * RP) {
        // quarters of rank a sit in lanes 4a .. 4a + 3 of one wave: (q0 + q1) + (q2 + q3)
        csum += __shfl_xor(csum, 1, 64);
        csum += __shfl_xor(csum, 2, 64);
        const int a = tid >> 2;
        if ((tid & 3) == 0 && a < p.r) p.Psum[(size_t)s * p.rpad + a] = csum;
    }

    // the WPC waves of a column block add their products in wave order through the free LDS
    static_assert(WPC == 2, "two waves per column block");
    float* red = lds.red;
    if (kb0 == 1) {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int v = 0; v < 16; v++) red[((ub * MB + mb) * 16 + v) * 64 + lane] = acc[mb][v];
    }
    __syncthreads();
    if (kb0 != 0) return;
    float* slice = p.P + (size_t)s * p.pstride;
    const int ul = ub * 32 + c;
    if (ul >= ncols) return;
    float* dst = slice + (size_t)(col0 + ul) * p.ldr;
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int a0 = mb * 32 + g * 8 + h * 4;
            float o[4];
#pragma unroll
            for (int v = 0; v < 4; v++) o[v] = acc[mb][g * 4 + v] + red[((ub * MB + mb) * 16 + g * 4 + v) * 64 + lane];
            if (a0 < p.ldr) *reinterpret_cast<float4*>(dst + a0) = make_float4(o[0], o[1], o[2], o[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// Rank-contiguous BF16 shadow of a factor: Xr[j*rps + a] = bf16(X(j, a)), X(j, a) at
// X[j*xs_j + a*xs_a] (W: j = i, xs_j = 1, xs_a = n; H: j = u, xs_j = r, xs_a = 1); zero for
// a in [r, rps) and for j in [count, rows).
__global__ __launch_bounds__(NTHREADS) void k_shadow_rank_bf16(const float* __restrict__ X, long xs_j, long xs_a,
                                                               int count, int r, int rows, uint16_t* __restrict__ Xr,
                                                               int rps) {
    const long total = (long)rows * rps;
    for (long e = (long)blockIdx.x * NTHREADS + threadIdx.x; e < total; e += (long)gridDim.x * NTHREADS) {
        const long j = e / rps, a = e - j * rps;
        Xr[e] = (j < count && a < r) ? __builtin_bit_cast(uint16_t, (__bf16)X[j * xs_j + a * xs_a]) : (uint16_t)0;
    }
}

// KL update in row form, for partials P[s][j*ldr + a] as the pass above writes them:
//   X(j, a) *= (sum_s P[s][j*ldr + a]) / max(sum_s Psum[s][a], eps)
// (nmf_kernels.cu:289-291, 337-339: the multiplier over the clamped sum, :104,190), both sums
// in slice order, then both BF16 shadows of the new X: Xk[a*ldk + j] (K-contiguous, the A
// operand of the other pass) and Xr[j*rps + a] (rank-contiguous).  JT rows per workgroup; X
// staged through LDS in whichever order its layout reads coalesced; each thread owns (row jj,
// 4 consecutive a) items and issues all its slice loads (eight slices at a time) before adding.
template <int JT, int RMAX>
__global__ __launch_bounds__(NTHREADS) void k_update_kl_rows_bf16(float* __restrict__ X, long xs_j, long xs_a,
                                                                  const float* __restrict__ P,
                                                                  const float* __restrict__ Psum, int count, int r,
                                                                  int ldr, int rpad, int S, long pstride, float eps,
                                                                  uint16_t* __restrict__ Xk, int ldk,
                                                                  uint16_t* __restrict__ Xr, int rps) {
    constexpr int XS = RMAX + 1;                   // LDS row pitch (r <= RMAX)
    constexpr int MAXI = (JT * RMAX / 4 + NTHREADS - 1) / NTHREADS;
    constexpr int XU = (JT * RMAX + NTHREADS - 1) / NTHREADS;
    __shared__ float sX[JT * XS];
    __shared__ float ssum[RMAX];
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * JT, nj = min(JT, count - j0);
    const int rq = ldr >> 2;
    const int items = JT * rq;
    // element idx of the workgroup's X rows -> (row jj, rank a), in the order its layout reads coalesced
    auto xpos = [&](int idx, int& jj, int& a) {
        if (xs_a == 1) jj = idx / r, a = idx - jj * r;
        else a = idx / JT, jj = idx - a * JT;
    };

    // 1. every load: the X rows, the slice partials of this thread's items, the slice sums
    float xr[XU];
#pragma unroll
    for (int u = 0; u < XU; u++) {
        const int idx = tid + u * NTHREADS;
        int jj, a;
        xpos(idx, jj, a);
        xr[u] = (idx < JT * r && jj < nj) ? X[(size_t)(j0 + jj) * xs_j + (size_t)a * xs_a] : 0.f;
    }
    float4 num[MAXI];
    const long ps4 = pstride >> 2;
    for (int s0 = 0; s0 < S; s0 += 8) {
        float4 x[MAXI][8];
#pragma unroll
        for (int k = 0; k < MAXI; k++) {
            const int item = tid + k * NTHREADS, jj = item / rq, q = item - jj * rq;
            const bool ok = item < items && jj < nj;
            const float4* src = reinterpret_cast<const float4*>(P + (size_t)(j0 + jj) * ldr) + q;
#pragma unroll
            for (int b = 0; b < 8; b++)
                x[k][b] = (ok && s0 + b < S) ? src[(size_t)(s0 + b) * ps4] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < MAXI; k++) {
            if (s0 == 0) num[k] = x[k][0];
            else num[k].x += x[k][0].x, num[k].y += x[k][0].y, num[k].z += x[k][0].z, num[k].w += x[k][0].w;
#pragma unroll
            for (int b = 1; b < 8; b++)
                if (s0 + b < S)
                    num[k].x += x[k][b].x, num[k].y += x[k][b].y, num[k].z += x[k][b].z, num[k].w += x[k][b].w;
        }
    }
    if (tid < r) {
        float sm = 0.f;
        for (int s0 = 0; s0 < S; s0 += 8) {
            float y[8];
#pragma unroll
            for (int b = 0; b < 8; b++) y[b] = s0 + b < S ? Psum[(size_t)(s0 + b) * rpad + tid] : 0.f;
#pragma unroll
            for (int b = 0; b < 8; b++)
                if (s0 + b < S) sm += y[b];
        }
        ssum[tid] = sm < eps ? eps : sm;
    }
#pragma unroll
    for (int u = 0; u < XU; u++) {
        const int idx = tid + u * NTHREADS;
        int jj, a;
        xpos(idx, jj, a);
        if (idx < JT * r && jj < nj) sX[jj * XS + a] = xr[u];
    }
    __syncthreads();
    // 2. the new values
#pragma unroll
    for (int k = 0; k < MAXI; k++) {
        const int item = tid + k * NTHREADS, jj = item / rq, q = item - jj * rq;
        if (item >= items || jj >= nj) continue;
        const float nv[4] = {num[k].x, num[k].y, num[k].z, num[k].w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int a = 4 * q + e;
            if (a < r) sX[jj * XS + a] *= (nv[e] / ssum[a]);
        }
    }
    __syncthreads();
    // 3. X and its two shadows, each written in its coalesced order
    for (int idx = tid; idx < JT * r; idx += NTHREADS) {
        int jj, a;
        xpos(idx, jj, a);
        if (jj < nj) X[(size_t)(j0 + jj) * xs_j + (size_t)a * xs_a] = sX[jj * XS + a];
    }
    for (int idx = tid; idx < JT * r; idx += NTHREADS) {
        const int a = idx / JT, jj = idx - a * JT;
        if (jj < nj) Xk[(size_t)a * ldk + j0 + jj] = __builtin_bit_cast(uint16_t, (__bf16)sX[jj * XS + a]);
    }
    for (int idx = tid; idx < JT * r; idx += NTHREADS) {
        const int jj = idx / r, a = idx - jj * r;
        if (jj < nj) Xr[(size_t)(j0 + jj) * rps + a] = __builtin_bit_cast(uint16_t, (__bf16)sX[jj * XS + a]);
    }
}

// ---------------------------------------------------------------------------
template <int MB, int NT>
static void launch_kl(const KlLpArgs& a, hipStream_t st) {
    if (a.snt_k0 < a.K)
        hipLaunchKernelGGL((k_pass_kl_bf16<MB, NT, true>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS * NT), 0, st, a);
    else
        hipLaunchKernelGGL((k_pass_kl_bf16<MB, NT, false>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS * NT), 0, st,
                           a);
}

hipError_t launch_pass_kl_bf16(const KlLpArgs& a, hipStream_t st) {
    if (a.r < 1 || a.r > 128 || a.nt != 2 || a.kslice % KKL || a.rps != kl_rank_pitch(a.r) || !(a.eps > 0.f))
        return hipErrorInvalidValue;
    if (a.rps == 32)
        launch_kl<1, 2>(a, st);
    else if (a.rps == 64)
        launch_kl<2, 2>(a, st);
    else
        launch_kl<4, 2>(a, st);
    return hipGetLastError();
}

hipError_t launch_shadow_rank_bf16(const float* X, long xs_j, long xs_a, int count, int r, int rows, uint16_t* Xr,
                                   int rps, hipStream_t st) {
    if (rows < count || rps < r) return hipErrorInvalidValue;
    const long total = (long)rows * rps;
    long blocks = (total + NTHREADS - 1) / NTHREADS;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_shadow_rank_bf16, dim3((unsigned)blocks), dim3(NTHREADS), 0, st, X, xs_j, xs_a, count, r, rows,
                       Xr, rps);
    return hipGetLastError();
}

hipError_t launch_update_kl_rows_bf16(float* X, long xs_j, long xs_a, const float* P, const float* Psum, int count,
                                      int r, int ldr, int rpad, int S, long pstride, float eps, uint16_t* Xk, int ldk,
                                      uint16_t* Xr, int rps, hipStream_t st) {
    if (r < 1 || r > 128 || ldr % 4 || ldr < r || rpad < r || pstride % 4 || S < 1) return hipErrorInvalidValue;
    // r <= 64: 32 rows per workgroup for long factors (at least 256 workgroups), else 16; above, 16
#define KLU(JT, RMAX)                                                                                             \
    hipLaunchKernelGGL((k_update_kl_rows_bf16<JT, RMAX>), dim3((count + JT - 1) / JT), dim3(NTHREADS), 0, st, X, \
                       xs_j, xs_a, P, Psum, count, r, ldr, rpad, S, pstride, eps, Xk, ldk, Xr, rps)
    if (r > 64)
        KLU(16, 128);
    else if (count >= 8192)
        KLU(32, 64);
    else
        KLU(16, 64);
#undef KLU
    return hipGetLastError();
}

}  // namespace gpumlib_nmf
