// FP32 MFMA pass kernels of the multiplicative-update iteration.
//
// H pass (one workgroup = 128 columns u of V x one K slice of rows i):
//   Euclidean:  P[s](a,u) = sum_i W(i,a) V(i,u)              (W^T V, reference
//               nmf_multiplicative_euclidian.cpp:36 via DeviceMatrix.h:217-227)
//               plus one extra tile with V replaced by W      (W^T W, :39)
//   KL:         P[s](a,u) = sum_i W(i,a) V(i,u) / ((W H)(i,u) + eps)
//               (the numerator of nmf_kernels.cu:295-341, with W H formed per
//               tile on MFMA instead of being materialised by a separate SGEMM
//               as in nmf_multiplicative_divergence.cpp:37)
// W pass (one workgroup = 128 rows i of V x one K slice of columns u):
//   Euclidean:  P[s](i,a) = sum_u V(i,u) H(a,u)               (V H^T, :56)
//               plus one extra tile with V replaced by H      (H H^T, :59-61)
//   KL:         P[s](i,a) = sum_u V(i,u) H(a,u) / ((W H)(i,u) + eps)
//               (numerator of nmf_kernels.cu:247-293)
//
// Every workgroup has 4 waves; a wave owns 32 columns (H pass) or 32 rows
// (W pass) of the tile and all r (padded to MB*32) rank rows, i.e. MB 32x32
// MFMA accumulators.  The grouped launches also come with 64-column (64-row)
// tiles (TW = 64): there two waves share each 32-column block, each taking
// half of the k steps of every chunk, and the pair's partials are added in
// LDS after the loop (more, narrower work items for problems with few columns
// or rows).  K is staged through double-buffered LDS in steps of KC
// with one barrier per step; the next one or two steps' global loads are in
// flight (in registers) while the current step's MFMAs run.
#include "nmf_common.hpp"
#include "nmf_internal.hpp"

#include <type_traits>

namespace gpumlib_nmf {

// ---------------------------------------------------------------------------
// Workgroup-wide deterministic sum of one double per thread; thread 0 returns it.
__device__ double block_sum_f64(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NTHREADS / 64; i++) t += red[i];
    }
    return t;
}

// 64-wide tiles: the wave of k half 1 hands its partial blocks to the wave of k half 0 that owns
// the same 32-column block (cw), which adds them (half 0 + half 1, a fixed order).  Every thread
// of the workgroup calls it; red holds 2 * MB * 16 * 64 floats and is free again on return.
template <int MB>
__device__ __forceinline__ void pair_reduce(f32x16 (&acc)[MB], float* red, int cw, int kh, int lane) {
    if (kh == 1) {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int v = 0; v < 16; v++) red[((cw * MB + mb) * 16 + v) * 64 + lane] = acc[mb][v];
    }
    __syncthreads();
    if (kh == 0) {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[mb][v] += red[((cw * MB + mb) * 16 + v) * 64 + lane];
    }
    __syncthreads();
}

// Rank b that lane half h feeds at step kk of a W H product on the 32x32x2 MFMAs, in the passes
// that hold a rank-contiguous copy of the A operand (r <= 64): blocks of 32 ranks, each lane half
// taking 16 consecutive ones, so that a lane's operands of one block are 16 contiguous floats (four
// 16-byte LDS reads instead of sixteen 4-byte ones).  Block by block, so a problem with r <= 32
// sums in the same order on the kernels built for one or for two rank blocks.
__device__ __forceinline__ int wh_rank(int kk, int h) { return (kk >> 4) * 32 + h * 16 + (kk & 15); }

// ---------------------------------------------------------------------------
// H pass.  KL: ratio B operand; OBJ: fused objective of the state (W, H) on
// entry.
// Euclidean gram W^T W: its nb x nb blocks of 32 x 32 are spread over the first
// tiles of every slice (block t on tile t; tiles past the last column tile of V
// carry a gram block only), and inside a tile the 4 waves each take a quarter of
// every chunk's k range, so no workgroup carries more than an eighth extra MFMA
// work and the grid stays one shape.
// FUSE (grouped full-K launches, one slice): the item's partial is the whole product, so the
// workgroup applies the H update to its columns itself instead of writing partials.
// Euclidean: H(a,u) *= WtV(a,u) / (sum_b G(a,b) H(b,u) + eps) with G = W^T W from p.G (computed by
// the gram launch before this one, G(a,b) at p.G[b*rpad + a]); KL: H(a,u) *= num(a,u) /
// max(colsum_a, eps), the column sums of W taken over the whole of K in this item.
template <int MB, bool KL, int OBJ, bool FUSE = false, int TW = TILE>
__device__ __forceinline__ void pass_h_body(const PassArgs& p, const int w, const int nb) {
    constexpr int RP = MB * 32;
    constexpr bool NEED_WH = KL || (OBJ != OBJ_NONE);
    static_assert(TW == 64 || TW == 128, "tile of 64 or 128 columns");
    constexpr int NCB = TW / 32;                        // 32-column blocks of the tile
    constexpr int KH = 4 / NCB;                         // waves per block, each on its k groups
    constexpr int VL = TW * KC / 4 / NTHREADS;          // float4 of the V chunk per thread
    // TRK: the W H product reads W from a rank-contiguous copy of the chunk (wh_rank order)
    constexpr bool TRK = NEED_WH && MB <= 2;
    // SWA: the plain Euclidean pass keeps the W chunk unpadded, its 16-byte groups XOR-swizzled by
    // row (as free of bank conflicts as the padded rows, 3 KB less LDS per workgroup)
    constexpr bool SWA = !NEED_WH && !FUSE;
    constexpr int LDA = SWA ? KC : LDK;
    __shared__ __attribute__((aligned(16))) float sA[2][RP][LDA];  // W chunk, [a][i]
    __shared__ __attribute__((aligned(16))) float sB[2][TW][LDK];  // V (or W) chunk, [u][i]
    __shared__ __attribute__((aligned(16))) float sAt[TRK ? 2 : 1][TRK ? KC : 1][RP + 4];  // W chunk, [i][a]
    __shared__ double red[NTHREADS / 64];
    static_assert(KH == 1 || 2 * TW * LDK >= 2 * MB * 16 * 64, "pair reduction fits the V chunk buffers");

    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: branches on it stay scalar
    const int cw = wave % NCB, kh = wave / NCB;  // the wave's column block and k half (TW = 64)
    const bool owner = KH == 1 || kh == 0;       // holds the block's total after the pair reduction
    const int c = lane & 31, h = lane >> 5;
    const int vt = (p.m + TW - 1) / TW;
    const int col0 = t * TW;
    const int ncols = (t < vt) ? min(TW, p.m - col0) : 0;
    const int nblk = nb * nb;
    const bool gcar = (!FUSE) && (!KL) && p.G != nullptr && t < nblk;  // the plan gives tiles >= nblk
    const int kbeg = s * p.kslice;
    const int kend = min(p.n, kbeg + p.kslice);
    const bool vec = p.vec != 0;
    auto ac = [](int row, int col) __attribute__((always_inline)) { return SWA ? (col ^ ((row & 7) << 2)) : col; };

    // Per-wave B operand of the W H product: H(b, u) for b = wh_rank(kk, h) (TRK) or 2*kk + h,
    // u = this lane's column.
    float hreg[NEED_WH ? MB * 16 : 1];
    if (NEED_WH) {
        const int u = col0 + cw * 32 + c;
#pragma unroll
        for (int kk = 0; kk < MB * 16; kk++) {
            const int b = TRK ? wh_rank(kk, h) : 2 * kk + h;
            hreg[kk] = (u < p.m && b < p.r) ? p.H[(size_t)u * p.r + b] : 0.f;
        }
    }

    // Register staging of the next chunks' global loads.  The plain Euclidean pass keeps
    // two chunks in flight (two register sets; the LDS is double-buffered); the KL and
    // objective variants, whose registers also hold H and the ratio, keep one.
    constexpr bool DEEP = !KL && OBJ == OBJ_NONE;
    float4 ra0[MB], rb0[VL], ra1[DEEP ? MB : 1], rb1[DEEP ? VL : 1];
    auto gload = [&](float4 (&rb)[VL], float4 (&ra)[MB], int k0) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < VL; j++) {
            const int q = tid + NTHREADS * j, col = q >> 3, kq = (q & 7) * 4;
            rb[j] = load4(p.V + (size_t)(col0 + col) * p.ldv + k0 + kq, col < ncols, k0 + kq, kend, vec);
        }
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const int q = tid + NTHREADS * j, a = q >> 3, kq = (q & 7) * 4;
            ra[j] = load4(p.W + (size_t)a * p.n + k0 + kq, a < p.r, k0 + kq, kend, vec);
        }
    };
    auto sstore = [&](const float4 (&rb)[VL], const float4 (&ra)[MB], int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < VL; j++) {
            const int q = tid + NTHREADS * j, col = q >> 3, kq = (q & 7) * 4;
            *reinterpret_cast<float4*>(&sB[buf][col][kq]) = rb[j];
        }
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const int q = tid + NTHREADS * j, a = q >> 3, kq = (q & 7) * 4;
            *reinterpret_cast<float4*>(&sA[buf][a][ac(a, kq)]) = ra[j];
            if constexpr (TRK) {
                sAt[buf][kq][a] = ra[j].x;
                sAt[buf][kq + 1][a] = ra[j].y;
                sAt[buf][kq + 2][a] = ra[j].z;
                sAt[buf][kq + 3][a] = ra[j].w;
            }
        }
    };

    f32x16 acc[MB], gacc;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
        for (int v = 0; v < 16; v++) acc[mb][v] = 0.f;
    for (int v = 0; v < 16; v++) gacc[v] = 0.f;
    const int gmb = gcar ? t / nb : 0, gnb = gcar ? t - (t / nb) * nb : 0;  // carried gram block
    double objacc = 0.0;
    float colsum = 0.f;  // KL: sum over this slice of W(i, tid) (tile 0 only)
    const bool do_sums = KL && (FUSE || t == 0) && (tid < RP);

    // The gram-carrying and the plain workgroups run separate copies of the loop, so
    // that no MFMA accumulator sits under a data-dependent branch.
    const int nchunks = (kend - kbeg + KC - 1) / KC;
    // One chunk out of LDS buffer ch & 1.
    auto chunk = [&](auto gram_tag, int ch) __attribute__((always_inline)) {
        constexpr bool GR = decltype(gram_tag)::value;
        const int buf = ch & 1;

        if (do_sums) {
#pragma unroll
            for (int j = 0; j < KC; j += 4) {
                const float4 x = *reinterpret_cast<const float4*>(&sA[buf][tid][j]);
                colsum += (x.x + x.y) + (x.z + x.w);
            }
        }

        if (NEED_WH) {
            // wh(i, u) for the 32 rows of this step and the wave's 32 columns.
            f32x16 wh;
            for (int v = 0; v < 16; v++) wh[v] = 0.f;
            if constexpr (TRK) {
#pragma unroll
                for (int blk = 0; blk < MB; blk++) {
                    if (blk >= nb) continue;
#pragma unroll
                    for (int q4 = 0; q4 < 4; q4++) {
                        const float4 w4 = *reinterpret_cast<const float4*>(&sAt[buf][c][blk * 32 + h * 16 + q4 * 4]);
                        const int kk = blk * 16 + q4 * 4;
                        wh = mfma32x2(w4.x, hreg[kk], wh);
                        wh = mfma32x2(w4.y, hreg[kk + 1], wh);
                        wh = mfma32x2(w4.z, hreg[kk + 2], wh);
                        wh = mfma32x2(w4.w, hreg[kk + 3], wh);
                    }
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < MB * 16; kk++)
                    if (kk < nb * 16) wh = mfma32x2(sA[buf][2 * kk + h][c], hreg[kk], wh);
            }
            // wh[v] is row i = acc_row(v, h) of the chunk, column c of the wave's block; with
            // 64-column tiles the wave takes the k groups g of its half only.
#pragma unroll
            for (int g = 0; g < 4; g++) {
                if (KH == 2 && (g >> 1) != kh) continue;
                const float4 vv = *reinterpret_cast<const float4*>(&sB[buf][cw * 32 + c][g * 8 + h * 4]);
                const float vj[4] = {vv.x, vv.y, vv.z, vv.w};
                float rj[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (OBJ != OBJ_NONE) objacc += (double)obj_term_f32<OBJ>(vj[j], wh[g * 4 + j], p.eps);
                    if (KL) rj[j] = kl_ratio(vj[j], wh[g * 4 + j], p.eps);
                }
                if (KL) {
#pragma unroll
                    for (int mb = 0; mb < MB; mb++) {
                        if (mb >= nb) continue;
                        const float4 wa = *reinterpret_cast<const float4*>(&sA[buf][mb * 32 + c][g * 8 + h * 4]);
                        acc[mb] = mfma32x2(wa.x, rj[0], acc[mb]);
                        acc[mb] = mfma32x2(wa.y, rj[1], acc[mb]);
                        acc[mb] = mfma32x2(wa.z, rj[2], acc[mb]);
                        acc[mb] = mfma32x2(wa.w, rj[3], acc[mb]);
                    }
                }
            }
        }
        if (!KL) {
#pragma unroll
            for (int gg = 0; gg < 4 / KH; gg++) {
                const int g = kh * (4 / KH) + gg;
                const float4 bv = *reinterpret_cast<const float4*>(&sB[buf][cw * 32 + c][g * 8 + h * 4]);
#pragma unroll
                for (int mb = 0; mb < MB; mb++) {
                    if (mb >= nb) continue;
                    const float4 av = *reinterpret_cast<const float4*>(&sA[buf][mb * 32 + c][ac(c, g * 8 + h * 4)]);
                    acc[mb] = mfma32x2(av.x, bv.x, acc[mb]);
                    acc[mb] = mfma32x2(av.y, bv.y, acc[mb]);
                    acc[mb] = mfma32x2(av.z, bv.z, acc[mb]);
                    acc[mb] = mfma32x2(av.w, bv.w, acc[mb]);
                }
            }
        }
        if constexpr (GR) {  // this wave's k group (wave*8 .. wave*8+7) of the chunk for the carried block
            const float4 av = *reinterpret_cast<const float4*>(&sA[buf][gmb * 32 + c][ac(c, wave * 8 + h * 4)]);
            const float4 bv = *reinterpret_cast<const float4*>(&sA[buf][gnb * 32 + c][ac(c, wave * 8 + h * 4)]);
            gacc = mfma32x2(av.x, bv.x, gacc);
            gacc = mfma32x2(av.y, bv.y, gacc);
            gacc = mfma32x2(av.z, bv.z, gacc);
            gacc = mfma32x2(av.w, bv.w, gacc);
        }
    };
    auto mainloop = [&](auto gram_tag) __attribute__((always_inline)) {
        if constexpr (DEEP) {
            // chunk ch is in LDS, the register set passed holds chunk ch + 1 (in flight), the
            // other set chunk ch + 2; after the multiply chunk ch + 1 goes to the other LDS
            // buffer and its register set takes the loads of chunk ch + 3
            auto step = [&](int ch, float4 (&rbn)[VL], float4 (&ran)[MB]) __attribute__((always_inline)) {
                chunk(gram_tag, ch);
                if (ch + 1 < nchunks) {
                    sstore(rbn, ran, (ch + 1) & 1);
                    if (ch + 3 < nchunks) gload(rbn, ran, kbeg + (ch + 3) * KC);
                }
                __syncthreads();
            };
            gload(rb0, ra0, kbeg);
            if (nchunks > 1) gload(rb1, ra1, kbeg + KC);
            sstore(rb0, ra0, 0);
            __syncthreads();
            if (nchunks > 2) gload(rb0, ra0, kbeg + 2 * KC);
            for (int ch = 0; ch < nchunks; ch += 2) {
                step(ch, rb1, ra1);
                if (ch + 1 < nchunks) step(ch + 1, rb0, ra0);
            }
        } else {
            gload(rb0, ra0, kbeg);
            sstore(rb0, ra0, 0);
            __syncthreads();
            for (int ch = 0; ch < nchunks; ch++) {
                if (ch + 1 < nchunks) gload(rb0, ra0, kbeg + (ch + 1) * KC);
                chunk(gram_tag, ch);
                if (ch + 1 < nchunks) sstore(rb0, ra0, (ch + 1) & 1);
                __syncthreads();
            }
        }
    };
    if (gcar)
        mainloop(std::true_type{});
    else
        mainloop(std::false_type{});
    if constexpr (KH == 2) pair_reduce<MB>(acc, &sB[0][0][0], cw, kh, lane);

    if constexpr (FUSE) {
        // acc[mb][v] = WtV or the KL numerator at (a = mb*32 + acc_row(v,h), u = col0 + cw*32 + c)
        constexpr int GS = RP + 1;
        float* sG = &sA[0][0][0];  // LDS is free after the loop's last barrier
        const int ul = cw * 32 + c, u = col0 + ul;
        const bool ucol = ul < ncols;
        if (KL) {
            if (do_sums) sG[tid] = (colsum < p.eps) ? p.eps : colsum;
        } else {
            const int rpad = round_up(p.r, 4);
            for (int e = tid; e < RP * RP; e += NTHREADS) {
                const int a = e / RP, b = e - a * RP;
                float g = 0.f;
                if (a < p.r && b < p.r)
                    for (int gs = 0; gs < p.gsl; gs++) g += p.G[(size_t)gs * p.r * rpad + (size_t)b * rpad + a];
                sG[a * GS + b] = g;
            }
        }
        __syncthreads();
        float hb[KL ? 1 : MB * 16];  // B operand of G H: H(b = 2kk + h, u)
        if constexpr (!KL) {
#pragma unroll
            for (int kk = 0; kk < MB * 16; kk++) {
                const int b = 2 * kk + h;
                if constexpr (NEED_WH && !TRK)
                    hb[kk] = hreg[kk];
                else
                    hb[kk] = (ucol && b < p.r) ? p.H[(size_t)u * p.r + b] : 0.f;
            }
        }
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
            if (mb >= nb || !owner) continue;
            f32x16 d;
            for (int v = 0; v < 16; v++) d[v] = 0.f;
            if constexpr (!KL) {
#pragma unroll
                for (int kk = 0; kk < MB * 16; kk++)
                    if (kk < nb * 16) d = mfma32x2(sG[(mb * 32 + c) * GS + 2 * kk + h], hb[kk], d);
            }
            if (!ucol) continue;
            float* hcol = const_cast<float*>(p.H) + (size_t)u * p.r;  // the problem's own H (JobArgs::H)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int a0 = mb * 32 + g * 8 + h * 4;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int a = a0 + j;
                    if (a >= p.r) continue;
                    const float num = acc[mb][g * 4 + j];
                    if (KL)
                        hcol[a] *= (num / sG[a]);
                    else
                        hcol[a] *= (num / (d[g * 4 + j] + p.eps));
                }
            }
        }
    } else {
        // acc[mb][v] = P(a = mb*32 + acc_row(v,h), u = cw*32 + c)
        const int ul = cw * 32 + c;
        if (owner && ul < ncols) {
            float* dst = p.P + (size_t)s * p.pstride + (size_t)(col0 + ul) * p.ldr;
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
        if (do_sums && tid < p.r) p.Psum[(size_t)s * round_up(p.r, 4) + tid] = colsum;
        if (gcar) {
            // gram(a, b) at column u = m + b of the slice (ldr == rpad); the 4 waves' k quarters
            // are summed in wave order through LDS (free after the loop)
            float* gred = &sB[0][0][0];
            float* gdst = p.P + (size_t)s * p.pstride + (size_t)p.m * p.ldr;
#pragma unroll
            for (int v = 0; v < 16; v++) gred[(wave * 64 + lane) * 16 + v] = gacc[v];
            __syncthreads();
            for (int e = tid; e < 64 * 16; e += NTHREADS) {
                const int le = e >> 4, v = e & 15;
                const float g = ((gred[le * 16 + v] + gred[(64 + le) * 16 + v]) + gred[(128 + le) * 16 + v]) +
                                gred[(192 + le) * 16 + v];
                const int a = gmb * 32 + acc_row(v, le >> 5), b = gnb * 32 + (le & 31);
                if (a < p.ldr && b < p.r) gdst[(size_t)b * p.ldr + a] = g;
            }
            gram_finish4(gdst - (size_t)s * p.pstride, p.pstride, p.S, p.S * nblk, p.r * p.ldr, p.G, p.gcount);
        }
    }
    if (OBJ != OBJ_NONE) {
        const double tot = block_sum_f64(objacc, red);
        if (tid == 0) p.objp[w] = tot;
    }
}

// ---------------------------------------------------------------------------
// W pass, Euclidean rule: P[s](i,a) = sum_u V(i,u) H(a,u), and the gram H H^T spread
// over the first tiles of every slice as in the H pass (k quarters per wave).
// FUSE: as in the H pass, the workgroup applies W(i,a) *= VHt(i,a) / (sum_b W(i,b) G(b,a) + eps)
// to its rows, G = H H^T from p.G (the gram launch before this one).
template <int MB, bool FUSE = false, int TW = TILE>
__device__ __forceinline__ void pass_w_euc_body(const PassArgs& p, const int w, const int nb) {
    constexpr int RP = MB * 32;
    static_assert(TW == 64 || TW == 128, "tile of 64 or 128 rows");
    constexpr int NCB = TW / 32;    // 32-row blocks of the tile
    constexpr int KH = 4 / NCB;     // waves per block, each on its k steps
    constexpr int VQ = TW / 4;      // float4 per chunk row u
    constexpr int VL = KC * VQ / NTHREADS;
    // rows k and k+1 land 32 banks apart: by padding (fused pass, whose gram staging reuses the
    // buffers) or, unpadded, by swapping the two 32-float halves of every odd row (16 KB less LDS
    // per workgroup at r = 64)
    constexpr bool SW = !FUSE;
    constexpr int PA = SW ? TW : TW + 32;
    constexpr int PB = SW ? RP : RP + 32;
    constexpr int XA = SW ? 32 : 0;
    constexpr int XB = (SW && RP % 64 == 0) ? 32 : 0;
    __shared__ __attribute__((aligned(16))) float sA[2][KC][PA];  // [u][i]
    __shared__ __attribute__((aligned(16))) float sB[2][KC][PB];  // [u][a]
    static_assert(KH == 1 || 2 * KC * PB >= 2 * MB * 16 * 64, "pair reduction fits the H chunk buffers");

    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: branches on it stay scalar
    const int cw = wave % NCB, kh = wave / NCB;  // the wave's row block and k half (TW = 64)
    const bool owner = KH == 1 || kh == 0;
    const int c = lane & 31, h = lane >> 5;
    const int vt = (p.n + TW - 1) / TW;
    const float* A = p.V;
    const int lda = p.n;
    const int row0 = t * TW;
    const int nrows = (t < vt) ? min(TW, p.n - row0) : 0;
    const int nblk = nb * nb;
    const bool gcar = !FUSE && p.G != nullptr && t < nblk;  // the plan gives tiles >= nblk
    const int kbeg = s * p.kslice;
    const int kend = min(p.m, kbeg + p.kslice);
    const bool vec = p.vec != 0;
    constexpr int BQ = RP / 4;                // float4 per H column chunk
    constexpr int BLOADS = (KC * BQ + NTHREADS - 1) / NTHREADS;

    // two chunks of global loads in flight (two register sets, double-buffered LDS) up to
    // r = 96; one at r = 128, where a second set would not fit two workgroups per CU
    constexpr bool DEEP = MB <= 3;
    float4 ra0[VL], rb0[BLOADS], ra1[DEEP ? VL : 1], rb1[DEEP ? BLOADS : 1];
    auto gload = [&](float4 (&ra)[VL], float4 (&rb)[BLOADS], int k0) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < VL; j++) {
            const int q = tid + NTHREADS * j, uu = q / VQ, iq = (q % VQ) * 4;
            ra[j] = load4(A + (size_t)(k0 + uu) * lda + row0 + iq, k0 + uu < kend, iq, nrows, vec);
        }
#pragma unroll
        for (int j = 0; j < BLOADS; j++) {
            const int q = tid + NTHREADS * j, uu = q / BQ, aq = (q % BQ) * 4;
            rb[j] = load4(p.H + (size_t)(k0 + uu) * p.r + aq, (uu < KC) && (k0 + uu < kend), aq, p.r, vec);
        }
    };
    auto sstore = [&](const float4 (&ra)[VL], const float4 (&rb)[BLOADS], int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < VL; j++) {
            const int q = tid + NTHREADS * j, uu = q / VQ, iq = (q % VQ) * 4;
            *reinterpret_cast<float4*>(&sA[buf][uu][iq ^ ((uu & 1) ? XA : 0)]) = ra[j];
        }
#pragma unroll
        for (int j = 0; j < BLOADS; j++) {
            const int q = tid + NTHREADS * j, uu = q / BQ, aq = (q % BQ) * 4;
            if (uu < KC) *reinterpret_cast<float4*>(&sB[buf][uu][aq ^ ((uu & 1) ? XB : 0)]) = rb[j];
        }
    };

    f32x16 acc[MB], gacc;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
        for (int v = 0; v < 16; v++) acc[mb][v] = 0.f;
    for (int v = 0; v < 16; v++) gacc[v] = 0.f;
    const int gmb = gcar ? t / nb : 0, gnb = gcar ? t - (t / nb) * nb : 0;  // carried gram block

    // separate loop copies for gram-carrying and plain workgroups (no MFMA under a branch)
    const int nchunks = (kend - kbeg + KC - 1) / KC;
    auto chunk = [&](auto gram_tag, int ch) __attribute__((always_inline)) {
        constexpr bool GR = decltype(gram_tag)::value;
        const int buf = ch & 1;
        {
#pragma unroll
            for (int k2 = 0; k2 < KC / 2 / KH; k2++) {
                const int kk = kh * (KC / 2 / KH) + k2;  // the wave's k steps of the chunk
                const float av = sA[buf][2 * kk + h][(cw * 32 + c) ^ (h ? XA : 0)];
#pragma unroll
                for (int mb = 0; mb < MB; mb++)
                    if (mb < nb) acc[mb] = mfma32x2(av, sB[buf][2 * kk + h][(mb * 32 + c) ^ (h ? XB : 0)], acc[mb]);
            }
            if constexpr (GR) {  // this wave's k quarter (steps wave*4 .. wave*4+3) for the carried block
#pragma unroll
                for (int j = 0; j < KC / 8; j++) {
                    const int kk = wave * (KC / 8) + j;
                    gacc = mfma32x2(sB[buf][2 * kk + h][(gmb * 32 + c) ^ (h ? XB : 0)], sB[buf][2 * kk + h][(gnb * 32 + c) ^ (h ? XB : 0)], gacc);
                }
            }
        }
    };
    auto mainloop = [&](auto gram_tag) __attribute__((always_inline)) {
        if constexpr (DEEP) {  // the schedule of the H pass
            auto step = [&](int ch, float4 (&ran)[VL], float4 (&rbn)[BLOADS]) __attribute__((always_inline)) {
                chunk(gram_tag, ch);
                if (ch + 1 < nchunks) {
                    sstore(ran, rbn, (ch + 1) & 1);
                    if (ch + 3 < nchunks) gload(ran, rbn, kbeg + (ch + 3) * KC);
                }
                __syncthreads();
            };
            gload(ra0, rb0, kbeg);
            if (nchunks > 1) gload(ra1, rb1, kbeg + KC);
            sstore(ra0, rb0, 0);
            __syncthreads();
            if (nchunks > 2) gload(ra0, rb0, kbeg + 2 * KC);
            for (int ch = 0; ch < nchunks; ch += 2) {
                step(ch, ra1, rb1);
                if (ch + 1 < nchunks) step(ch + 1, ra0, rb0);
            }
        } else {
            gload(ra0, rb0, kbeg);
            sstore(ra0, rb0, 0);
            __syncthreads();
            for (int ch = 0; ch < nchunks; ch++) {
                if (ch + 1 < nchunks) gload(ra0, rb0, kbeg + (ch + 1) * KC);
                chunk(gram_tag, ch);
                if (ch + 1 < nchunks) sstore(ra0, rb0, (ch + 1) & 1);
                __syncthreads();
            }
        }
    };
    if (gcar)
        mainloop(std::true_type{});
    else
        mainloop(std::false_type{});
    if constexpr (KH == 2) pair_reduce<MB>(acc, &sB[0][0][0], cw, kh, lane);

    if constexpr (FUSE) {
        // acc[mb][v] = VHt(i = row0 + cw*32 + acc_row(v,h), a = mb*32 + c)
        constexpr int GS = RP + 1;
        float* sG = &sA[0][0][0];  // free after the loop's last barrier; sG[b*GS + a] = G(b, a)
        const int rpad = round_up(p.r, 4);
        for (int e = tid; e < RP * RP; e += NTHREADS) {
            const int b = e / RP, a = e - b * RP;
            float g = 0.f;
            if (a < p.r && b < p.r)
                for (int gs = 0; gs < p.gsl; gs++) g += p.G[(size_t)gs * p.r * rpad + (size_t)a * rpad + b];
            sG[b * GS + a] = g;
        }
        __syncthreads();
        if (!owner) return;
        const int il = cw * 32 + c;  // A operand row of W G: W(row0 + il, b = 2kk + h)
        float wa[MB * 16];
#pragma unroll
        for (int kk = 0; kk < MB * 16; kk++) {
            const int b = 2 * kk + h;
            wa[kk] = (il < nrows && b < p.r) ? p.W[(size_t)b * p.n + row0 + il] : 0.f;
        }
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
            if (mb >= nb) continue;
            f32x16 d;
            for (int v = 0; v < 16; v++) d[v] = 0.f;
#pragma unroll
            for (int kk = 0; kk < MB * 16; kk++)
                if (kk < nb * 16) d = mfma32x2(wa[kk], sG[(2 * kk + h) * GS + mb * 32 + c], d);
            const int a = mb * 32 + c;
            if (a >= p.r) continue;
            float* wcol = const_cast<float*>(p.W) + (size_t)a * p.n + row0;  // the problem's own W
#pragma unroll
            for (int g = 0; g < 4; g++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int i = cw * 32 + g * 8 + h * 4 + j;
                    if (i < nrows) wcol[i] *= (acc[mb][g * 4 + j] / (d[g * 4 + j] + p.eps));
                }
            }
        }
        return;
    }
    // acc[mb][v] = P(i = cw*32 + acc_row(v,h), a = mb*32 + c)
    const int orow0 = row0;
    const int rlim = p.npad - row0;
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
        const int a = mb * 32 + c;
        if (a >= p.r || !owner) continue;
        float* dst = p.P + (size_t)s * p.pstride + (size_t)a * p.ldp + orow0;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int il = cw * 32 + g * 8 + h * 4;
            if (il < rlim) {
                float4 o = make_float4(acc[mb][g * 4], acc[mb][g * 4 + 1], acc[mb][g * 4 + 2], acc[mb][g * 4 + 3]);
                *reinterpret_cast<float4*>(dst + il) = o;
            }
        }
    }
    if (gcar) {
        // gram(a, b) = H H^T(a, b) at P[s][b*ldp + npad + a]; reduced into G[b*rpad + a]
        const int rpad = round_up(p.r, 4);
        const int ldp = p.ldp, npad = p.npad;
        float* gred = &sA[0][0][0];  // free after the loop
        float* gdst = p.P + (size_t)s * p.pstride + npad;
#pragma unroll
        for (int v = 0; v < 16; v++) gred[(wave * 64 + lane) * 16 + v] = gacc[v];
        __syncthreads();
        for (int e = tid; e < 64 * 16; e += NTHREADS) {
            const int le = e >> 4, v = e & 15;
            const float g = ((gred[le * 16 + v] + gred[(64 + le) * 16 + v]) + gred[(128 + le) * 16 + v]) +
                            gred[(192 + le) * 16 + v];
            const int a = gmb * 32 + acc_row(v, le >> 5), b = gnb * 32 + (le & 31);
            if (a < rpad && b < p.r) gdst[(size_t)b * ldp + a] = g;
        }
        gram_finish(p.P, p.pstride, p.S, p.S * nblk, p.r * rpad,
                    [=](int e) { return (long)(e / rpad) * ldp + npad + (e % rpad); }, p.G, p.gcount);
    }
}

// ---------------------------------------------------------------------------
// W pass, KL rule: P[s](i,a) = sum_u V(i,u) H(a,u) / ((W H)(i,u) + eps).
// (W H)^T is formed per step as H^T W^T so that the ratio lands in the A-operand
// layout of the second product without a shuffle.
// FUSE: the workgroup applies W(i,a) *= num(i,a) / max(rowsum_a, eps) to its rows, the row sums
// of H taken over the whole of K in this item.
template <int MB, bool FUSE = false, int TW = TILE>
__device__ __forceinline__ void pass_w_kl_body(const PassArgs& p, const int w, const int nb) {
    constexpr int RP = MB * 32;
    static_assert(TW == 64 || TW == 128, "tile of 64 or 128 rows");
    constexpr int NCB = TW / 32;    // 32-row blocks of the tile
    constexpr int KH = 4 / NCB;     // waves per block, each on its k groups
    constexpr int VQ = TW / 4;      // float4 per chunk row u
    constexpr int VL = KC * VQ / NTHREADS;
    constexpr int PV = TW + 8;  // rows u and u+4 land 32 banks apart
    constexpr bool TRK = MB <= 2;  // the H^T W^T product reads H from a rank-contiguous copy (wh_rank order)
    __shared__ __attribute__((aligned(16))) float sV[2][KC][PV];   // [u][i]
    __shared__ __attribute__((aligned(16))) float sH[2][RP][LDK];  // [a][u]
    __shared__ __attribute__((aligned(16))) float sHu[TRK ? 2 : 1][TRK ? KC : 1][RP + 4];  // [u][a]
    static_assert(KH == 1 || 2 * KC * PV >= 2 * MB * 16 * 64, "pair reduction fits the V chunk buffers");

    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: branches on it stay scalar
    const int cw = wave % NCB, kh = wave / NCB;  // the wave's row block and k half (TW = 64)
    const bool owner = KH == 1 || kh == 0;
    const int c = lane & 31, h = lane >> 5;
    const int row0 = t * TW;
    const int nrows = min(TW, p.n - row0);
    const int kbeg = s * p.kslice;
    const int kend = min(p.m, kbeg + p.kslice);
    const bool vec = p.vec != 0;
    // H chunk (u in the chunk, all rank rows a): each thread loads 4 consecutive rank rows of
    // one column u (16 bytes) and stores them down its column of sH; consecutive threads take
    // consecutive u, so each of the 4 scalar stores of a wave lands on consecutive banks
    constexpr int BQ = RP / 4;
    constexpr int BLOADS = (KC * BQ + NTHREADS - 1) / NTHREADS;

    // B operand of H^T W^T: W(i, b) for i = this lane's row, b = wh_rank(kk, h) (TRK) or 2*kk + h.
    float wreg[MB * 16];
    {
        const int i = row0 + cw * 32 + c;
#pragma unroll
        for (int kk = 0; kk < MB * 16; kk++) {
            const int b = TRK ? wh_rank(kk, h) : 2 * kk + h;
            wreg[kk] = (i < p.n && b < p.r) ? p.W[(size_t)b * p.n + i] : 0.f;
        }
    }

    float4 ra[VL], rb[BLOADS];
    auto gload = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < VL; j++) {
            const int q = tid + NTHREADS * j, uu = q / VQ, iq = (q % VQ) * 4;
            ra[j] = load4(p.V + (size_t)(k0 + uu) * p.n + row0 + iq, k0 + uu < kend, iq, nrows, vec);
        }
#pragma unroll
        for (int j = 0; j < BLOADS; j++) {
            const int q = tid + NTHREADS * j, uu = q % KC, aq = (q / KC) * 4;
            rb[j] = load4(p.H + (size_t)(k0 + uu) * p.r + aq, (aq < RP) && (k0 + uu < kend), aq, p.r, vec);
        }
    };
    auto sstore = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < VL; j++) {
            const int q = tid + NTHREADS * j, uu = q / VQ, iq = (q % VQ) * 4;
            *reinterpret_cast<float4*>(&sV[buf][uu][iq]) = ra[j];
        }
#pragma unroll
        for (int j = 0; j < BLOADS; j++) {
            const int q = tid + NTHREADS * j, uu = q % KC, aq = (q / KC) * 4;
            if (aq < RP) {
                sH[buf][aq + 0][uu] = rb[j].x;
                sH[buf][aq + 1][uu] = rb[j].y;
                sH[buf][aq + 2][uu] = rb[j].z;
                sH[buf][aq + 3][uu] = rb[j].w;
                if constexpr (TRK) *reinterpret_cast<float4*>(&sHu[buf][uu][aq]) = rb[j];
            }
        }
    };

    f32x16 acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
        for (int v = 0; v < 16; v++) acc[mb][v] = 0.f;
    float rowsum = 0.f;  // sum over this slice of H(tid, u) (tile 0 only, or every tile when fused)
    const bool do_sums = (FUSE || t == 0) && (tid < RP);

    const int nchunks = (kend - kbeg + KC - 1) / KC;
    gload(kbeg);
    sstore(0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
        const int buf = ch & 1;
        if (ch + 1 < nchunks) gload(kbeg + (ch + 1) * KC);
        if (do_sums) {
#pragma unroll
            for (int j = 0; j < KC; j += 4) {
                const float4 x = *reinterpret_cast<const float4*>(&sH[buf][tid][j]);
                rowsum += (x.x + x.y) + (x.z + x.w);
            }
        }
        // whT(u, i): rows u = acc_row(v,h) of this step, column i = cw*32 + c; with 64-row
        // tiles the wave takes the k groups g of its half only.
        f32x16 wt;
        for (int v = 0; v < 16; v++) wt[v] = 0.f;
        if constexpr (TRK) {
#pragma unroll
            for (int blk = 0; blk < MB; blk++) {
                if (blk >= nb) continue;
#pragma unroll
                for (int q4 = 0; q4 < 4; q4++) {
                    const float4 h4 = *reinterpret_cast<const float4*>(&sHu[buf][c][blk * 32 + h * 16 + q4 * 4]);
                    const int kk = blk * 16 + q4 * 4;
                    wt = mfma32x2(h4.x, wreg[kk], wt);
                    wt = mfma32x2(h4.y, wreg[kk + 1], wt);
                    wt = mfma32x2(h4.z, wreg[kk + 2], wt);
                    wt = mfma32x2(h4.w, wreg[kk + 3], wt);
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < MB * 16; kk++)
                if (kk < nb * 16) wt = mfma32x2(sH[buf][2 * kk + h][c], wreg[kk], wt);
        }
#pragma unroll
        for (int g = 0; g < 4; g++) {
            if (KH == 2 && (g >> 1) != kh) continue;
            float rj[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float vv = sV[buf][g * 8 + h * 4 + j][cw * 32 + c];
                rj[j] = kl_ratio(vv, wt[g * 4 + j], p.eps);
            }
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                if (mb >= nb) continue;
                const float4 hb = *reinterpret_cast<const float4*>(&sH[buf][mb * 32 + c][g * 8 + h * 4]);
                acc[mb] = mfma32x2(rj[0], hb.x, acc[mb]);
                acc[mb] = mfma32x2(rj[1], hb.y, acc[mb]);
                acc[mb] = mfma32x2(rj[2], hb.z, acc[mb]);
                acc[mb] = mfma32x2(rj[3], hb.w, acc[mb]);
            }
        }
        if (ch + 1 < nchunks) sstore(buf ^ 1);
        __syncthreads();
    }
    if constexpr (KH == 2) pair_reduce<MB>(acc, &sV[0][0][0], cw, kh, lane);

    if constexpr (FUSE) {
        float* sS = &sH[0][0][0];  // free after the loop's last barrier
        if (do_sums) sS[tid] = (rowsum < p.eps) ? p.eps : rowsum;
        __syncthreads();
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
            const int a = mb * 32 + c;
            if (mb >= nb || a >= p.r || !owner) continue;
            float* wcol = const_cast<float*>(p.W) + (size_t)a * p.n + row0;  // the problem's own W
#pragma unroll
            for (int g = 0; g < 4; g++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int i = cw * 32 + g * 8 + h * 4 + j;
                    if (i < nrows) wcol[i] *= (acc[mb][g * 4 + j] / sS[a]);
                }
            }
        }
        return;
    }
    const int rlim = p.npad - row0;
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
        const int a = mb * 32 + c;
        if (a >= p.r || !owner) continue;
        float* dst = p.P + (size_t)s * p.pstride + (size_t)a * p.ldp + row0;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int il = cw * 32 + g * 8 + h * 4;
            if (il < rlim) {
                float4 o = make_float4(acc[mb][g * 4], acc[mb][g * 4 + 1], acc[mb][g * 4 + 2], acc[mb][g * 4 + 3]);
                *reinterpret_cast<float4*>(dst + il) = o;
            }
        }
    }
    if (do_sums && tid < p.r) p.Psum[(size_t)s * round_up(p.r, 4) + tid] = rowsum;
}

// ---------------------------------------------------------------------------
// Single-problem kernels: one workgroup per work item, XCD-aware order.
template <int MB, bool KL, int OBJ>
__global__ __launch_bounds__(NTHREADS, 2) void k_pass_h_f32(PassArgs p) {
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= p.tiles * p.S) return;
    pass_h_body<MB, KL, OBJ>(p, w, MB);  // MB = ceil(r/32)
}

template <int MB>
__global__ __launch_bounds__(NTHREADS, 2) void k_pass_w_euc_f32(PassArgs p) {
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= p.tiles * p.S) return;
    pass_w_euc_body<MB>(p, w, MB);
}

template <int MB>
__global__ __launch_bounds__(NTHREADS, 2) void k_pass_w_kl_f32(PassArgs p) {
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= p.tiles * p.S) return;
    pass_w_kl_body<MB>(p, w, MB);
}

// ---------------------------------------------------------------------------
// Grouped kernels: a persistent grid strides over the work items of every
// problem in the table.  Problems have r <= 64 (two rank blocks), and each runs
// exactly the code of its single-problem launch on its own workspace.
__device__ __forceinline__ int group_find(const int* first, int J, int item) {
    int lo = 0, hi = J - 1;  // largest j with first[j] <= item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// MB: rank blocks of 32 the kernel is built for (1: every problem of the table has r <= 32;
// 2: r <= 64); a problem runs ceil(r/32) of them.  TW: columns (H pass) or rows (W pass) per
// work item, 128 or 64; the table's problems are planned with tiles of that width.
template <bool KL, int OBJ, int MB, int TW>
__global__ __launch_bounds__(NTHREADS, 2) void k_group_pass_h(const JobArgs* __restrict__ jobs, const int* table) {
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find(first, J, item);
        const PassArgs p = jobs[idx[j]].h;
        __syncthreads();
        pass_h_body<MB, KL, OBJ, false, TW>(p, item - first[j], (p.r + 31) >> 5);
    }
}

template <bool KL, int MB, int TW>
__global__ __launch_bounds__(NTHREADS, 2) void k_group_pass_w(const JobArgs* __restrict__ jobs, const int* table) {
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find(first, J, item);
        const PassArgs p = jobs[idx[j]].w;
        __syncthreads();
        if (KL)
            pass_w_kl_body<MB, false, TW>(p, item - first[j], (p.r + 31) >> 5);
        else
            pass_w_euc_body<MB, false, TW>(p, item - first[j], (p.r + 31) >> 5);
    }
}

template <int MB, int TW>
static hipError_t launch_group_pass_h_mb(const JobArgs* jobs, const int* table, int kl, int obj, int grid,
                                         hipStream_t st) {
#define LG(K, O)                                                                                         \
    hipLaunchKernelGGL((k_group_pass_h<K, O, MB, TW>), dim3(grid), dim3(NTHREADS), 0, st, jobs, table); \
    return hipGetLastError();
    if (kl) {
        if (obj == OBJ_NONE) { LG(true, OBJ_NONE) }
        if (obj == OBJ_KL) { LG(true, OBJ_KL) }
    } else {
        if (obj == OBJ_NONE) { LG(false, OBJ_NONE) }
        if (obj == OBJ_FROB) { LG(false, OBJ_FROB) }
    }
#undef LG
    return hipErrorInvalidValue;
}

hipError_t launch_group_pass_h(const JobArgs* jobs, const int* table, int kl, int obj, int mb, int tw, int grid,
                               hipStream_t st) {
    if (mb == 1 && tw == 128) return launch_group_pass_h_mb<1, 128>(jobs, table, kl, obj, grid, st);
    if (mb == 2 && tw == 128) return launch_group_pass_h_mb<2, 128>(jobs, table, kl, obj, grid, st);
    if (mb == 1 && tw == 64) return launch_group_pass_h_mb<1, 64>(jobs, table, kl, obj, grid, st);
    if (mb == 2 && tw == 64) return launch_group_pass_h_mb<2, 64>(jobs, table, kl, obj, grid, st);
    return hipErrorInvalidValue;
}

template <int MB, int TW>
static hipError_t launch_group_pass_w_mb(const JobArgs* jobs, const int* table, int kl, int grid, hipStream_t st) {
    if (kl)
        hipLaunchKernelGGL((k_group_pass_w<true, MB, TW>), dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    else
        hipLaunchKernelGGL((k_group_pass_w<false, MB, TW>), dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    return hipGetLastError();
}

hipError_t launch_group_pass_w(const JobArgs* jobs, const int* table, int kl, int mb, int tw, int grid,
                               hipStream_t st) {
    if (mb == 1 && tw == 128) return launch_group_pass_w_mb<1, 128>(jobs, table, kl, grid, st);
    if (mb == 2 && tw == 128) return launch_group_pass_w_mb<2, 128>(jobs, table, kl, grid, st);
    if (mb == 1 && tw == 64) return launch_group_pass_w_mb<1, 64>(jobs, table, kl, grid, st);
    if (mb == 2 && tw == 64) return launch_group_pass_w_mb<2, 64>(jobs, table, kl, grid, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Fused full-K grouped launches (one slice per problem and pass, r <= 64): every
// work item is one column tile (H) or row tile (W) of one problem over the whole
// contraction, and applies that tile's update itself (pass bodies with FUSE).  The
// Euclidean updates take their gram from a gram launch that precedes the pass.

// G = X^T X over the contraction range [k0, k1) of one factor (k0 a multiple of 128), written
// G(a,b) at G[b*rpad + a]: WHICH 0: X = W (n x r, column-major, k = i); WHICH 1: X = H^T (H r x m,
// k = u).  Stages of KG contraction elements through LDS (loads of the next stage in flight);
// the 4 waves take interleaved k steps of every stage, and their partial blocks are added in
// wave order.
template <int WHICH>
__device__ __forceinline__ void gram_body(const float* __restrict__ X, int k0, int k1, int r, int ld,
                                          float* __restrict__ G, bool vec) {
    constexpr int RP = 64, KG = 128, XP = RP + 4;
    constexpr int LOADS = RP * KG / 4 / NTHREADS;  // float4 per thread per stage
    __shared__ __attribute__((aligned(16))) float sX[KG][XP];  // sX[k][a] = X(k, a)
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, h = lane >> 5;
    const int nb = (r + 31) >> 5;
    float4 rx[LOADS];
    auto gload = [&](int ks) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < LOADS; j++) {
            const int q = tid + NTHREADS * j;
            if (WHICH == 0) {  // W(k, a) = W[a*ld + k]: 4 consecutive k of row a
                const int a = q / (KG / 4), kq = (q % (KG / 4)) * 4;
                rx[j] = load4(X + (size_t)a * ld + ks + kq, a < r, ks + kq, k1, vec);
            } else {  // H(a, k) = H[k*ld + a]: 4 consecutive a of column k
                const int k = q / (RP / 4), aq = (q % (RP / 4)) * 4;
                rx[j] = load4(X + (size_t)(ks + k) * ld + aq, ks + k < k1, aq, r, vec);
            }
        }
    };
    auto sstore = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < LOADS; j++) {
            const int q = tid + NTHREADS * j;
            if (WHICH == 0) {
                const int a = q / (KG / 4), kq = (q % (KG / 4)) * 4;
                sX[kq][a] = rx[j].x;
                sX[kq + 1][a] = rx[j].y;
                sX[kq + 2][a] = rx[j].z;
                sX[kq + 3][a] = rx[j].w;
            } else {
                const int k = q / (RP / 4), aq = (q % (RP / 4)) * 4;
                *reinterpret_cast<float4*>(&sX[k][aq]) = rx[j];
            }
        }
    };
    f32x16 gacc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
            for (int v = 0; v < 16; v++) gacc[x][y][v] = 0.f;
    const int stages = (k1 - k0 + KG - 1) / KG;
    if (stages > 0) gload(k0);
    for (int st = 0; st < stages; st++) {
        sstore();
        __syncthreads();
        if (st + 1 < stages) gload(k0 + (st + 1) * KG);
#pragma unroll
        for (int j = 0; j < KG / 8; j++) {
            const int k = 2 * (wave + 4 * j) + h;
#pragma unroll
            for (int x = 0; x < 2; x++) {
                if (x >= nb) continue;
                const float av = sX[k][x * 32 + c];
#pragma unroll
                for (int y = 0; y < 2; y++)
                    if (y < nb) gacc[x][y] = mfma32x2(av, sX[k][y * 32 + c], gacc[x][y]);
            }
        }
        __syncthreads();
    }
    // gacc[x][y][v] = partial G(a = x*32 + acc_row(v,h), b = y*32 + c) of this wave
    float* red = &sX[0][0];  // 4 waves x 64 lanes x 16 (free after the last barrier)
    const int rpad = round_up(r, 4);
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            if (x >= nb || y >= nb) continue;
#pragma unroll
            for (int v = 0; v < 16; v++) red[(wave * 64 + lane) * 16 + v] = gacc[x][y][v];
            __syncthreads();
            for (int e = tid; e < 64 * 16; e += NTHREADS) {
                const int le = e >> 4, v = e & 15;
                const float g = ((red[le * 16 + v] + red[(64 + le) * 16 + v]) + red[(128 + le) * 16 + v]) +
                                red[(192 + le) * 16 + v];
                const int a = x * 32 + acc_row(v, le >> 5), b = y * 32 + (le & 31);
                if (a < r && b < r) G[(size_t)b * rpad + a] = g;
            }
            __syncthreads();
        }
}

template <int WHICH>
__global__ __launch_bounds__(NTHREADS) void k_group_gram(const JobArgs* __restrict__ jobs, const int* table) {
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find(first, J, item);
        const JobArgs& ja = jobs[idx[j]];
        const PassArgs& p = WHICH == 0 ? ja.h : ja.w;
        // K slice s of the gram: [s*gks, min(K, (s+1)*gks)), its partial at G + s*r*rpad
        const int s = item - first[j], K = WHICH == 0 ? p.n : p.m;
        const int k0 = s * p.gks, k1 = min(K, k0 + p.gks);
        float* G = p.G + (size_t)s * p.r * round_up(p.r, 4);
        __syncthreads();
        if (WHICH == 0)
            gram_body<0>(p.W, k0, k1, p.r, p.n, G, p.vec != 0);
        else
            gram_body<1>(p.H, k0, k1, p.r, p.r, G, p.vec != 0);
    }
}

template <bool KL, int OBJ, int MB, int TW>
__global__ __launch_bounds__(NTHREADS, 2) void k_group_fused_h(const JobArgs* __restrict__ jobs, const int* table) {
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find(first, J, item);
        const PassArgs p = jobs[idx[j]].h;
        __syncthreads();
        pass_h_body<MB, KL, OBJ, true, TW>(p, item - first[j], (p.r + 31) >> 5);
    }
}

template <bool KL, int MB, int TW>
__global__ __launch_bounds__(NTHREADS, 2) void k_group_fused_w(const JobArgs* __restrict__ jobs, const int* table) {
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find(first, J, item);
        const PassArgs p = jobs[idx[j]].w;
        __syncthreads();
        if (KL)
            pass_w_kl_body<MB, true, TW>(p, item - first[j], (p.r + 31) >> 5);
        else
            pass_w_euc_body<MB, true, TW>(p, item - first[j], (p.r + 31) >> 5);
    }
}

hipError_t launch_group_gram(const JobArgs* jobs, const int* table, int which, int grid, hipStream_t st) {
    if (which == 0)
        hipLaunchKernelGGL(k_group_gram<0>, dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    else
        hipLaunchKernelGGL(k_group_gram<1>, dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    return hipGetLastError();
}

template <int MB, int TW>
static hipError_t launch_group_fused_h_mb(const JobArgs* jobs, const int* table, int kl, int obj, int grid,
                                          hipStream_t st) {
#define LF(K, O)                                                                                          \
    hipLaunchKernelGGL((k_group_fused_h<K, O, MB, TW>), dim3(grid), dim3(NTHREADS), 0, st, jobs, table); \
    return hipGetLastError();
    if (kl) {
        if (obj == OBJ_NONE) { LF(true, OBJ_NONE) }
        if (obj == OBJ_KL) { LF(true, OBJ_KL) }
    } else {
        if (obj == OBJ_NONE) { LF(false, OBJ_NONE) }
        if (obj == OBJ_FROB) { LF(false, OBJ_FROB) }
    }
#undef LF
    return hipErrorInvalidValue;
}

hipError_t launch_group_fused_h(const JobArgs* jobs, const int* table, int kl, int obj, int mb, int tw, int grid,
                                hipStream_t st) {
    if (mb == 1 && tw == 128) return launch_group_fused_h_mb<1, 128>(jobs, table, kl, obj, grid, st);
    if (mb == 2 && tw == 128) return launch_group_fused_h_mb<2, 128>(jobs, table, kl, obj, grid, st);
    if (mb == 1 && tw == 64) return launch_group_fused_h_mb<1, 64>(jobs, table, kl, obj, grid, st);
    if (mb == 2 && tw == 64) return launch_group_fused_h_mb<2, 64>(jobs, table, kl, obj, grid, st);
    return hipErrorInvalidValue;
}

template <int MB, int TW>
static hipError_t launch_group_fused_w_mb(const JobArgs* jobs, const int* table, int kl, int grid, hipStream_t st) {
    if (kl)
        hipLaunchKernelGGL((k_group_fused_w<true, MB, TW>), dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    else
        hipLaunchKernelGGL((k_group_fused_w<false, MB, TW>), dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    return hipGetLastError();
}

hipError_t launch_group_fused_w(const JobArgs* jobs, const int* table, int kl, int mb, int tw, int grid,
                                hipStream_t st) {
    if (mb == 1 && tw == 128) return launch_group_fused_w_mb<1, 128>(jobs, table, kl, grid, st);
    if (mb == 2 && tw == 128) return launch_group_fused_w_mb<2, 128>(jobs, table, kl, grid, st);
    if (mb == 1 && tw == 64) return launch_group_fused_w_mb<1, 64>(jobs, table, kl, grid, st);
    if (mb == 2 && tw == 64) return launch_group_fused_w_mb<2, 64>(jobs, table, kl, grid, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Host launchers.
template <int MB>
static hipError_t launch_h_mb(const PassArgs& a, int kl, int obj, hipStream_t st) {
    const int grid = xcd_grid(a.tiles * a.S);
#define LH(KLV, OBJV)                                                                          \
    hipLaunchKernelGGL((k_pass_h_f32<MB, KLV, OBJV>), dim3(grid), dim3(NTHREADS), 0, st, a); \
    return hipGetLastError();
    if (!kl) {
        // the Euclidean pass carries no objective (it would need W H for that alone: the
        // objective pass on split factors computes it)
        if (obj != OBJ_NONE) return hipErrorInvalidValue;
        LH(false, OBJ_NONE)
    }
    switch (obj) {
        case OBJ_NONE: LH(true, OBJ_NONE)
        case OBJ_FROB: LH(true, OBJ_FROB)
        case OBJ_KL: LH(true, OBJ_KL)
        case OBJ_QUALITY: LH(true, OBJ_QUALITY)
        default: return hipErrorInvalidValue;
    }
#undef LH
}

hipError_t launch_pass_h_f32(const PassArgs& a, int kl, int obj, hipStream_t st) {
    const int rp = round_up(a.r, 32);
    if (rp <= 32) return launch_h_mb<1>(a, kl, obj, st);
    if (rp <= 64) return launch_h_mb<2>(a, kl, obj, st);
    if (rp <= 96) return launch_h_mb<3>(a, kl, obj, st);
    if (rp <= 128) return launch_h_mb<4>(a, kl, obj, st);
    return hipErrorInvalidValue;
}

hipError_t launch_pass_w_f32(const PassArgs& a, int kl, hipStream_t st) {
    const int rp = round_up(a.r, 32);
    const int grid = xcd_grid(a.tiles * a.S);
#define LW(K, MBV)                                                                      \
    hipLaunchKernelGGL((K<MBV>), dim3(grid), dim3(NTHREADS), 0, st, a); \
    return hipGetLastError();
    if (kl) {
        if (rp <= 32) { LW(k_pass_w_kl_f32, 1) }
        if (rp <= 64) { LW(k_pass_w_kl_f32, 2) }
        if (rp <= 96) { LW(k_pass_w_kl_f32, 3) }
        if (rp <= 128) { LW(k_pass_w_kl_f32, 4) }
    } else {
        if (rp <= 32) { LW(k_pass_w_euc_f32, 1) }
        if (rp <= 64) { LW(k_pass_w_euc_f32, 2) }
        if (rp <= 96) { LW(k_pass_w_euc_f32, 3) }
        if (rp <= 128) { LW(k_pass_w_euc_f32, 4) }
    }
#undef LW
    return hipErrorInvalidValue;
}

}  // namespace gpumlib_nmf
