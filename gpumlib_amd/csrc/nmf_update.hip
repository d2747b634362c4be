// Update ("epilogue") kernels and the reference-shaped element-wise / reduction
// kernels.  All reductions are in a fixed order, so results do not depend on
// timing or on what else runs on the device.
#include "nmf_common.hpp"
#include "nmf_fp8.hpp"
#include "nmf_internal.hpp"

namespace gpumlib_nmf {


// ---------------------------------------------------------------------------
// Sum of S slice partials at stride pstride, in slice order (loads unrolled).
__device__ __forceinline__ float sum_slices(const float* src, int S, long pstride) {
    float v = 0.f;
    int s = 0;
    for (; s + 4 <= S; s += 4) {
        const float x0 = src[(size_t)s * pstride], x1 = src[(size_t)(s + 1) * pstride];
        const float x2 = src[(size_t)(s + 2) * pstride], x3 = src[(size_t)(s + 3) * pstride];
        v += x0;
        v += x1;
        v += x2;
        v += x3;
    }
    for (; s < S; s++) v += src[(size_t)s * pstride];
    return v;
}

// Euclidean H update, reference nmf_multiplicative_euclidian.cpp:36-45 and
// nmf_kernels.cu:46:  H(a,u) *= WtV(a,u) / (WtWH(a,u) + eps), WtWH = WtW H.
// WtV(a,u) = sum_s P[s][u*ldr + a];  WtW(a,b) = G[b*ldr + a] (reduced by the pass).
__device__ __forceinline__ void update_h_euc_body(float* __restrict__ H, const float* __restrict__ P,
                                                  const float* __restrict__ G, int m, int r, int ldr, int S,
                                                  long pstride, float eps, int blk, float* sm) {
    const int rs = r + 1;
    float* sG = sm;           // sG[a*rs + b] = WtW(a,b)
    float* sH = sm + r * rs;  // sH[uu*rs + b] = H(b, u0 + uu)
    const int tid = threadIdx.x;
    const int u0 = blk * UPD_COLS;
    const int ncol = min(UPD_COLS, m - u0);
    for (int idx = tid; idx < r * r; idx += NTHREADS) {
        const int b = idx / r, a = idx - b * r;
        sG[a * rs + b] = G[(size_t)b * ldr + a];
    }
    for (int idx = tid; idx < ncol * r; idx += NTHREADS) {
        const int uu = idx / r, a = idx - uu * r;
        sH[uu * rs + a] = H[(size_t)(u0 + uu) * r + a];
    }
    __syncthreads();
    for (int idx = tid; idx < ncol * r; idx += NTHREADS) {
        const int uu = idx / r, a = idx - uu * r;
        const int u = u0 + uu;
        const float wtv = sum_slices(P + (size_t)u * ldr + a, S, pstride);
        float d = 0.f;
        for (int b = 0; b < r; b++) d += sG[a * rs + b] * sH[uu * rs + b];
        H[(size_t)u * r + a] = sH[uu * rs + a] * (wtv / (d + eps));
    }
}

// Euclidean W update, reference nmf_multiplicative_euclidian.cpp:53-66:
// W(i,a) *= VHt(i,a) / (WHHt(i,a) + eps), WHHt = W HHt.
// VHt(i,a) = sum_s P[s][a*ldp + i];  HHt(b,a) = G[a*rpad + b] (reduced by the pass).
__device__ __forceinline__ void update_w_euc_body(float* __restrict__ W, const float* __restrict__ P,
                                                  const float* __restrict__ G, int n, int r, int ldp, int S,
                                                  long pstride, float eps, int blk, float* sm) {
    const int rs = r + 1, rpad = round_up(r, 4);
    float* sG = sm;           // sG[b*rs + a] = HHt(b,a)
    float* sW = sm + r * rs;  // sW[ii*rs + b] = W(i0 + ii, b)
    const int tid = threadIdx.x;
    const int i0 = blk * UPD_ROWS;
    const int nrow = min(UPD_ROWS, n - i0);
    for (int idx = tid; idx < r * r; idx += NTHREADS) {
        const int a = idx / r, b = idx - a * r;
        sG[b * rs + a] = G[(size_t)a * rpad + b];
    }
    for (int idx = tid; idx < UPD_ROWS * r; idx += NTHREADS) {
        const int ii = idx % UPD_ROWS, b = idx / UPD_ROWS;
        if (ii < nrow) sW[ii * rs + b] = W[(size_t)b * n + i0 + ii];
    }
    __syncthreads();
    for (int idx = tid; idx < UPD_ROWS * r; idx += NTHREADS) {
        const int ii = idx % UPD_ROWS, a = idx / UPD_ROWS;
        if (ii >= nrow) continue;
        const int i = i0 + ii;
        const float vht = sum_slices(P + (size_t)a * ldp + i, S, pstride);
        float d = 0.f;
        for (int b = 0; b < r; b++) d += sW[ii * rs + b] * sG[b * rs + a];
        W[(size_t)a * n + i] = sW[ii * rs + a] * (vht / (d + eps));
    }
}

// ---------------------------------------------------------------------------
// Staging of a small operand into LDS: every thread issues U loads before it stores
// any of them, so a workgroup pays about one memory round trip per U * NTHREADS
// elements instead of one per NTHREADS.
template <int U, typename Ld, typename St>
__device__ __forceinline__ void stage_lds(int total, Ld ld, St st) {
    for (int base = threadIdx.x; base < total; base += U * NTHREADS) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = base + u * NTHREADS;
            v[u] = idx < total ? ld(idx) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = base + u * NTHREADS;
            if (idx < total) st(idx, v[u]);
        }
    }
}

// Two-phase staging of a small operand: load() issues this thread's U loads into
// registers, store() writes them to LDS.  Loads issued between the two phases share
// the same memory round trip.
template <int U>
struct RegStage {
    float v[U];
    template <typename Ld>
    __device__ __forceinline__ void load(int total, Ld ld) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = threadIdx.x + u * NTHREADS;
            v[u] = idx < total ? ld(idx) : 0.f;
        }
    }
    template <typename St>
    __device__ __forceinline__ void store(int total, St st) const {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = threadIdx.x + u * NTHREADS;
            if (idx < total) st(idx, v[u]);
        }
    }
};

// Register staging in 16-byte groups: group g covers elements 4g .. 4g+3.  ptr(g) is
// the group's 16-byte-aligned address (nullptr: zeros); mask(g, j) says whether
// element j of the group exists (the others are staged as zeros).
template <int U>
struct RegStage4 {
    float4 v[U];
    // v = sum over ns <= kMaxGramSlices slices (stride ps floats) of the group at ptr(g),
    // in slice order; all ns U loads are issued before the first add.
    template <typename Ptr, typename Mask>
    __device__ __forceinline__ void load_sum(int groups, int ns, long ps, Ptr ptr, Mask mask) {
        constexpr int MS = kMaxGramSlices;
        float4 x[U][MS];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int g = threadIdx.x + u * NTHREADS;
            const float* q = g < groups ? ptr(g) : nullptr;
#pragma unroll
            for (int sl = 0; sl < MS; sl++)
                x[u][sl] = (q && sl < ns) ? *reinterpret_cast<const float4*>(q + (size_t)sl * ps)
                                          : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            float4 t = x[u][0];
#pragma unroll
            for (int sl = 1; sl < MS; sl++)
                if (sl < ns) t.x += x[u][sl].x, t.y += x[u][sl].y, t.z += x[u][sl].z, t.w += x[u][sl].w;
            const int g = threadIdx.x + u * NTHREADS;
            v[u] = make_float4(mask(g, 0) ? t.x : 0.f, mask(g, 1) ? t.y : 0.f, mask(g, 2) ? t.z : 0.f,
                               mask(g, 3) ? t.w : 0.f);
        }
    }
    template <typename St>
    __device__ __forceinline__ void store(int groups, St st) const {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int g = threadIdx.x + u * NTHREADS;
            if (g < groups) {
                st(4 * g, v[u].x);
                st(4 * g + 1, v[u].y);
                st(4 * g + 2, v[u].z);
                st(4 * g + 3, v[u].w);
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Single-problem Euclidean updates, tiled so that every thread issues all of its
// slice loads before it waits (the update is a short, latency-bound pass over
// the partials: one round trip per workgroup, not one per element).
//
// Row form, for partials laid out P[s][j*ldr + a] (a contiguous per row j):
//   X(j,a) <- X(j,a) * (N(j,a) / (sum_b X(j,b) C(b,a) + eps)),  N(j,a) = sum_s P[s][j*ldr + a]
// H update (nmf_multiplicative_euclidian.cpp:36-45, nmf_kernels.cu:46): j = u,
//   X(u,a) = H(a,u) at H[u*r + a], C(b,a) = WtW(a,b) = G[b*ldg + a]  (gt = 0)
// W update from a pass over V^T (nmf_multiplicative_euclidian.cpp:53-66): j = i,
//   X(i,a) = W(i,a) at W[a*n + i], C(b,a) = HHt(b,a) = G[a*ldg + b]  (gt = 1)
// Sh (optional) receives the BF16 shadow of the new X: X(j,a) at Sh[a*ldsh + j]; with
// shplane > 0, its three BF16 pieces (bf16x2_split3) in planes shplane elements apart, the
// workgroup's rows past count written as zeros (ldsh, a multiple of 128, covers them).
// Each thread owns (row jj, 4 consecutive a); JT rows per workgroup.  RP is r rounded
// up to 32: the staged C and X are zero beyond r, so the inner product runs over RP
// terms (the added terms are exact zeros), on the matrix cores.
template <int JT, int RP>
__global__ __launch_bounds__(NTHREADS) void k_update_rows_euc(float* __restrict__ X, long xs_j, long xs_a,
                                                              const float* __restrict__ P, const float* __restrict__ G,
                                                              int ldg, int gt, int count, int r, int ldr, int S,
                                                              long pstride, float eps, uint16_t* __restrict__ Sh,
                                                              int ldsh, uint8_t* __restrict__ Q, int ldq,
                                                              uint8_t* __restrict__ Qsc, int nblk, int add,
                                                              long gps, int gsl, long shplane) {
    extern __shared__ __attribute__((aligned(16))) float smr[];
    constexpr int RS = RP + 4, NQ = RP / 4, ITEMS = JT * NQ, MAXI = (ITEMS + NTHREADS - 1) / NTHREADS;
    float* sG = smr;            // sG[b*RS + a] = C(b,a), zero for a or b >= r
    float* sX = sG + RP * RS;   // sX[jj*RS + b] = X(j0+jj, b), zero for b >= r
    float* sN = sX + JT * RS;   // new X(j0+jj, a)
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * JT, nj = min(JT, count - j0);
    const int rq = round_up(r, 4) / 4;  // float4 groups of a row of P (ldr = 4*rq)

    // C and the X rows, zero padded to RP.  Up to r = 64 their loads are issued first,
    // into registers, so that they and the slice loads below take one round trip.
    // gsl = 0: G is the reduced gram; gsl in [1, kMaxGramSlices] (r <= 64 only): G holds the grams of
    // the pass's gsl K slices at stride gps, summed here in slice order.
    auto ld_g = [&](int idx) {  // C(b,a) in staging order: along b for gt (coalesced), along a otherwise
        const int hi = idx / RP, lo = idx - hi * RP;
        return (hi < r && lo < r) ? G[(size_t)hi * ldg + lo] : 0.f;
    };
    // 16-byte groups of C in the same order (RP and ldg are multiples of 4, G 16-byte aligned)
    auto ptr_g4 = [&](int g) -> const float* {
        const int hi = (4 * g) / RP, lo = 4 * g - hi * RP;
        return (hi < r && lo < r) ? G + (size_t)hi * ldg + lo : nullptr;
    };
    auto mask_g4 = [&](int g, int j) {
        const int hi = (4 * g) / RP, lo = 4 * g - hi * RP + j;
        return hi < r && lo < r;
    };
    auto st_g = [&](int idx, float v) {
        const int hi = idx / RP, lo = idx - hi * RP;
        if (gt) sG[lo * RS + hi] = v; else sG[hi * RS + lo] = v;
    };
    auto ld_x = [&](int idx) {
        if (xs_a == 1) {
            const int jj = idx / RP, b = idx - jj * RP;
            return (jj < nj && b < r) ? X[(size_t)(j0 + jj) * xs_j + b] : 0.f;
        }
        const int jj = idx % JT, b = idx / JT;
        return (jj < nj && b < r) ? X[(size_t)(j0 + jj) * xs_j + (size_t)b * xs_a] : 0.f;
    };
    auto st_x = [&](int idx, float v) {
        if (xs_a == 1) {
            const int jj = idx / RP, b = idx - jj * RP;
            sX[jj * RS + b] = v;
        } else {
            const int jj = idx % JT, b = idx / JT;
            sX[jj * RS + b] = v;
        }
    };
    constexpr int UG = (RP * RP + NTHREADS - 1) / NTHREADS, UX = (JT * RP + NTHREADS - 1) / NTHREADS;
    constexpr int UG4 = (RP * RP / 4 + NTHREADS - 1) / NTHREADS;
    constexpr bool EARLY = UG <= 16;
    RegStage4<EARLY ? UG4 : 1> stg_g;
    RegStage<EARLY ? UX : 1> stg_x;
    if constexpr (EARLY) {
        stg_g.load_sum(RP * RP / 4, gsl > 0 ? gsl : 1, gps, ptr_g4, mask_g4);
        stg_x.load(JT * RP, ld_x);
    }

    // 1. issue the slice loads of this thread's items
    float4 num[MAXI];
#pragma unroll
    for (int k = 0; k < MAXI; k++) {
        num[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        const int item = tid + k * NTHREADS, jj = item / NQ, q = item - jj * NQ;
        if (item >= ITEMS || jj >= nj || q >= rq) continue;
        const float4* src = reinterpret_cast<const float4*>(P + (size_t)(j0 + jj) * ldr) + q;
        const long ps4 = pstride >> 2;
        int s = 0;
        for (; s + 4 <= S; s += 4) {
            const float4 x0 = src[(size_t)s * ps4], x1 = src[(size_t)(s + 1) * ps4];
            const float4 x2 = src[(size_t)(s + 2) * ps4], x3 = src[(size_t)(s + 3) * ps4];
            num[k].x += x0.x, num[k].y += x0.y, num[k].z += x0.z, num[k].w += x0.w;
            num[k].x += x1.x, num[k].y += x1.y, num[k].z += x1.z, num[k].w += x1.w;
            num[k].x += x2.x, num[k].y += x2.y, num[k].z += x2.z, num[k].w += x2.w;
            num[k].x += x3.x, num[k].y += x3.y, num[k].z += x3.z, num[k].w += x3.w;
        }
        for (; s < S; s++) {
            const float4 x0 = src[(size_t)s * ps4];
            num[k].x += x0.x, num[k].y += x0.y, num[k].z += x0.z, num[k].w += x0.w;
        }
    }
    // 2. C and the X rows into LDS
    if constexpr (EARLY) {
        stg_g.store(RP * RP / 4, st_g);
        stg_x.store(JT * RP, st_x);
    } else {
        stage_lds<8>(RP * RP, ld_g, st_g);
        stage_lds<8>(JT * RP, ld_x, st_x);
    }
    __syncthreads();
    // 3. denominators D = X C on the matrix cores: 16 x 16 blocks (row block jb of the
    // tile, column block ab of a) dealt to the 4 waves in turn, the whole contraction in
    // order per block (16x16x4 FP32 MFMAs: exact products, FP32 accumulation).  Small
    // blocks keep every wave busy and each dependent chain short (RP / 4 MFMAs).
    // D goes to sN, where each item then replaces its own d by the new value.
    {
        constexpr int NJB = JT / 16, NAB = RP / 16;
        const int lane = tid & 63, c = lane & 15, g = lane >> 4;
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        for (int blk = wave; blk < NJB * NAB; blk += NTHREADS / 64) {
            const int jb = blk / NAB, ab = blk - jb * NAB;
            // lane (c, g): A = X(row jb*16 + c, k + g), B = C(k + g, col ab*16 + c);
            // result register v = D(row jb*16 + 4g + v, col ab*16 + c)
            const float* xa = sX + (jb * 16 + c) * RS + g;
            const float* gb = sG + g * RS + ab * 16 + c;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
            for (int k = 0; k < RP; k += 4)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[k], gb[k * RS], acc, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; v++) sN[(jb * 16 + 4 * g + v) * RS + ab * 16 + c] = acc[v];
        }
    }
    __syncthreads();
    // 4. new values
#pragma unroll
    for (int k = 0; k < MAXI; k++) {
        const int item = tid + k * NTHREADS, jj = item / NQ, q = item - jj * NQ;
        if (item >= ITEMS || jj >= nj) continue;
        const float* xr = sX + jj * RS;
        const float4 d = *reinterpret_cast<const float4*>(sN + jj * RS + 4 * q);
        const float4 xo = *reinterpret_cast<const float4*>(xr + 4 * q);
        float4 o;
        if (add) {
            o.x = additive_rule(xo.x, num[k].x, d.x);
            o.y = additive_rule(xo.y, num[k].y, d.y);
            o.z = additive_rule(xo.z, num[k].z, d.z);
            o.w = additive_rule(xo.w, num[k].w, d.w);
        } else {
            o.x = xo.x * (num[k].x / (d.x + eps));
            o.y = xo.y * (num[k].y / (d.y + eps));
            o.z = xo.z * (num[k].z / (d.z + eps));
            o.w = xo.w * (num[k].w / (d.w + eps));
        }
        *reinterpret_cast<float4*>(sN + jj * RS + 4 * q) = o;
    }
    __syncthreads();
    // 5. coalesced writes of X and of its shadow
    if (xs_a == 1) {
        for (int idx = tid; idx < nj * r; idx += NTHREADS) {
            const int jj = idx / r, a = idx - jj * r;
            X[(size_t)(j0 + jj) * xs_j + a] = sN[jj * RS + a];
        }
    } else {
        for (int idx = tid; idx < JT * r; idx += NTHREADS) {
            const int jj = idx % JT, a = idx / JT;
            if (jj < nj) X[(size_t)(j0 + jj) * xs_j + (size_t)a * xs_a] = sN[jj * RS + a];
        }
    }
    if (Sh && shplane) {
        // pairs of rows: 4-byte aligned (j0 and ldsh are even)
        for (int idx = tid; idx < JT / 2 * r; idx += NTHREADS) {
            const int jj = 2 * (idx % (JT / 2)), a = idx / (JT / 2);
            const float x0 = jj < nj ? sN[jj * RS + a] : 0.f, x1 = jj + 1 < nj ? sN[(jj + 1) * RS + a] : 0.f;
            uint32_t w1, w2, w3;
            bf16x2_split3(x0, x1, w1, w2, w3);
            uint32_t* dst = reinterpret_cast<uint32_t*>(Sh + (size_t)a * ldsh + j0 + jj);
            dst[0] = w1;
            dst[shplane / 2] = w2;
            dst[shplane] = w3;
        }
    } else if (Sh) {
        for (int idx = tid; idx < JT * r; idx += NTHREADS) {
            const int jj = idx % JT, a = idx / JT;
            if (jj < nj) Sh[(size_t)a * ldsh + j0 + jj] = __builtin_bit_cast(uint16_t, (__bf16)sN[jj * RS + a]);
        }
    }
    if constexpr (JT == 32) {
        // FP8 shadow: the workgroup's 32 rows are one scale block of every rank column a
        // (X(j,a) at Q[a*ldq + j], scale at Qsc[a*nblk + j/32]); rows past count are zeros
        // Four consecutive lanes share a column a (r <= 64 here, so 256 threads cover every
        // column once): each takes 8 of the 32 rows, the column's amax is combined across
        // the four lanes, and each lane writes its 8 bytes.
        if (Q) {
            const int a = tid >> 2, part = tid & 3;
            float v8[8];
            float amax = 0.f;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const int jj = 8 * part + b;
                v8[b] = (a < r && jj < nj) ? sN[jj * RS + a] : 0.f;
                amax = fmaxf(amax, fabsf(v8[b]));
            }
            amax = fmaxf(amax, __shfl_xor(amax, 1, 64));
            amax = fmaxf(amax, __shfl_xor(amax, 2, 64));
            if (a < r) {
                const int sc = e8m0_for_amax(amax);
                if (part == 0) Qsc[(size_t)a * nblk + j0 / 32] = (uint8_t)sc;
                uint32_t word[2];
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    uint32_t w = 0;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int jj = 8 * part + 4 * q + b;
                        const uint32_t e = jj < nj ? f32_to_e4m3(e8m0_unscale(v8[4 * q + b], sc)) : 0u;
                        w |= e << (8 * b);
                    }
                    word[q] = w;
                }
                *reinterpret_cast<uint2*>(Q + (size_t)a * ldq + j0 + 8 * part) = make_uint2(word[0], word[1]);
            }
        }
    }
}

// Column form, for the FP32 W pass layout P[s][a*ldp + i] (i contiguous per a):
//   W(i,a) <- W(i,a) * (VHt(i,a) / (sum_b W(i,b) HHt(b,a) + eps)),  HHt(b,a) = G[a*rpad + b]
// (nmf_multiplicative_euclidian.cpp:53-66).  JT rows i per workgroup; each thread owns
// 4 consecutive rows i of one rank column a (i fastest across lanes), issues all its
// slice loads before waiting, and runs the inner product over RP (r rounded up to 32)
// zero-padded terms, fully unrolled.  W rows are staged transposed (sWt[b][i]) so the
// 4 rows of a thread are one 16-byte LDS read.
template <int JT, int RP>
__global__ __launch_bounds__(NTHREADS) void k_update_w_euc_cols(float* __restrict__ W, const float* __restrict__ P,
                                                                const float* __restrict__ G, int n, int r, int ldp,
                                                                int S, long pstride, float eps, int add) {
    extern __shared__ __attribute__((aligned(16))) float smc[];
    constexpr int RS = RP + 4, JS = JT + 4, NI4 = JT / 4, ITEMS = NI4 * RP, MAXI = (ITEMS + NTHREADS - 1) / NTHREADS;
    float* sG = smc;             // sG[b*RS + a] = HHt(b,a), zero for a or b >= r
    float* sWt = sG + RP * RS;   // sWt[b*JS + ii] = W(i0+ii, b), zero outside
    const int tid = threadIdx.x;
    const int rpad = round_up(r, 4);
    const int i0 = blockIdx.x * JT, nrow = min(JT, n - i0);

    // HHt and the W rows, zero padded; up to r = 64 their loads go out before the slice
    // loads (one memory round trip for all of them).
    auto ld_g = [&](int idx) {
        const int a = idx / RP, b = idx - a * RP;
        return (a < r && b < r) ? G[(size_t)a * rpad + b] : 0.f;
    };
    auto st_g = [&](int idx, float v) {
        const int a = idx / RP, b = idx - a * RP;
        sG[b * RS + a] = v;
    };
    auto ld_w = [&](int idx) {
        const int ii = idx % JT, b = idx / JT;
        return (ii < nrow && b < r) ? W[(size_t)b * n + i0 + ii] : 0.f;
    };
    auto st_w = [&](int idx, float v) {
        const int ii = idx % JT, b = idx / JT;
        sWt[b * JS + ii] = v;
    };
    constexpr int UG = (RP * RP + NTHREADS - 1) / NTHREADS, UW = (RP * JT + NTHREADS - 1) / NTHREADS;
    constexpr bool EARLY = UG <= 16;
    RegStage<EARLY ? UG : 1> stg_g;
    RegStage<EARLY ? UW : 1> stg_w;
    if constexpr (EARLY) {
        stg_g.load(RP * RP, ld_g);
        stg_w.load(RP * JT, ld_w);
    }

    float4 num[MAXI];
#pragma unroll
    for (int k = 0; k < MAXI; k++) {
        num[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        const int item = tid + k * NTHREADS, i4 = item % NI4, a = item / NI4;
        if (item >= ITEMS || a >= r || 4 * i4 >= nrow) continue;
        // rows up to round_up(n, 4) exist in every slice (i0 and ldp are multiples of 4)
        const float4* src = reinterpret_cast<const float4*>(P + (size_t)a * ldp + i0 + 4 * i4);
        const long ps4 = pstride >> 2;
        int s = 0;
        for (; s + 4 <= S; s += 4) {
            const float4 x0 = src[(size_t)s * ps4], x1 = src[(size_t)(s + 1) * ps4];
            const float4 x2 = src[(size_t)(s + 2) * ps4], x3 = src[(size_t)(s + 3) * ps4];
            num[k].x += x0.x, num[k].y += x0.y, num[k].z += x0.z, num[k].w += x0.w;
            num[k].x += x1.x, num[k].y += x1.y, num[k].z += x1.z, num[k].w += x1.w;
            num[k].x += x2.x, num[k].y += x2.y, num[k].z += x2.z, num[k].w += x2.w;
            num[k].x += x3.x, num[k].y += x3.y, num[k].z += x3.z, num[k].w += x3.w;
        }
        for (; s < S; s++) {
            const float4 x0 = src[(size_t)s * ps4];
            num[k].x += x0.x, num[k].y += x0.y, num[k].z += x0.z, num[k].w += x0.w;
        }
    }
    if constexpr (EARLY) {
        stg_g.store(RP * RP, st_g);
        stg_w.store(RP * JT, st_w);
    } else {
        stage_lds<8>(RP * RP, ld_g, st_g);
        stage_lds<8>(RP * JT, ld_w, st_w);
    }
    __syncthreads();
    float4 out[MAXI];
#pragma unroll
    for (int k = 0; k < MAXI; k++) {
        const int item = tid + k * NTHREADS, i4 = item % NI4, a = item / NI4;
        out[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (item >= ITEMS || a >= r || 4 * i4 >= nrow) continue;
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 16
        for (int b = 0; b < RP; b++) {
            const float4 w4 = *reinterpret_cast<const float4*>(sWt + b * JS + 4 * i4);
            const float g = sG[b * RS + a];
            d.x += w4.x * g, d.y += w4.y * g, d.z += w4.z * g, d.w += w4.w * g;
        }
        const float4 wo = *reinterpret_cast<const float4*>(sWt + a * JS + 4 * i4);
        if (add) {
            out[k].x = additive_rule(wo.x, num[k].x, d.x);
            out[k].y = additive_rule(wo.y, num[k].y, d.y);
            out[k].z = additive_rule(wo.z, num[k].z, d.z);
            out[k].w = additive_rule(wo.w, num[k].w, d.w);
        } else {
            out[k].x = wo.x * (num[k].x / (d.x + eps));
            out[k].y = wo.y * (num[k].y / (d.y + eps));
            out[k].z = wo.z * (num[k].z / (d.z + eps));
            out[k].w = wo.w * (num[k].w / (d.w + eps));
        }
    }
#pragma unroll
    for (int k = 0; k < MAXI; k++) {
        const int item = tid + k * NTHREADS, i4 = item % NI4, a = item / NI4;
        if (item >= ITEMS || a >= r) continue;
        const int i = i0 + 4 * i4;
        float* dst = W + (size_t)a * n + i;
        if (4 * i4 < nrow) dst[0] = out[k].x;
        if (4 * i4 + 1 < nrow) dst[1] = out[k].y;
        if (4 * i4 + 2 < nrow) dst[2] = out[k].z;
        if (4 * i4 + 3 < nrow) dst[3] = out[k].w;
    }
}

// KL H update, reference nmf_kernels.cu:337-339 with the clamped sums of :104.
// add != 0: the additive rule of nmf_kernels_additive_divergence.cu:178-190,
//   H <- max(0, H + (H / sum) * (num - rawsum)),  num - rawsum = sum_i W(i,a) (ratio - 1).
// ssum holds [clamped sums | raw sums] (2r floats).
__device__ __forceinline__ void update_h_kl_body(float* __restrict__ H, const float* __restrict__ P,
                                                 const float* __restrict__ Psum, int m, int r, int ldr, int S,
                                                 long pstride, float eps, int blk, float* ssum, int add) {
    const int tid = threadIdx.x;
    const int rpad = round_up(r, 4);
    for (int a = tid; a < r; a += NTHREADS) {
        float v = 0.f;
        for (int s = 0; s < S; s++) v += Psum[(size_t)s * rpad + a];
        ssum[a] = (v < eps) ? eps : v;
        ssum[r + a] = v;
    }
    __syncthreads();
    const int u0 = blk * UPD_COLS;
    const int ncol = min(UPD_COLS, m - u0);
    for (int idx = tid; idx < ncol * r; idx += NTHREADS) {
        const int uu = idx / r, a = idx - uu * r;
        const int u = u0 + uu;
        const float num = sum_slices(P + (size_t)u * ldr + a, S, pstride);
        float* h = H + (size_t)u * r + a;
        if (add)
            *h = additive_kl_rule(*h, ssum[a], num - ssum[r + a]);
        else
            *h *= (num / ssum[a]);
    }
}

__global__ __launch_bounds__(NTHREADS) void k_update_h_kl(float* __restrict__ H, const float* __restrict__ P,
                                                          const float* __restrict__ Psum, int m, int r, int ldr,
                                                          int S, long pstride, float eps, int add) {
    __shared__ float ssum[1024];
    update_h_kl_body(H, P, Psum, m, r, ldr, S, pstride, eps, blockIdx.x, ssum, add);
}

// KL W update, reference nmf_kernels.cu:289-291 with the clamped sums of :190.
// add != 0: the additive rule of nmf_kernels_additive_divergence.cu:109-121.
__device__ __forceinline__ void update_w_kl_body(float* __restrict__ W, const float* __restrict__ P,
                                                 const float* __restrict__ Psum, int n, int r, int ldp, int S,
                                                 long pstride, float eps, int blk, float* ssum, int add) {
    const int tid = threadIdx.x;
    const int rpad = round_up(r, 4);
    for (int a = tid; a < r; a += NTHREADS) {
        float v = 0.f;
        for (int s = 0; s < S; s++) v += Psum[(size_t)s * rpad + a];
        ssum[a] = (v < eps) ? eps : v;
        ssum[r + a] = v;
    }
    __syncthreads();
    const int i0 = blk * UPD_ROWS;
    const int nrow = min(UPD_ROWS, n - i0);
    for (int idx = tid; idx < UPD_ROWS * r; idx += NTHREADS) {
        const int ii = idx % UPD_ROWS, a = idx / UPD_ROWS;
        if (ii >= nrow) continue;
        const int i = i0 + ii;
        const float num = sum_slices(P + (size_t)a * ldp + i, S, pstride);
        float* w = W + (size_t)a * n + i;
        if (add)
            *w = additive_kl_rule(*w, ssum[a], num - ssum[r + a]);
        else
            *w *= (num / ssum[a]);
    }
}

__global__ __launch_bounds__(NTHREADS) void k_update_w_kl(float* __restrict__ W, const float* __restrict__ P,
                                                          const float* __restrict__ Psum, int n, int r, int ldp,
                                                          int S, long pstride, float eps, int add) {
    __shared__ float ssum[1024];
    update_w_kl_body(W, P, Psum, n, r, ldp, S, pstride, eps, blockIdx.x, ssum, add);
}

__global__ void k_zero_i32(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t launch_zero_i32(int* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

// Deterministic reduction of count doubles (fixed strided order, then a fixed tree).
__global__ __launch_bounds__(NTHREADS) void k_reduce_f64(const double* __restrict__ in, int count, double* out) {
    __shared__ double red[NTHREADS / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += NTHREADS) v += in[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);  // wave64 shuffle tree
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < NTHREADS / 64; w++) t += red[w];
        out[0] = t;
    }
}

// The same reduction, the total written as a float (the reference quality's type).
__global__ __launch_bounds__(NTHREADS) void k_reduce_f64_to_f32(const double* __restrict__ in, int count, float* out) {
    __shared__ double red[NTHREADS / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += NTHREADS) v += in[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < NTHREADS / 64; w++) t += red[w];
        out[0] = (float)t;
    }
}

// ---------------------------------------------------------------------------
// Reference-shaped kernels over materialised operands.

// nmf_kernels.cu:43-47
__global__ void k_mu_update(const float* __restrict__ nm, const float* __restrict__ dm, float* __restrict__ mm,
                            long elements, float eps) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < elements; k += (long)gridDim.x * blockDim.x)
        mm[k] *= nm[k] / (dm[k] + eps);
}

// nmf_kernels.cu:31-39
__global__ void k_add_update(float* __restrict__ X, const float* __restrict__ d1, const float* __restrict__ d2,
                             long elements) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < elements; k += (long)gridDim.x * blockDim.x)
        X[k] = additive_rule(X[k], d1[k], d2[k]);
}

// Column sums of W (one workgroup per column a), clamped at eps: nmf_kernels.cu:63-109.
__global__ __launch_bounds__(NTHREADS) void k_sum_w(const float* __restrict__ W, int n, float eps, float* out) {
    __shared__ float red[NTHREADS];
    const int a = blockIdx.x;
    float v = 0.f;
    for (int i = threadIdx.x; i < n; i += NTHREADS) v += W[(size_t)a * n + i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = NTHREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[a] = (red[0] < eps) ? eps : red[0];
}

// Row sums of H (r x m, column-major), clamped at eps: nmf_kernels.cu:149-195.
// One workgroup per block of up to 64 rank rows, any r: thread t sums row a0 + t % rb
// over the columns u = g, g + groups, ... (g = t / rb), so a wave reads rb consecutive
// floats of a column; the column groups are then added in group order.
constexpr int kSumHThreads = 1024;
__global__ __launch_bounds__(kSumHThreads) void k_sum_h(const float* __restrict__ H, int r, int m, float eps,
                                                        float* out) {
    __shared__ float red[kSumHThreads];
    const int a0 = blockIdx.x * 64;
    const int rb = min(64, r - a0);
    const int groups = kSumHThreads / rb;
    const int al = threadIdx.x % rb, g = threadIdx.x / rb;
    float v = 0.f;
    if (g < groups) {
#pragma unroll 8
        for (int u = g; u < m; u += groups) v += H[(size_t)u * r + a0 + al];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < rb) {
        float s = 0.f;
        for (int j = 0; j < groups; j++) s += red[j * rb + threadIdx.x];
        out[a0 + threadIdx.x] = (s < eps) ? eps : s;
    }
}

// Reference quality sum_k V[k] log10(WH[k] + eps) - WH[k] (nmf_kernels_quality.cu:29-71),
// accumulated in double over many workgroups.
__global__ __launch_bounds__(NTHREADS) void k_quality(const float* __restrict__ V, const float* __restrict__ WH,
                                                      long count, float eps, double* partials) {
    __shared__ double red[NTHREADS / 64];
    double acc = 0.0;
    for (long k = (long)blockIdx.x * NTHREADS + threadIdx.x; k < count; k += (long)gridDim.x * NTHREADS) {
        const float wh = WH[k];
        acc += (double)V[k] * log10((double)(wh + eps)) - (double)wh;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < NTHREADS / 64; i++) t += red[i];
        partials[blockIdx.x] = t;
    }
}

// KL H update with a materialised WH: the reference kernel contract of
// nmf_kernels.cu:295-341.  One workgroup per column u of H.
__global__ __launch_bounds__(NTHREADS) void k_kl_update_h_wh(float* __restrict__ H, const float* __restrict__ W,
                                                             const float* __restrict__ V, const float* __restrict__ WH,
                                                             const float* __restrict__ sumW, int n, int m, int r,
                                                             float eps, int add) {
    __shared__ float sR[NTHREADS];
    const int u = blockIdx.x, tid = threadIdx.x;
    for (int a0 = 0; a0 < r; a0 += NTHREADS) {
        const int a = a0 + tid;
        float acc = 0.f;
        for (int i0 = 0; i0 < n; i0 += NTHREADS) {
            const int i = i0 + tid;
            // additive form (nmf_kernels_additive_divergence.cu:160-169): the term is ratio - 1
            sR[tid] = (i < n) ? V[(size_t)u * n + i] / (WH[(size_t)u * n + i] + eps) - (add ? 1.f : 0.f) : 0.f;
            __syncthreads();
            if (a < r) {
                const int lim = min(NTHREADS, n - i0);
                for (int j = 0; j < lim; j++) acc += W[(size_t)a * n + i0 + j] * sR[j];
            }
            __syncthreads();
        }
        if (a < r) {
            float* h = H + (size_t)u * r + a;
            if (add)
                *h = additive_kl_rule(*h, sumW[a], acc);
            else
                *h *= (acc / sumW[a]);
        }
    }
}

// KL W update with a materialised WH: the reference kernel contract of
// nmf_kernels.cu:247-293.  One thread per row i, one rank column at a time.
__global__ __launch_bounds__(NTHREADS) void k_kl_update_w_wh(float* __restrict__ W, const float* __restrict__ H,
                                                             const float* __restrict__ V, const float* __restrict__ WH,
                                                             const float* __restrict__ sumH, int n, int m, int r,
                                                             float eps, int add) {
    const int i = blockIdx.x * NTHREADS + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < r; a++) {
        float acc = 0.f;
        for (int u = 0; u < m; u++) {
            const size_t k = (size_t)u * n + i;
            acc += H[(size_t)u * r + a] * (V[k] / (WH[k] + eps) - (add ? 1.f : 0.f));
        }
        float* w = W + (size_t)a * n + i;
        if (add)
            *w = additive_kl_rule(*w, sumH[a], acc);
        else
            *w *= (acc / sumH[a]);
    }
}

// ---------------------------------------------------------------------------
// Grouped updates (r <= 64 per problem) and per-problem objective reductions.
__device__ __forceinline__ int group_find_u(const int* first, int J, int item) {
    int lo = 0, hi = J - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

constexpr int kGroupSmemFloats = (64 + UPD_ROWS * 2) * 65;

__global__ __launch_bounds__(NTHREADS) void k_group_update_h(const JobArgs* __restrict__ jobs, const int* table) {
    __shared__ float sm[kGroupSmemFloats];
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find_u(first, J, item);
        const JobArgs& ja = jobs[idx[j]];
        const PassArgs& p = ja.h;
        __syncthreads();
        if (ja.rule == 0)
            update_h_euc_body(ja.H, p.P, p.G, p.m, p.r, p.ldr, p.S, p.pstride, p.eps, item - first[j], sm);
        else
            update_h_kl_body(ja.H, p.P, p.Psum, p.m, p.r, p.ldr, p.S, p.pstride, p.eps, item - first[j], sm, 0);
    }
}

__global__ __launch_bounds__(NTHREADS) void k_group_update_w(const JobArgs* __restrict__ jobs, const int* table) {
    __shared__ float sm[kGroupSmemFloats];
    const int J = table[0], T = table[1];
    const int* first = table + 2;
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int item = blockIdx.x; item < T; item += gridDim.x) {
        const int j = group_find_u(first, J, item);
        const JobArgs& ja = jobs[idx[j]];
        const PassArgs& p = ja.w;
        __syncthreads();
        if (ja.rule == 0)
            update_w_euc_body(ja.W, p.P, p.G, p.n, p.r, p.ldp, p.S, p.pstride, p.eps, item - first[j], sm);
        else
            update_w_kl_body(ja.W, p.P, p.Psum, p.n, p.r, p.ldp, p.S, p.pstride, p.eps, item - first[j], sm, 0);
    }
}

// One workgroup per problem: obj[slot] = sum of its H-pass objective partials (fixed order).
__global__ __launch_bounds__(NTHREADS) void k_group_reduce_obj(const JobArgs* __restrict__ jobs, const int* table,
                                                               int slot) {
    __shared__ double red[NTHREADS];
    const int J = table[0];
    const int* idx = table + 3 + kGroupMaxJobs;
    for (int j = blockIdx.x; j < J; j += gridDim.x) {
        const JobArgs& ja = jobs[idx[j]];
        double v = 0.0;
        for (int i = threadIdx.x; i < ja.nobj; i += NTHREADS) v += ja.h.objp[i];
        __syncthreads();
        red[threadIdx.x] = v;
        __syncthreads();
        for (int off = NTHREADS / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) ja.obj[slot] = red[0];
    }
}

hipError_t launch_group_update_h(const JobArgs* jobs, const int* table, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_group_update_h, dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    return hipGetLastError();
}

hipError_t launch_group_update_w(const JobArgs* jobs, const int* table, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_group_update_w, dim3(grid), dim3(NTHREADS), 0, st, jobs, table);
    return hipGetLastError();
}

hipError_t launch_group_reduce_obj(const JobArgs* jobs, const int* table, int slot, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_group_reduce_obj, dim3(grid), dim3(NTHREADS), 0, st, jobs, table, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
static size_t upd_smem(int r, int rows) { return (size_t)(r * (r + 1) + rows * (r + 1)) * sizeof(float); }

template <typename K>
static hipError_t set_smem(K kernel, size_t sm) {
    if (sm <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
}

// Row-form update: JT = 32 rows per workgroup for long counts (about 512 or more
// workgroups), else 16; RP = r rounded up to 32.
template <int JT, int RP>
static hipError_t launch_rows_t(float* X, long xs_j, long xs_a, const float* P, const float* G, int ldg, int gt,
                                int count, int r, int ldr, int S, long pstride, float eps, uint16_t* Sh, int ldsh,
                                uint8_t* Q, int ldq, uint8_t* Qsc, int nblk, int add, long gps, int gsl,
                                long shplane, hipStream_t st) {
    const size_t sm = (size_t)(RP + 2 * JT) * (RP + 4) * sizeof(float);
    hipError_t e = set_smem(k_update_rows_euc<JT, RP>, sm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_update_rows_euc<JT, RP>), dim3((count + JT - 1) / JT), dim3(NTHREADS), sm, st, X, xs_j, xs_a,
                       P, G, ldg, gt, count, r, ldr, S, pstride, eps, Sh, ldsh, Q, ldq, Qsc, nblk, add, gps, gsl,
                       shplane);
    return hipGetLastError();
}

// Q (FP8 shadow) needs 32 rows per workgroup (one scale block); it is given for r <= 64.
// gsl > 0 (gram slices summed by the kernel) needs the register-staged gram: r <= 64.
static hipError_t launch_rows(float* X, long xs_j, long xs_a, const float* P, const float* G, int ldg, int gt,
                              int count, int r, int ldr, int S, long pstride, float eps, uint16_t* Sh, int ldsh,
                              uint8_t* Q, int ldq, uint8_t* Qsc, int nblk, int add, long gps, int gsl,
                              hipStream_t st, long shplane = 0) {
    if (r < 1 || r > 128 || (Q && (r > 64 || ldq % 32 || ldq < count))) return hipErrorInvalidValue;
    if (shplane && (!Sh || shplane % 2 || ldsh % 128 || ldsh < count || shplane < (long)r * ldsh))
        return hipErrorInvalidValue;
    if (gsl < 0 || gsl > kMaxGramSlices || (gsl > 0 && r > 64)) return hipErrorInvalidValue;
    const bool wide = count >= 16384 || Q != nullptr;
#define LR(JT, RP)                                                                                                \
    return launch_rows_t<JT, RP>(X, xs_j, xs_a, P, G, ldg, gt, count, r, ldr, S, pstride, eps, Sh, ldsh, Q, ldq, \
                                 Qsc, nblk, add, gps, gsl, shplane, st)
    if (r <= 32) { if (wide) LR(32, 32); LR(16, 32); }
    if (r <= 64) { if (wide) LR(32, 64); LR(16, 64); }
    LR(16, 128);
#undef LR
}

hipError_t launch_update_h_euc(float* H, const float* P, const float* G, int m, int r, int ldr, int S, long pstride,
                               float eps, uint16_t* Hs, int ldhs, hipStream_t st, int add, int gsl, long shplane) {
    return launch_rows(H, r, 1, P, G, ldr, 0, m, r, ldr, S, pstride, eps, Hs, ldhs, nullptr, 0, nullptr, 0, add,
                       pstride, gsl, st, shplane);
}

hipError_t launch_update_w_euc_t(float* W, const float* P, const float* G, int n, int r, int ldr, int S, long pstride,
                                 float eps, uint16_t* Wb, int ldwb, hipStream_t st, int add, int gsl, long shplane) {
    return launch_rows(W, 1, n, P, G, ldr, 1, n, r, ldr, S, pstride, eps, Wb, ldwb, nullptr, 0, nullptr, 0, add,
                       pstride, gsl, st, shplane);
}

hipError_t launch_update_h_euc_q8(float* H, const float* P, const float* G, int m, int r, int ldr, int S,
                                  long pstride, float eps, uint8_t* Hq, int ldq, uint8_t* Hsc, hipStream_t st,
                                  int gsl) {
    return launch_rows(H, r, 1, P, G, ldr, 0, m, r, ldr, S, pstride, eps, nullptr, 0, Hq, ldq, Hsc, ldq / 32, 0,
                       pstride, gsl, st);
}

hipError_t launch_update_w_euc_t_q8(float* W, const float* P, const float* G, int n, int r, int ldr, int S,
                                    long pstride, float eps, uint8_t* Wq, int ldq, uint8_t* Wsc, hipStream_t st,
                                    int gsl) {
    return launch_rows(W, 1, n, P, G, ldr, 1, n, r, ldr, S, pstride, eps, nullptr, 0, Wq, ldq, Wsc, ldq / 32, 0,
                       pstride, gsl, st);
}

template <int JT, int RP>
static hipError_t launch_cols_t(float* W, const float* P, const float* G, int n, int r, int ldp, int S, long pstride,
                                float eps, int add, hipStream_t st) {
    const size_t sm = (size_t)(RP * (RP + 4) + RP * (JT + 4)) * sizeof(float);
    hipError_t e = set_smem(k_update_w_euc_cols<JT, RP>, sm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_update_w_euc_cols<JT, RP>), dim3((n + JT - 1) / JT), dim3(NTHREADS), sm, st, W, P, G, n, r,
                       ldp, S, pstride, eps, add);
    return hipGetLastError();
}

hipError_t launch_update_w_euc(float* W, const float* P, const float* G, int n, int r, int ldp, int S, long pstride,
                               float eps, hipStream_t st, int add) {
    if (r < 1 || r > 128) return hipErrorInvalidValue;
    const bool wide = n >= 16384;
    if (r <= 32) return wide ? launch_cols_t<32, 32>(W, P, G, n, r, ldp, S, pstride, eps, add, st)
                             : launch_cols_t<16, 32>(W, P, G, n, r, ldp, S, pstride, eps, add, st);
    if (r <= 64) return wide ? launch_cols_t<32, 64>(W, P, G, n, r, ldp, S, pstride, eps, add, st)
                             : launch_cols_t<16, 64>(W, P, G, n, r, ldp, S, pstride, eps, add, st);
    return launch_cols_t<16, 128>(W, P, G, n, r, ldp, S, pstride, eps, add, st);
}

// ---------------------------------------------------------------------------
// Single-problem KL updates (nmf_kernels.cu:289-291, 337-339; additive form
// nmf_kernels_additive_divergence.cu:109-121, 178-190) over the split-K partials.  Each
// thread owns 16-byte groups of partials and issues the loads of KLB slices of all its
// groups (and of the clamped sums) before it adds any of them, so a workgroup pays about
// S / KLB memory round trips instead of one per element; slices are added in slice order,
// as the per-element loop did (bit-identical results).
constexpr int KLB = 8;

// Threads t < rpad / 4 also sum the S slices of Psum (stride rpad) for rank rows 4t .. 4t+3:
// the clamped sums max(sum, eps) and the raw sums (additive form) go to LDS.
// H update, row form: thread -> (column u = u0 + t / rq, rank rows 4 (t % rq) .. +3),
// partials P[s][u*ldr + a]; JT = NTHREADS / rq columns per workgroup.
__global__ __launch_bounds__(NTHREADS) void k_update_h_kl_v(float* __restrict__ H, const float* __restrict__ P,
                                                            const float* __restrict__ Psum, int m, int r, int ldr,
                                                            int S, long pstride, float eps, int add) {
    __shared__ float ssum[2 * 128];
    const int tid = threadIdx.x, rq = ldr >> 2, jt = NTHREADS / rq;
    const int rpad = round_up(r, 4);
    const int u = blockIdx.x * jt + tid / rq, q = tid % rq;
    const bool item = (tid / rq) < jt && u < m && 4 * q < r;
    const bool sumt = tid < rq;
    const float4* P4 = reinterpret_cast<const float4*>(P) + (size_t)u * rq + q;
    const float4* S4 = reinterpret_cast<const float4*>(Psum) + tid;
    const long ps4 = pstride >> 2, ss4 = rpad >> 2;
    float4 num = make_float4(0.f, 0.f, 0.f, 0.f), sm = num;
    for (int s0 = 0; s0 < S; s0 += KLB) {
        float4 x[KLB], y[KLB];
#pragma unroll
        for (int b = 0; b < KLB; b++) {
            const bool ok = s0 + b < S;
            x[b] = (item && ok) ? P4[(size_t)(s0 + b) * ps4] : make_float4(0.f, 0.f, 0.f, 0.f);
            y[b] = (sumt && ok) ? S4[(size_t)(s0 + b) * ss4] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int b = 0; b < KLB; b++) {
            if (s0 + b >= S) break;
            num.x += x[b].x, num.y += x[b].y, num.z += x[b].z, num.w += x[b].w;
            sm.x += y[b].x, sm.y += y[b].y, sm.z += y[b].z, sm.w += y[b].w;
        }
    }
    if (sumt) {
        const float v[4] = {sm.x, sm.y, sm.z, sm.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int a = 4 * tid + j;
            if (a < r) {
                ssum[a] = (v[j] < eps) ? eps : v[j];
                ssum[128 + a] = v[j];
            }
        }
    }
    __syncthreads();
    if (!item) return;
    const float nv[4] = {num.x, num.y, num.z, num.w};
    float* h = H + (size_t)u * r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int a = 4 * q + j;
        if (a >= r) break;
        if (add)
            h[a] = additive_kl_rule(h[a], ssum[a], nv[j] - ssum[128 + a]);
        else
            h[a] *= (nv[j] / ssum[a]);
    }
}

// W update, column form: partials P[s][a*ldp + i] (i contiguous); thread item -> (rank
// column a, rows 4 g .. 4 g + 3 of the workgroup's JT rows); up to 2 items per thread.
__global__ __launch_bounds__(NTHREADS) void k_update_w_kl_v(float* __restrict__ W, const float* __restrict__ P,
                                                            const float* __restrict__ Psum, int n, int r, int ldp,
                                                            int S, long pstride, float eps, int add, int jt,
                                                            int vec) {
    __shared__ float ssum[2 * 128];
    const int tid = threadIdx.x, rpad = round_up(r, 4), rq = rpad >> 2;
    const int gq = jt >> 2;  // row quads per workgroup
    const int i0 = blockIdx.x * jt;
    const bool sumt = tid < rq;
    const long ss4 = rpad >> 2;
    const float4* S4 = reinterpret_cast<const float4*>(Psum) + tid;
    int ia[2], aa[2];
    bool it[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int idx = tid + k * NTHREADS;
        aa[k] = idx / gq;
        ia[k] = i0 + 4 * (idx % gq);
        it[k] = aa[k] < r && ia[k] < n;
    }
    float4 num[2] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
    float4 sm = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s0 = 0; s0 < S; s0 += KLB) {
        float4 x[2][KLB], y[KLB];
#pragma unroll
        for (int b = 0; b < KLB; b++) {
            const bool ok = s0 + b < S;
#pragma unroll
            for (int k = 0; k < 2; k++)
                x[k][b] = (it[k] && ok) ? *reinterpret_cast<const float4*>(P + (size_t)(s0 + b) * pstride +
                                                                            (size_t)aa[k] * ldp + ia[k])
                                        : make_float4(0.f, 0.f, 0.f, 0.f);
            y[b] = (sumt && ok) ? S4[(size_t)(s0 + b) * ss4] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int b = 0; b < KLB; b++) {
            if (s0 + b >= S) break;
#pragma unroll
            for (int k = 0; k < 2; k++)
                num[k].x += x[k][b].x, num[k].y += x[k][b].y, num[k].z += x[k][b].z, num[k].w += x[k][b].w;
            sm.x += y[b].x, sm.y += y[b].y, sm.z += y[b].z, sm.w += y[b].w;
        }
    }
    if (sumt) {
        const float v[4] = {sm.x, sm.y, sm.z, sm.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int a = 4 * tid + j;
            if (a < r) {
                ssum[a] = (v[j] < eps) ? eps : v[j];
                ssum[128 + a] = v[j];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; k++) {
        if (!it[k]) continue;
        const int a = aa[k];
        float* w = W + (size_t)a * n + ia[k];
        const float nv[4] = {num[k].x, num[k].y, num[k].z, num[k].w};
        const float cs = ssum[a], raw = ssum[128 + a];
        if (vec) {
            float4 o = *reinterpret_cast<float4*>(w);
            float ov[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
            for (int j = 0; j < 4; j++) ov[j] = add ? additive_kl_rule(ov[j], cs, nv[j] - raw) : ov[j] * (nv[j] / cs);
            *reinterpret_cast<float4*>(w) = make_float4(ov[0], ov[1], ov[2], ov[3]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (ia[k] + j >= n) break;
                w[j] = add ? additive_kl_rule(w[j], cs, nv[j] - raw) : w[j] * (nv[j] / cs);
            }
        }
    }
}

hipError_t launch_update_h_kl(float* H, const float* P, const float* Psum, int m, int r, int ldr, int S, long pstride,
                              float eps, hipStream_t st, int add) {
    if (r > 128 || ldr != round_up(r, 4)) return hipErrorInvalidValue;
    const int jt = NTHREADS / (ldr >> 2);
    hipLaunchKernelGGL(k_update_h_kl_v, dim3((m + jt - 1) / jt), dim3(NTHREADS), 0, st, H, P, Psum, m, r, ldr, S,
                       pstride, eps, add);
    return hipGetLastError();
}

hipError_t launch_update_w_kl(float* W, const float* P, const float* Psum, int n, int r, int ldp, int S, long pstride,
                              float eps, hipStream_t st, int add) {
    if (r > 128) return hipErrorInvalidValue;
    // rows per workgroup: (jt / 4) row quads x r rank columns <= 2 items per thread
    const int rpad = round_up(r, 4);
    int jt = 4 * ((2 * NTHREADS) / rpad);
    if (jt > 256) jt = 256;
    const int vec = (n % 4 == 0) && (((uintptr_t)W & 15) == 0);
    hipLaunchKernelGGL(k_update_w_kl_v, dim3((n + jt - 1) / jt), dim3(NTHREADS), 0, st, W, P, Psum, n, r, ldp, S,
                       pstride, eps, add, jt, vec);
    return hipGetLastError();
}

hipError_t launch_reduce_f64(const double* in, int count, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_f64, dim3(1), dim3(NTHREADS), 0, st, in, count, out);
    return hipGetLastError();
}

hipError_t launch_mu_update(const float* nm, const float* dm, float* mm, long elements, float eps, hipStream_t st) {
    if (elements <= 0) return hipSuccess;
    long blocks = (elements + NTHREADS - 1) / NTHREADS;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_mu_update, dim3((unsigned)blocks), dim3(NTHREADS), 0, st, nm, dm, mm, elements, eps);
    return hipGetLastError();
}

hipError_t launch_add_update(float* X, const float* d1, const float* d2, long elements, hipStream_t st) {
    if (elements <= 0) return hipSuccess;
    long blocks = (elements + NTHREADS - 1) / NTHREADS;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_add_update, dim3((unsigned)blocks), dim3(NTHREADS), 0, st, X, d1, d2, elements);
    return hipGetLastError();
}

hipError_t launch_sum_w(const float* W, int n, int r, float eps, float* out, hipStream_t st) {
    if (r <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_w, dim3(r), dim3(NTHREADS), 0, st, W, n, eps, out);
    return hipGetLastError();
}

hipError_t launch_sum_h(const float* H, int r, int m, float eps, float* out, hipStream_t st) {
    if (r <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_h, dim3((r + 63) / 64), dim3(kSumHThreads), 0, st, H, r, m, eps, out);
    return hipGetLastError();
}

hipError_t launch_quality(const float* V, const float* WH, long count, float eps, double* partials, int npartials,
                          double* out, hipStream_t st, float* out_f32) {
    hipLaunchKernelGGL(k_quality, dim3(npartials), dim3(NTHREADS), 0, st, V, WH, count, eps, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (out_f32) {
        hipLaunchKernelGGL(k_reduce_f64_to_f32, dim3(1), dim3(NTHREADS), 0, st, partials, npartials, out_f32);
        return hipGetLastError();
    }
    return launch_reduce_f64(partials, npartials, out, st);
}

hipError_t launch_kl_update_h_wh(float* H, const float* W, const float* V, const float* WH, const float* sumW, int n,
                                 int m, int r, float eps, hipStream_t st, int add) {
    if (m <= 0 || r <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kl_update_h_wh, dim3(m), dim3(NTHREADS), 0, st, H, W, V, WH, sumW, n, m, r, eps, add);
    return hipGetLastError();
}

hipError_t launch_kl_update_w_wh(float* W, const float* H, const float* V, const float* WH, const float* sumH, int n,
                                 int m, int r, float eps, hipStream_t st, int add) {
    if (n <= 0 || r <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kl_update_w_wh, dim3((n + NTHREADS - 1) / NTHREADS), dim3(NTHREADS), 0, st, W, H, V, WH, sumH,
                       n, m, r, eps, add);
    return hipGetLastError();
}

}  // namespace gpumlib_nmf
