// Shared definitions for the gfx950 NMF kernels.
//
// Storage is column-major, as in GPUMLib (src/algorithms/nmf/nmf_base.h:109-111):
//   V is n x m, V(i,u) = V[u*n + i];  W is n x r, W(i,a) = W[a*n + i];
//   H is r x m, H(a,u) = H[u*r + a].
// The two long contractions of one multiplicative-update iteration are
//   "H pass": reduce over the rows i of V   -> r x m   (W^T V, W^T W, KL numerator of H)
//   "W pass": reduce over the columns u of V -> n x r  (V H^T, H H^T, KL numerator of W)
// Each is split over K into S slices written as FP32 partials; a small update
// kernel ("epilogue") sums the S partials in slice order (deterministic) and
// applies the multiplicative rule.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpumlib_nmf {


// K depth staged through LDS per pipeline step, and the LDS row pitch used for
// K-contiguous tiles (4 floats of padding keep 16-byte alignment for 128-bit LDS
// reads and spread 8-lane groups over distinct banks).
constexpr int KC = 32;
constexpr int LDK = KC + 4;
// Output columns (H pass) or rows (W pass) of V handled by one workgroup.
constexpr int TILE = 128;
constexpr int NTHREADS = 256;
// Columns of H / rows of W handled by one update workgroup.
constexpr int TILE_LP = 64;   // columns per workgroup of the BF16 / FP8 passes
constexpr int UPD_COLS = 32;
constexpr int UPD_ROWS = 32;

// Rank padding: the fused kernels hold r rounded up to 32-row MFMA blocks.
__host__ __device__ inline int round_up(int x, int q) { return (x + q - 1) / q * q; }
inline int xcd_grid(int work) { return (work + 7) / 8 * 8; }

// Element (k, col) of a tiled BF16 V store orientation with ntiles column tiles: blocks of
// 64 columns x 256 contraction elements, block (k / 256, col / 64) at k / 256 * ntiles +
// col / 64, column-major inside the block.
__host__ __device__ inline size_t bf16_tiled_index(int k, int col, int ntiles) {
    return (((size_t)(k >> 8) * ntiles + (col >> 6)) * 64 + (col & 63)) * 256 + (k & 255);
}

// The same for the FP8 V store: blocks of 64 columns x 512 contraction elements (32 KB).
__host__ __device__ inline size_t fp8_tiled_index(int k, int col, int ntiles) {
    return (((size_t)(k >> 9) * ntiles + (col >> 6)) * 64 + (col & 63)) * 512 + (k & 511);
}

// The same for the FP32 V store of the split-BF16 path: blocks of 64 columns x 128
// contraction elements (32 KB).
__host__ __device__ inline size_t x3_tiled_index(int k, int col, int ntiles) {
    return (((size_t)(k >> 7) * ntiles + (col >> 6)) * 64 + (col & 63)) * 128 + (k & 127);
}

// Objective selector of the objective-carrying passes and kernels.
enum ObjKind { OBJ_NONE = 0, OBJ_FROB = 1, OBJ_KL = 2, OBJ_QUALITY = 3 };

#if defined(__HIPCC__)
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// One 32x32x2 FP32 MFMA.  Operand lane layout (lane l, c = l % 32, h = l / 32):
//   A[row c][k h], B[k h][col c];  result register v holds D[row 8*(v/4) + 4*h + v%4][col c].
__device__ __forceinline__ f32x16 mfma32x2(float a, float b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
}

// Load of the streamed operand of a pass (V) with the non-temporal cache policy (nt), for V
// too large to stay cached from one iteration to the next: the stream then does not displace
// the L2-resident factor chunks every workgroup re-reads.  A read probe of the BF16 pass's
// pattern (one contiguous 64 KB chunk per workgroup and step, 256 workgroups of 512 threads,
// staged through LDS) over 537 MB on the MI355X: 6.7 TB/s with the nt policy, 6.1 TB/s
// without.  V that fits the 256 MB last-level cache is read with the default policy (with nt
// its reuse across iterations is lost: the KL pass at 4096 x 4096 ran 7% slower).
template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    return __builtin_nontemporal_load(p);
}

// KL ratio V / (WH + eps) of the fused passes (nmf_kernels.cu:271,274,319,322), as V times
// the hardware reciprocal of WH + eps (v_rcp_f32, within 1 ulp): the ratio only feeds a
// sum over the contraction, where the <= 1.5 ulp difference per term from a correctly
// rounded division is far below the FP32 accumulation error (objective after 300 iterations
// at 4096 x 4096 r = 32: 2.8e-8 relative change), and it saves the eight-instruction
// division sequence per element.  The reference-shaped launchers keep the IEEE division.
__device__ __forceinline__ float kl_ratio(float v, float wh, float eps) {
    return v * __builtin_amdgcn_rcpf(wh + eps);
}

// Additive Euclidean element rule (nmf_kernels.cu:31-39): x + (x / d2) * (d1 - d2),
// set to 0 when negative (a NaN is kept: `v < 0` is false for it).  No eps.
__device__ __forceinline__ float additive_rule(float x, float d1, float d2) {
    const float v = x + (x / d2) * (d1 - d2);
    return v < 0.f ? 0.f : v;
}

// Additive KL rule (nmf_kernels_additive_divergence.cu:109-121, 178-190):
// x + (x / sum) * t, t = sum_k F(k) (ratio(k) - 1), set to 0 when negative.
__device__ __forceinline__ float additive_kl_rule(float x, float sum, float t) {
    const float v = x + (x / sum) * t;
    return v < 0.f ? 0.f : v;
}

// Per-element objective terms, returned in FP64 for an FP64 sum (ObjKind): Frobenius
// (v - wh)^2 and KL v ln(v / wh) - v + wh (0 ln 0 = 0), both formed in FP64; the reference
// quality v log10(wh + eps) - wh formed in FP32 as the reference forms it (its value is
// reported as a float).
template <int OBJ>
__device__ __forceinline__ double obj_term(float v, float wh, float eps) {
    if (OBJ == OBJ_FROB) {
        const double d = (double)v - (double)wh;
        return d * d;
    } else if (OBJ == OBJ_KL) {
        const double dv = v, dwh = wh;
        const double t = (v > 0.f) ? dv * log(dv / dwh) : 0.0;
        return t - dv + dwh;
    } else if (OBJ == OBJ_QUALITY) {
        // reference: V[k] * log10(WH[k] + eps) - WH[k]  (nmf_kernels_quality.cu:35)
        return (double)fmaf(v, log10f(wh + eps), -wh);
    }
    return 0.0;
}

// The same terms formed in FP32 without cancellation, for the passes that carry the objective
// beside their multiply (the fused KL and the fused batch passes): each term within a few
// units of 2^-24 of its value, summed in FP64 by the caller.  Frobenius: (v - wh)^2 (v - wh is
// exact when v and wh are within a factor 2 of each other, else within half an ulp of it).  KL
// (v > 0): v ln(v/wh) - v + wh = v (y - ln(1 + y)) with y = (wh - v) / v, where y - ln(1 + y),
// which tends to y^2 / 2, is summed from its series sum_{k>=2} (-y)^k / k for |y| < 1/4 (terms
// up to k = 13: truncation below 1e-8 of the value) and taken as y - ln(wh / v) otherwise (the
// ratio formed directly, not as 1 + y, which loses its digits when wh << v): against the exact
// term on 300 000 sampled pairs at most 2.1e-6 relative, next to |y| = 1/4; v <= 0: wh - v, as
// the FP64 form.
template <int OBJ>
__device__ __forceinline__ float obj_term_f32(float v, float wh, float eps) {
    if (OBJ == OBJ_FROB) {
        const float d = v - wh;
        return d * d;
    } else if (OBJ == OBJ_KL) {
        if (!(v > 0.f)) return wh - v;
        const float y = (wh - v) / v;
        float f;
        if (fabsf(y) < 0.25f) {
            float s = 1.f / 13.f;
            s = 1.f / 12.f - y * s;
            s = 1.f / 11.f - y * s;
            s = 1.f / 10.f - y * s;
            s = 1.f / 9.f - y * s;
            s = 1.f / 8.f - y * s;
            s = 1.f / 7.f - y * s;
            s = 1.f / 6.f - y * s;
            s = 1.f / 5.f - y * s;
            s = 1.f / 4.f - y * s;
            s = 1.f / 3.f - y * s;
            s = 1.f / 2.f - y * s;
            f = y * y * s;
        } else {
            f = y - logf(wh / v);
        }
        return v * f;
    } else if (OBJ == OBJ_QUALITY) {
        return fmaf(v, log10f(wh + eps), -wh);
    }
    return 0.f;
}

// Three-way BF16 split of the pair (a, b), round to nearest even at every step:
// x = x1 + x2 + x3 + e with x1 = bf16(x), x2 = bf16(x - x1), x3 = bf16(x - x1 - x2); both
// differences are exact in FP32, |x2| <= 2^-9 |x|, |x3| <= 2^-18 |x|, |e| <= 2^-27 |x|
// (normal range).  Piece p of (a, b) comes packed in wp, a in the low half.
__device__ __forceinline__ void bf16x2_split3(float a, float b, uint32_t& w1, uint32_t& w2, uint32_t& w3) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    auto pack = [](f32x2_t x) { return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2_t)); };
    auto lo = [](uint32_t w) { return __uint_as_float(w << 16); };
    auto hi = [](uint32_t w) { return __uint_as_float(w & 0xffff0000u); };
    const f32x2_t x = {a, b};
    w1 = pack(x);
    const f32x2_t r1 = {a - lo(w1), b - hi(w1)};
    w2 = pack(r1);
    const f32x2_t r2 = {r1[0] - lo(w2), r1[1] - hi(w2)};
    w3 = pack(r2);
}

// Row of the 32x32 MFMA result held in accumulator register v by half-wave h.
__device__ __forceinline__ int acc_row(int v, int h) { return ((v >> 2) << 3) + (h << 2) + (v & 3); }

// Workgroups are handed to the 8 XCDs round-robin by hardware block id.  This
// maps block ids so that consecutive work items (which share a K slice and so a
// slice of W or H) run on the same XCD and hit the same L2.
__device__ __forceinline__ int xcd_work_index(int bid, int grid) {
    const int per = grid >> 3;  // grid is a multiple of 8
    return (bid & 7) * per + (bid >> 3);
}

// Guarded 4-float load of p[0..3] where element j exists iff ok && k + j < kend.
// vec: the address is 16-byte aligned whenever k is a multiple of 4.
__device__ __forceinline__ float4 load4(const float* p, bool ok, int k, int kend, bool vec) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!ok) return v;
    if (vec && k + 3 < kend) return *reinterpret_cast<const float4*>(p);
    if (k < kend) v.x = p[0];
    if (k + 1 < kend) v.y = p[1];
    if (k + 2 < kend) v.z = p[2];
    if (k + 3 < kend) v.w = p[3];
    return v;
}
// Gram reduction by the last arriving workgroup.  The workgroups that compute
// parts of the gram (arrivals of them per launch) have written their parts of
// the S slice partials; the one that arrives last sums the S partials in slice
// order (the same order every time) into G and resets the counter for the next
// pass.  off(e) is the offset of gram element e inside a slice.  Device-scope
// fences make the other XCDs' partials visible.
template <typename OffFn>
__device__ void gram_finish(const float* P, long pstride, int S, int arrivals, int count, OffFn off, float* G,
                            int* counter) {
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(counter, 1) == arrivals - 1) ? 1 : 0;
    __syncthreads();
    if (!last) return;
    __threadfence();
    // 4 elements per thread per round, 4 slices per step: 16 loads in flight
    const int stride = blockDim.x;
    for (int e0 = threadIdx.x; e0 < count; e0 += 4 * stride) {
        long o[4];
        float g[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int e = e0 + j * stride;
            o[j] = e < count ? off(e) : off(e0);
            g[j] = 0.f;
        }
        int s = 0;
        for (; s + 4 <= S; s += 4) {
            float x[4][4];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int j = 0; j < 4; j++) x[q][j] = P[(size_t)(s + q) * pstride + o[j]];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int j = 0; j < 4; j++) g[j] += x[q][j];
        }
        for (; s < S; s++)
#pragma unroll
            for (int j = 0; j < 4; j++) g[j] += P[(size_t)s * pstride + o[j]];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (e0 + j * stride < count) G[e0 + j * stride] = g[j];
    }
    if (threadIdx.x == 0) atomicExch(counter, 0);
}

// Same for a gram stored contiguously in every slice: P points at slice 0's
// gram, count floats (a multiple of 4, 16-byte aligned) per slice.  Each thread
// sums up to QJ float4 groups, 4 slices per step (4 QJ loads of 16 bytes in
// flight), adding the slices in order.  The last-arriving workgroup runs this
// while its slot could start another tile, so its round trips lengthen the pass:
// QJ = 8 takes an r = 64 gram over 8 slices in two round trips, for 128 VGPRs,
// which only kernels already held to 2 waves per SIMD can spend.
template <int QJ = 2>
__device__ inline void gram_finish4(const float* P, long pstride, int S, int arrivals, int count, float* G,
                                    int* counter) {
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(counter, 1) == arrivals - 1) ? 1 : 0;
    __syncthreads();
    if (!last) return;
    __threadfence();
    const int nq = count >> 2, stride = blockDim.x;
    const float4* P4 = reinterpret_cast<const float4*>(P);
    const long ps4 = pstride >> 2;
    for (int q0 = threadIdx.x; q0 < nq; q0 += QJ * stride) {
        int q[QJ];
        float4 g[QJ];
#pragma unroll
        for (int j = 0; j < QJ; j++) {
            q[j] = (q0 + j * stride < nq) ? q0 + j * stride : q0;
            g[j] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        int s = 0;
        for (; s + 4 <= S; s += 4) {
            float4 x[4][QJ];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < QJ; j++) x[i][j] = P4[(size_t)(s + i) * ps4 + q[j]];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < QJ; j++)
                    g[j].x += x[i][j].x, g[j].y += x[i][j].y, g[j].z += x[i][j].z, g[j].w += x[i][j].w;
        }
        for (; s < S; s++)
#pragma unroll
            for (int j = 0; j < QJ; j++) {
                const float4 x = P4[(size_t)s * ps4 + q[j]];
                g[j].x += x.x, g[j].y += x.y, g[j].z += x.z, g[j].w += x.w;
            }
#pragma unroll
        for (int j = 0; j < QJ; j++)
            if (q0 + j * stride < nq) reinterpret_cast<float4*>(G)[q0 + j * stride] = g[j];
    }
    if (threadIdx.x == 0) atomicExch(counter, 0);
}
#endif  // __HIPCC__

// Launch description of one pass (H or W) of one problem.
struct PassArgs {
    const float* V;   // n x m
    const float* W;   // n x r
    const float* H;   // r x m
    float* P;         // partials, S slices of pstride floats
    float* Psum;      // KL: S slices of rpad floats (column sums of W / row sums of H)
    double* objp;     // per-workgroup objective partials (null: no objective)
    int n, m, r;
    int ldr;          // H pass partial: P[s][u*ldr + a], u in [0, m + r) (gram columns at u >= m)
    int ldp;          // W pass partial: P[s][a*ldp + i], i in [0, npad + r) (gram rows at i >= npad)
    int npad;         // round_up(n, 4)
    int kslice;       // K extent of one slice, multiple of KC
    int S;            // slices
    int tiles;        // tiles per slice (including the gram tile for the Euclidean rule)
    long pstride;     // floats per slice of P
    float eps;
    int vec;          // V, W, H rows/columns are 16-byte aligned (n % 4 == 0 and r % 4 == 0)
    int ldv;          // leading dimension of V (FP32 V: n)
    int vtiles;       // BF16 / FP8 storage: column tiles of the tiled V (ceil(m / 64))
    const uint8_t* vsc;  // FP8 storage: E8M0 scale of V's 32 x 32 tile (ub * nbi + ib)
    int nbi;             // FP8 storage: 32-row blocks of V
    float* G;            // reduced gram of this pass's factor: gram(a,b) at G[b*rpad + a]
    int* gcount;         // arrival counter of the gram-carrying workgroups (zero between passes)
    int gsl, gks;        // fused batch engine: the gram comes in gsl K slices of gks elements (slice s
                         // at G + s*r*rpad), summed in slice order where it is read
};

// Objective pass over the split factors (nmf_objective.hip): V (FP32 n x m with leading
// dimension ldv, or a tiled BF16 / FP8 store with its column tiles and FP8 tile scales),
// the three BF16 pieces of W^T (piece p, row i at Ws + (p*nrows + i)*rp) and of H (piece p,
// column u at Hs + (p*ncols + u)*rp), rp = split_rank(r), zero padded; work items are
// rblocks 128-row blocks x csplit ranges of the cblocks 32-column blocks, one FP64
// partial each in objp.
struct SplitObjArgs {
    const void* V;
    const uint8_t* vsc;
    int nbi, vtiles, ldv, vec;
    const uint16_t* Ws;
    const uint16_t* Hs;
    int nrows, ncols;
    int n, m, r, rp;
    float eps;
    int rblocks, cblocks, csplit;
    double* objp;
};

// FP8 K-contiguous pass: as LpPassArgs, with E8M0 scales.  A's scale for column a
// and K block kb is Asc[a * nbk + kb]; B's is Bsc[(col >> bcol_shift) * bs_col + kb * bs_k].
struct Fp8PassArgs {
    const uint8_t* A;
    const uint8_t* B;
    const uint8_t* Asc;
    const uint8_t* Bsc;
    float* P;
    int K, N, r;
    int lda, ldr;     // A: column a at A + a*lda;  B: tiled (fp8_tiled_index, `tiles` column tiles)
    int kslice, S, tiles;
    long pstride;
    int nbk;
    int bcol_shift, bs_col, bs_k;
    float* G;
    int* gcount;
    int gkeep;   // 1: leave the S gram partials (columns N.. of every slice) for the update to sum
    int snt_k0;  // B chunks from contraction index snt_k0 on: non-temporal policy (K or more: none)
};

// Low-precision K-contiguous pass: P[s](a, u) = sum_k A(k, a) B(k, u) over slice s.
// A: column a at A + a*lda; B: the tiled BF16 store of one V orientation (element (k, u)
// at B[bf16_tiled_index(k, u, tiles)]); the slice's gram (columns N .. N + r - 1 of P)
// takes both operands from A (block t on column tile t when tiles >= its block count, else
// all on tile 0).  P[s][u*ldr + a].
struct LpPassArgs {
    const uint16_t* A;
    const uint16_t* B;
    float* P;
    int K, N, r;
    int lda, ldr;
    int kslice, S, tiles;
    int nt;       // 64-column store blocks per workgroup tile (1, or 2 for r <= 64); tiles counts those tiles
    long pstride;
    float* G;
    int* gcount;
    int gkeep;   // 1: leave the S gram partials (columns N.. of every slice) for the update to sum
    long aplane; // FP32X pass (B: the tiled FP32 store, x3_tiled_index): element stride between
                 // the three BF16 piece planes of A
    int snt_k0;  // BF16 pass: B chunks from contraction index snt_k0 on are read with the non-temporal
                 // policy (stream_load); K or more: none
};

// KL pass over the BF16 V store: P[s](a, col) = sum_{k in slice s} A(k, a) R(k, col) with the ratio
// R = bf16(V(k, col) / ((A C)(k, col) + eps)), A C formed per 32 x 32 block on the matrix cores and
// never stored.  A comes twice, K-contiguous (Ak[a*lda + k], the second product) and
// rank-contiguous (Ar[k*rps + a], the A C product); C rank-contiguous (Cr[col*rps + a]).  Both
// rank-contiguous shadows are zero for a in [r, rps) and past the last row.  Tile 0 of every
// slice also writes the slice's column sums of A, Psum[s*rpad + a] = sum_k A(k, a).
// Row pitch of the rank-contiguous shadows of the KL rule with BF16 storage: the rank rounded up
// to the pass's rank blocks (32, 64 or 128).
__host__ __device__ inline int kl_rank_pitch(int r) { return r <= 32 ? 32 : (r <= 64 ? 64 : 128); }

struct KlLpArgs {
    const uint16_t* Ak;
    const uint16_t* Ar;
    const uint16_t* Cr;
    const uint16_t* B;   // the tiled store of one V orientation (bf16_tiled_index, vtiles column tiles)
    float* P;
    float* Psum;
    int K, N, r;
    int lda, rps, ldr, rpad;
    int kslice, S, tiles;
    int nt;              // 64-column store blocks per workgroup tile (2 for r <= 64)
    long pstride;
    float eps;
    int snt_k0;          // B chunks from contraction index snt_k0 on: non-temporal policy (K or more: none)
};

// One problem of a grouped (batched) launch: the arguments of its two passes and
// of its updates.  P, Psum and objp point into the problem's own workspace, so
// no value depends on which other problems share the launch.
struct JobArgs {
    PassArgs h;    // H pass (rows of V), also the H update's partials
    PassArgs w;    // W pass (columns of V)
    float* W;      // updated in place
    float* H;
    double* obj;   // objective outputs of the problem (two slots)
    int rule;      // 0 Euclidean, 1 KL
    int nobj;      // H-pass workgroups = objective partials
};

// Grouped launch table in device memory (capacity maxj problems):
//   t[0] = problems J, t[1] = work items T, t[2 .. 2+J] = first item of each
//   problem (t[2+J] = T), t[3+maxj .. 3+maxj+J) = index into the JobArgs array.
constexpr int kGroupMaxJobs = 1024;
inline int group_table_ints(int maxj) { return 3 + 2 * maxj; }


}  // namespace gpumlib_nmf
