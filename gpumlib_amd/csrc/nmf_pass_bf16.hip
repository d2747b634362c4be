// BF16 storage path: V held in BF16 in both orientations (V and V^T, each
// column-major with a 16-byte-aligned leading dimension), BF16 shadows of W and
// H as MFMA operands, FP32 accumulation and FP32 updates.
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

namespace gpumlib_nmf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int KB = 64;          // K per pipeline step (4 MFMA k-steps of 16)
constexpr int LDB = KB + 8;     // LDS row pitch in bf16 elements (144 bytes)

__device__ __forceinline__ f32x16 mfma_bf16(uint4 a, uint4 b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
}

// Guarded load of 8 bf16 p[0..7]; element j exists iff ok && k + j < kend.  Rows
// start 16-byte aligned (leading dimensions are multiples of 8).
__device__ __forceinline__ uint4 load8_bf16(const uint16_t* p, bool ok, int k, int kend) {
    if (!ok) return make_uint4(0, 0, 0, 0);
    if (k + 7 < kend) return *reinterpret_cast<const uint4*>(p);
    uint16_t e[8];
#pragma unroll
    for (int j = 0; j < 8; j++) e[j] = (k + j < kend) ? p[j] : (uint16_t)0;
    return make_uint4(e[0] | ((uint32_t)e[1] << 16), e[2] | ((uint32_t)e[3] << 16), e[4] | ((uint32_t)e[5] << 16),
                      e[6] | ((uint32_t)e[7] << 16));
}

// One workgroup per (slice s, column tile t); 4 waves, wave w owns columns
// w*32 .. w*32+31 of the 128-column tile and all MB row blocks.  The first
// column tile of every slice also accumulates the slice's gram A^T A from the
// A rows already staged in LDS (gram block (mb, nb) on wave mb*MB + nb, modulo
// 4), so the grid is exactly tiles x S workgroups of one shape and the gram
// costs no extra pass over A.  The last of the S gram workgroups reduces the
// partial grams into G.
template <int MB>
__global__ __launch_bounds__(NTHREADS, 2) void k_pass_tn_bf16(LpPassArgs p) {
    constexpr int RP = MB * 32;
    constexpr int GB = (MB * MB + 3) / 4;  // gram blocks per wave
    __shared__ __attribute__((aligned(16))) uint16_t sA[2][RP][LDB];
    __shared__ __attribute__((aligned(16))) uint16_t sB[2][TILE][LDB];

    const int total = p.tiles * p.S;
    const int w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= total) return;
    const int s = w / p.tiles, t = w - s * p.tiles;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = lane & 31, h = lane >> 5;
    const bool gram = (t == 0) && p.G != nullptr;
    const int col0 = t * TILE;
    const int ncols = min(TILE, p.N - col0);
    const int kbeg = s * p.kslice;
    const int kend = min(p.K, kbeg + p.kslice);

    uint4 ra[MB], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = tid + NTHREADS * j, col = q >> 3, kq = (q & 7) * 8;
            rb[j] = load8_bf16(p.B + (size_t)(col0 + col) * p.ldb + k0 + kq, col < ncols, k0 + kq, kend);
        }
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const int q = tid + NTHREADS * j, a = q >> 3, kq = (q & 7) * 8;
            ra[j] = load8_bf16(p.A + (size_t)a * p.lda + k0 + kq, a < p.r, k0 + kq, kend);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = tid + NTHREADS * j, col = q >> 3, kq = (q & 7) * 8;
            *reinterpret_cast<uint4*>(&sB[buf][col][kq]) = rb[j];
        }
#pragma unroll
        for (int j = 0; j < MB; j++) {
            const int q = tid + NTHREADS * j, a = q >> 3, kq = (q & 7) * 8;
            *reinterpret_cast<uint4*>(&sA[buf][a][kq]) = ra[j];
        }
    };

    f32x16 acc[MB], gacc[GB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
        for (int v = 0; v < 16; v++) acc[mb][v] = 0.f;
#pragma unroll
    for (int q = 0; q < GB; q++)
        for (int v = 0; v < 16; v++) gacc[q][v] = 0.f;

    const int nchunks = (kend - kbeg + KB - 1) / KB;
    gload(kbeg);
    sstore(0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
        const int buf = ch & 1;
        if (ch + 1 < nchunks) gload(kbeg + (ch + 1) * KB);
        // lane (c, h) supplies k = g*16 + h*8 + j (j < 8) for both operands
#pragma unroll
        for (int g = 0; g < KB / 16; g++) {
            const uint4 bv = *reinterpret_cast<const uint4*>(&sB[buf][wave * 32 + c][g * 16 + h * 8]);
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                const uint4 av = *reinterpret_cast<const uint4*>(&sA[buf][mb * 32 + c][g * 16 + h * 8]);
                acc[mb] = mfma_bf16(av, bv, acc[mb]);
            }
        }
        if (gram) {
#pragma unroll
            for (int q = 0; q < GB; q++) {
                const int blk = wave + 4 * q;
                if (blk < MB * MB) {
                    const int mb = blk / MB, nb = blk - (blk / MB) * MB;
#pragma unroll
                    for (int g = 0; g < KB / 16; g++) {
                        const uint4 av = *reinterpret_cast<const uint4*>(&sA[buf][mb * 32 + c][g * 16 + h * 8]);
                        const uint4 bv = *reinterpret_cast<const uint4*>(&sA[buf][nb * 32 + c][g * 16 + h * 8]);
                        gacc[q] = mfma_bf16(av, bv, gacc[q]);
                    }
                }
            }
        }
        if (ch + 1 < nchunks) sstore(buf ^ 1);
        __syncthreads();
    }

    // acc[mb][v] = P(a = mb*32 + acc_row(v,h), u = wave*32 + c)
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
        gram_finish4(p.P + (size_t)p.N * p.ldr, p.pstride, p.S, p.r * p.ldr, p.G, p.gcount);
    }
}

// ---------------------------------------------------------------------------
// V (n x m, FP32 column-major) -> Vb (column-major, ld ldv) and Vt (V^T column-major,
// ld ldt), both BF16 (round to nearest even).  64 x 64 tiles through LDS.
__global__ __launch_bounds__(NTHREADS) void k_pack_v_bf16(const float* __restrict__ V, int n, int m,
                                                          uint16_t* __restrict__ Vb, int ldv,
                                                          uint16_t* __restrict__ Vt, int ldt) {
    __shared__ uint16_t tile[64][66];
    const int i0 = blockIdx.x * 64, u0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int uu = ty; uu < 64; uu += 4) {
        const int i = i0 + tx, u = u0 + uu;
        uint16_t b = 0;
        if (i < n && u < m) {
            b = __builtin_bit_cast(uint16_t, (__bf16)V[(size_t)u * n + i]);
            Vb[(size_t)u * ldv + i] = b;
        }
        tile[uu][tx] = b;
    }
    __syncthreads();
    for (int ii = ty; ii < 64; ii += 4) {
        const int u = u0 + tx, i = i0 + ii;
        if (i < n && u < m) Vt[(size_t)i * ldt + u] = tile[tx][ii];
    }
}

// W (n x r, FP32) -> Wb (column-major, ld ldw);  H (r x m, FP32) -> Ht (H^T column-major, ld ldh).
__global__ __launch_bounds__(NTHREADS) void k_shadow_bf16(const float* __restrict__ W, const float* __restrict__ H,
                                                          int n, int m, int r, uint16_t* __restrict__ Wb, int ldw,
                                                          uint16_t* __restrict__ Ht, int ldh) {
    const long nw = (long)n * r, nh = (long)r * m;
    for (long e = (long)blockIdx.x * NTHREADS + threadIdx.x; e < nw + nh; e += (long)gridDim.x * NTHREADS) {
        if (e < nw) {
            const long a = e / n, i = e - a * n;
            Wb[a * ldw + i] = __builtin_bit_cast(uint16_t, (__bf16)W[e]);
        } else {
            const long f = e - nw, a = f / m, u = f - a * m;
            Ht[a * ldh + u] = __builtin_bit_cast(uint16_t, (__bf16)H[u * r + a]);
        }
    }
}

// ---------------------------------------------------------------------------
template <int MB>
static void launch_tn(const LpPassArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_pass_tn_bf16<MB>), dim3(xcd_grid(a.tiles * a.S)), dim3(NTHREADS), 0, st, a);
}

hipError_t launch_pass_tn_bf16(const LpPassArgs& a, hipStream_t st) {
    const int rp = round_up(a.r, 32);
    if (rp <= 32)
        launch_tn<1>(a, st);
    else if (rp <= 64)
        launch_tn<2>(a, st);
    else if (rp <= 128)
        launch_tn<4>(a, st);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pack_v_bf16(const float* V, int n, int m, uint16_t* Vb, int ldv, uint16_t* Vt, int ldt,
                              hipStream_t st) {
    dim3 grid((n + 63) / 64, (m + 63) / 64);
    hipLaunchKernelGGL(k_pack_v_bf16, grid, dim3(NTHREADS), 0, st, V, n, m, Vb, ldv, Vt, ldt);
    return hipGetLastError();
}

hipError_t launch_shadow_bf16(const float* W, const float* H, int n, int m, int r, uint16_t* Wb, int ldw, uint16_t* Ht,
                              int ldh, hipStream_t st) {
    const long total = (long)n * r + (long)r * m;
    long blocks = (total + NTHREADS - 1) / NTHREADS;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_shadow_bf16, dim3((unsigned)blocks), dim3(NTHREADS), 0, st, W, H, n, m, r, Wb, ldw, Ht, ldh);
    return hipGetLastError();
}

}  // namespace gpumlib_nmf
