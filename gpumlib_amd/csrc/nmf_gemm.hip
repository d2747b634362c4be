// General GEMM on the matrix cores for the container API:
//   C (M x N, column-major, ldc) = alpha * op(A) * op(B) + beta * C
// with op = 'N' or 'T', the contract of DeviceMatrix<float>::Multiply /
// MultiplyBySelfTranspose (reference src/memory/DeviceMatrix.h:195-227, which
// call cublasSgemm).  beta == 0 never reads C.
//
// Operand formats: FP32 (v_mfma_f32_32x32x2f32), BF16 bit patterns
// (v_mfma_f32_32x32x16_bf16) and OCP E4M3 with one E8M0 scale per operand
// (v_mfma_scale_f32_32x32x64_f8f6f4, the operand scales given to the instruction); the
// BF16 and FP8 forms have no reference counterpart.  Products of two BF16 or two scaled
// E4M3 values are exact in FP32 and accumulate in FP32, so the FP32 and BF16 forms are the
// FP32-accumulated GEMM of the stored values; the FP8 matrix cores sum each 64-term block of
// K at less than FP32 accuracy (measured: up to ~6e-5 relative on full-range operands).
//
// 128 x 128 output tile per workgroup of 4 waves (2 x 2, 64 x 64 each: 2 x 2 blocks of 32 x 32);
// per stage 16 KB of each operand (KT = 32 FP32, 64 BF16 or 128 E4M3 elements of K) staged
// through double-buffered LDS with K contiguous in every row, whichever operand dimension is
// contiguous in memory: 16-byte global loads run along that dimension (along K for 'T' A and
// 'N' B, written to LDS as they are; along M or N otherwise, scattered into K-contiguous rows).
// Loads of the next stage are in flight in registers while the current one is multiplied.
// Tiles crossing the matrix edge, operands whose rows are not 16-byte aligned and the last
// partial stage of K load element by element, zeros outside the matrix.  Each lane feeds the matrix cores a contiguous run
// of K (4, 8 or 32 elements) for its row and column, the same run for both operands, so the
// contraction covers every k once.
#include "nmf_common.hpp"
#include "nmf_internal.hpp"

namespace gpumlib_nmf {

typedef __bf16 gbf16x8 __attribute__((ext_vector_type(8)));
typedef int gi32x8 __attribute__((ext_vector_type(8)));

constexpr int GT = 128;     // output tile (rows and columns)
constexpr int GROWB = 128;  // bytes of K per staged row and stage
constexpr int GPITCH = GROWB + 16;

template <int FMT> struct GemmFmt;
template <> struct GemmFmt<0> { static constexpr int ES = 4; typedef uint32_t elem; };  // FP32
template <> struct GemmFmt<1> { static constexpr int ES = 2; typedef uint16_t elem; };  // BF16
template <> struct GemmFmt<2> { static constexpr int ES = 1; typedef uint8_t elem; };   // E4M3

// One operand tile of a stage: GT rows (the output dimension) x KT elements of K.  kc: K is the
// contiguous dimension in memory (row r at p + r*ld); otherwise the row dimension is (k at
// p + k*ld).  load() issues 16-byte loads along the contiguous dimension into registers (a
// stage wholly inside K, rows 16-byte aligned, and each chunk wholly inside or wholly outside
// the rows: the rows of a chunk along them are a multiple of EPC); chunks outside are zeros.
// store() writes them to LDS rows with K contiguous.  load_direct() is the general form: element by element, zeros
// outside the matrix, straight into LDS (edge tiles and the last partial stage of K).
template <int FMT>
struct TileLoader {
    static constexpr int ES = GemmFmt<FMT>::ES;
    static constexpr int KT = GROWB / ES;     // K per stage
    static constexpr int EPC = 16 / ES;       // elements per 16-byte chunk
    static constexpr int CHUNKS = GT * KT / EPC;
    static constexpr int CPT = CHUNKS / NTHREADS;  // chunks per thread (4)
    static_assert(CPT * NTHREADS == CHUNKS, "tile split");
    typedef typename GemmFmt<FMT>::elem E;
    uint4 v[CPT];

    __device__ __forceinline__ void load(const uint8_t* __restrict__ p, long ld, int kc, int r0, int rows, int k0) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int q = threadIdx.x + NTHREADS * j;
            int r, k;
            if (kc) r = q / (KT / EPC), k = (q % (KT / EPC)) * EPC;  // chunk along K
            else k = q / (GT / EPC), r = (q % (GT / EPC)) * EPC;     // chunk along the rows
            const size_t off = kc ? (size_t)(r0 + r) * ld + k0 + k : (size_t)(k0 + k) * ld + r0 + r;
            v[j] = r0 + r < rows ? *reinterpret_cast<const uint4*>(p + off * ES) : make_uint4(0u, 0u, 0u, 0u);
        }
    }
    __device__ __forceinline__ void store(uint8_t* s, int kc) const {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int q = threadIdx.x + NTHREADS * j;
            if (kc) {
                const int r = q / (KT / EPC), k = (q % (KT / EPC)) * EPC;
                *reinterpret_cast<uint4*>(s + r * GPITCH + k * ES) = v[j];
            } else {
                const int k = q / (GT / EPC), r = (q % (GT / EPC)) * EPC;
                const E* b = reinterpret_cast<const E*>(&v[j]);
#pragma unroll
                for (int e = 0; e < EPC; e++) *reinterpret_cast<E*>(s + (r + e) * GPITCH + k * ES) = b[e];
            }
        }
    }
    static __device__ void load_direct(uint8_t* s, const uint8_t* __restrict__ p, long ld, int kc, int r0, int rows,
                                       int k0, int K) {
        const E* pe = reinterpret_cast<const E*>(p);
        for (int e = threadIdx.x; e < GT * KT; e += NTHREADS) {
            int r, k;
            if (kc) r = e / KT, k = e - (e / KT) * KT;  // consecutive threads along K
            else k = e / GT, r = e - (e / GT) * GT;     // along the rows
            const int gr = r0 + r, gk = k0 + k;
            const E x = (gr < rows && gk < K) ? pe[kc ? (size_t)gr * ld + gk : (size_t)gk * ld + gr] : (E)0;
            *reinterpret_cast<E*>(s + r * GPITCH + k * ES) = x;
        }
    }
};

// One stage's products for a wave's 2 x 2 blocks: lane (c, h) takes row c of each block and
// the K run [h*R, h*R + R) of every MFMA step (R = 4 FP32 over 4 MFMAs, 8 BF16, 32 E4M3).
template <int FMT>
__device__ __forceinline__ void gemm_stage(const uint8_t* sa, const uint8_t* sb, int wm, int wn, int c, int h,
                                           int sca, int scb, f32x16 (&acc)[2][2]) {
    if constexpr (FMT == 0) {
#pragma unroll
        for (int g = 0; g < GROWB / 32; g++) {  // 8 K per group: 4 MFMAs of K = 2
            f32x4 a[2], b[2];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                a[x] = *reinterpret_cast<const f32x4*>(sa + (wm * 64 + x * 32 + c) * GPITCH + g * 32 + h * 16);
                b[x] = *reinterpret_cast<const f32x4*>(sb + (wn * 64 + x * 32 + c) * GPITCH + g * 32 + h * 16);
            }
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++) acc[x][y] = mfma32x2(a[x][j], b[y][j], acc[x][y]);
        }
    } else if constexpr (FMT == 1) {
#pragma unroll
        for (int g = 0; g < GROWB / 32; g++) {  // 16 K per MFMA
            uint4 a[2], b[2];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                a[x] = *reinterpret_cast<const uint4*>(sa + (wm * 64 + x * 32 + c) * GPITCH + g * 32 + h * 16);
                b[x] = *reinterpret_cast<const uint4*>(sb + (wn * 64 + x * 32 + c) * GPITCH + g * 32 + h * 16);
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(gbf16x8, a[x]),
                                                                        __builtin_bit_cast(gbf16x8, b[y]), acc[x][y],
                                                                        0, 0, 0);
        }
    } else {
#pragma unroll
        for (int g = 0; g < GROWB / 64; g++) {  // 64 K per MFMA: 32 bytes per lane
            gi32x8 a[2], b[2];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                const uint8_t* pa = sa + (wm * 64 + x * 32 + c) * GPITCH + g * 64 + h * 32;
                const uint8_t* pb = sb + (wn * 64 + x * 32 + c) * GPITCH + g * 64 + h * 32;
                const uint4 a0 = *reinterpret_cast<const uint4*>(pa), a1 = *reinterpret_cast<const uint4*>(pa + 16);
                const uint4 b0 = *reinterpret_cast<const uint4*>(pb), b1 = *reinterpret_cast<const uint4*>(pb + 16);
                a[x] = gi32x8{(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                b[x] = gi32x8{(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    acc[x][y] =
                        __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[x], b[y], acc[x][y], 0, 0, 0, sca, 0, scb);
        }
    }
}

// ta / tb: 1 for 'T'.  op(A)(i, k) is A[k*lda + i] ('N') or A[i*lda + k] ('T'); op(B)(k, j)
// is B[j*ldb + k] ('N') or B[k*ldb + j] ('T').  sca / scb: the E8M0 operand scales (E4M3 only).
// vec: bit 0 A, bit 1 B have 16-byte aligned rows along their contiguous dimension; cvec: so does C.
template <int FMT>
__global__ __launch_bounds__(NTHREADS) void k_gemm(int ta, int tb, int M, int N, int K, float alpha,
                                                   const void* __restrict__ A, long lda, int sca,
                                                   const void* __restrict__ B, long ldb, int scb, float beta,
                                                   float* __restrict__ C, long ldc, int vec, int cvec) {
    // stage buffers: A of buffer b at smem + b * GT * GPITCH, B of buffer b at smem + (2 + b) * GT * GPITCH;
    // after the last stage the first CPITCH * GT floats hold the C tile, [column][row]
    constexpr int CPITCH = GT + 4;
    static_assert(CPITCH * GT * 4 <= 4 * GT * GPITCH, "C tile fits in the stage buffers");
    __shared__ __attribute__((aligned(16))) uint8_t smem[4 * GT * GPITCH];
    auto sA = [&](int b) { return smem + b * GT * GPITCH; };
    auto sB = [&](int b) { return smem + (2 + b) * GT * GPITCH; };
    constexpr int KT = TileLoader<FMT>::KT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = lane & 31, h = lane >> 5;
    const int i0 = blockIdx.x * GT, j0 = blockIdx.y * GT;
    const int wm = wave & 1, wn = wave >> 1;
    // A rows are the output rows i (K contiguous when 'T'); B rows the output columns j
    // (K contiguous when 'N')
    const int akc = ta, bkc = !tb;
    const uint8_t* pa = static_cast<const uint8_t*>(A);
    const uint8_t* pb = static_cast<const uint8_t*>(B);
    f32x16 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
            for (int v = 0; v < 16; v++) acc[x][y][v] = 0.f;

    // Stages [0, nfast) of a tile entirely inside M x N with aligned rows take the register-
    // prefetched vector loads; the others (edge tiles, the last partial stage of K) load
    // element by element.
    TileLoader<FMT> la, lb;
    constexpr int EPC = TileLoader<FMT>::EPC;
    const int nst = (K + KT - 1) / KT;
    // a chunk along the rows lies wholly inside or outside them when their count is a multiple of EPC
    const bool fast = (vec == 3) && (akc || M % EPC == 0) && (bkc || N % EPC == 0);
    const int nfast = fast ? K / KT : 0;
    auto fill = [&](int buf, int st) __attribute__((always_inline)) {
        if (st < nfast) {
            la.store(sA(buf), akc);
            lb.store(sB(buf), bkc);
        } else {
            TileLoader<FMT>::load_direct(sA(buf), pa, lda, akc, i0, M, st * KT, K);
            TileLoader<FMT>::load_direct(sB(buf), pb, ldb, bkc, j0, N, st * KT, K);
        }
    };
    if (nst > 0) {
        if (nfast > 0) {
            la.load(pa, lda, akc, i0, M, 0);
            lb.load(pb, ldb, bkc, j0, N, 0);
        }
        fill(0, 0);
    }
    __syncthreads();
    for (int st = 0; st < nst; st++) {
        const int cur = st & 1;
        if (st + 1 < nfast) {
            la.load(pa, lda, akc, i0, M, (st + 1) * KT);
            lb.load(pb, ldb, bkc, j0, N, (st + 1) * KT);
        }
        gemm_stage<FMT>(sA(cur), sB(cur), wm, wn, c, h, sca, scb, acc);
        if (st + 1 < nst) fill(cur ^ 1, st + 1);
        __syncthreads();
    }
    // lane (c, h) holds rows acc_row(v, h) of column c of each block (runs of 4 rows): the tile
    // goes through LDS, [column][row], so that every column's rows leave in one contiguous run
    float* sc = reinterpret_cast<float*>(smem);
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int col = wn * 64 + y * 32 + c, row = wm * 64 + x * 32 + g * 8 + h * 4;
                *reinterpret_cast<f32x4*>(sc + col * CPITCH + row) =
                    f32x4{acc[x][y][g * 4], acc[x][y][g * 4 + 1], acc[x][y][g * 4 + 2], acc[x][y][g * 4 + 3]};
            }
    __syncthreads();
    // 32 threads per column, 4 rows each; 8 columns per pass of the workgroup
    const int rq = (tid & 31) * 4;
    for (int col = tid >> 5; col < GT; col += NTHREADS / 32) {
        const int j = j0 + col, i = i0 + rq;
        if (j >= N || i >= M) continue;
        const f32x4 a = *reinterpret_cast<const f32x4*>(sc + col * CPITCH + rq);
        float* cp = C + (size_t)j * ldc + i;
        if (cvec && i + 4 <= M) {
            f32x4 r = alpha * a;
            if (beta != 0.f) r += beta * *reinterpret_cast<const f32x4*>(cp);
            *reinterpret_cast<f32x4*>(cp) = r;
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (i + e < M) cp[e] = (beta == 0.f) ? alpha * a[e] : alpha * a[e] + beta * cp[e];
        }
    }
}

// fmt 0 FP32, 1 BF16, 2 E4M3 (sa, sb: E8M0 scales of A and B).
hipError_t launch_gemm(int fmt, char ta, char tb, int M, int N, int K, float alpha, const void* A, int lda, int sa,
                       const void* B, int ldb, int sb, float beta, float* C, int ldc, hipStream_t st) {
    auto flag = [](char t) { return (t == 'T' || t == 't') ? 1 : ((t == 'N' || t == 'n') ? 0 : -1); };
    const int fa = flag(ta), fb = flag(tb);
    if (fa < 0 || fb < 0 || M < 0 || N < 0 || K < 0 || fmt < 0 || fmt > 2) return hipErrorInvalidValue;
    if (sa < 0 || sa > 254 || sb < 0 || sb > 254) return hipErrorInvalidValue;
    if (ldc < (M > 1 ? M : 1)) return hipErrorInvalidValue;
    if (M == 0 || N == 0) return hipSuccess;
    const int es = fmt == 0 ? 4 : (fmt == 1 ? 2 : 1);
    auto rows16 = [&](const void* p, int ld) { return ((uintptr_t)p % 16 == 0) && ((size_t)ld * es % 16 == 0); };
    const int vec = (rows16(A, lda) ? 1 : 0) | (rows16(B, ldb) ? 2 : 0);
    const int cvec = ((uintptr_t)C % 16 == 0) && (ldc % 4 == 0);
    dim3 grid((M + GT - 1) / GT, (N + GT - 1) / GT);
    if (fmt == 0)
        hipLaunchKernelGGL(k_gemm<0>, grid, dim3(NTHREADS), 0, st, fa, fb, M, N, K, alpha, A, (long)lda, sa, B,
                           (long)ldb, sb, beta, C, (long)ldc, vec, cvec);
    else if (fmt == 1)
        hipLaunchKernelGGL(k_gemm<1>, grid, dim3(NTHREADS), 0, st, fa, fb, M, N, K, alpha, A, (long)lda, sa, B,
                           (long)ldb, sb, beta, C, (long)ldc, vec, cvec);
    else
        hipLaunchKernelGGL(k_gemm<2>, grid, dim3(NTHREADS), 0, st, fa, fb, M, N, K, alpha, A, (long)lda, sa, B,
                           (long)ldb, sb, beta, C, (long)ldc, vec, cvec);
    return hipGetLastError();
}

hipError_t launch_sgemm(char ta, char tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                        int ldb, float beta, float* C, int ldc, hipStream_t st) {
    return launch_gemm(0, ta, tb, M, N, K, alpha, A, lda, 127, B, ldb, 127, beta, C, ldc, st);
}

}  // namespace gpumlib_nmf
