// Runs the container GEMM surface of the drop-in API on the GPU and writes every
// case's operands and results for tests/test_gpu_dropin_gemm.py to check against
// an FP64 product:
//   * DeviceMatrix<float>::Multiply(A, B, C, alpha, beta) with row- and column-major
//     A and B (a row-major operand goes to the GEMM as transposed, DeviceMatrix.h:217-227),
//     row-major C with beta = 0 (made column-major) and column-major C with beta != 0,
//     and the reference's own use: W.ReplaceByTranspose() as the left operand
//     (nmf_multiplicative_euclidian.cpp:33-37);
//   * DeviceMatrix<float>::MultiplyBySelfTranspose(C, alpha, beta) (DeviceMatrix.h:195-205)
//     on both storage orders;
//   * the C++ helpers KernelSumW, KernelSumH and KernelNMFquality (nmf_kernels.h:44-73).
// Usage: dropin_gemm_probe OUT_DIR.  Each case k is written to OUT_DIR/case_k.bin:
//   int32 kind (0 Multiply, 1 MultiplyBySelfTranspose, 2 sums, 3 quality), M, N, K,
//   a_row, b_row, c_row_before, c_row_after; float32 alpha, beta; then the matrices in
//   logical row-major order (element (i, j) at i * cols + j):
//   Multiply: A (M x K), B (K x N), C0 (M x N), C (M x N);
//   MultiplyBySelfTranspose: A (M x K), C0 (M x M), C (M x M);
//   sums: W (M x K, i.e. n x r), H (K x N, r x m), sumW (K), sumH (K);
//   quality: V (M), WH (M), quality (1).
// Exit status 1 when a call reports an error.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "../../gpumlib_amd/src/cuda/nmf/nmf_kernels.h"
#include "../../gpumlib_amd/src/memory/DeviceArray.h"
#include "../../gpumlib_amd/src/memory/DeviceMatrix.h"
#include "../../gpumlib_amd/src/memory/HostArray.h"
#include "../../gpumlib_amd/src/memory/HostMatrix.h"

using namespace GPUMLib;

namespace {

std::string out_dir;
int case_no = 0;
int errors = 0;

// Deterministic values in [0.25, 1.25) (a small LCG; the values only need to be varied).
struct Lcg {
    unsigned long long s;
    float next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return 0.25f + (float)((s >> 40) & 0xFFFFFF) / 16777216.0f;
    }
};

HostMatrix<cudafloat> random_host(size_t rows, size_t cols, StoringOrder order, Lcg& g) {
    HostMatrix<cudafloat> x(rows, cols, order);
    for (size_t i = 0; i < rows; i++)
        for (size_t j = 0; j < cols; j++) x(i, j) = g.next();
    return x;
}

void put_logical(FILE* f, const HostMatrix<cudafloat>& x) {
    std::vector<float> v(x.Rows() * x.Columns());
    for (size_t i = 0; i < x.Rows(); i++)
        for (size_t j = 0; j < x.Columns(); j++) v[i * x.Columns() + j] = x(i, j);
    fwrite(v.data(), sizeof(float), v.size(), f);
}

FILE* open_case(int kind, int M, int N, int K, int arow, int brow, int crow0, int crow1, float alpha, float beta) {
    const std::string path = out_dir + "/case_" + std::to_string(case_no++) + ".bin";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        errors++;
        return nullptr;
    }
    const int hdr[8] = {kind, M, N, K, arow, brow, crow0, crow1};
    fwrite(hdr, sizeof(int), 8, f);
    fwrite(&alpha, sizeof(float), 1, f);
    fwrite(&beta, sizeof(float), 1, f);
    return f;
}

void check_device(const char* what) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        errors++;
    }
}

void multiply_case(int M, int N, int K, bool arow, bool brow, bool crow, float alpha, float beta, Lcg& g,
                   bool replace_by_transpose) {
    HostMatrix<cudafloat> hA;
    DeviceMatrix<cudafloat> A;
    if (replace_by_transpose) {
        // the reference's W^T: a column-major K x M matrix viewed as its (row-major) transpose
        HostMatrix<cudafloat> w = random_host(K, M, ColumnMajor, g);
        A = w;
        A.ReplaceByTranspose();
        hA = HostMatrix<cudafloat>(A);
    } else {
        hA = random_host(M, K, arow ? RowMajor : ColumnMajor, g);
        A = hA;
    }
    HostMatrix<cudafloat> hB = random_host(K, N, brow ? RowMajor : ColumnMajor, g);
    HostMatrix<cudafloat> hC0 = random_host(M, N, crow ? RowMajor : ColumnMajor, g);
    DeviceMatrix<cudafloat> B(hB), C(hC0);
    DeviceMatrix<cudafloat>::Multiply(A, B, C, alpha, beta);
    check_device("Multiply");
    HostMatrix<cudafloat> hC(C);
    FILE* f = open_case(0, M, N, K, A.IsRowMajor(), brow, crow, C.IsRowMajor(), alpha, beta);
    if (!f) return;
    put_logical(f, hA);
    put_logical(f, hB);
    put_logical(f, hC0);
    put_logical(f, hC);
    fclose(f);
}

void self_transpose_case(int M, int K, bool arow, bool crow, float alpha, float beta, Lcg& g) {
    HostMatrix<cudafloat> hA = random_host(M, K, arow ? RowMajor : ColumnMajor, g);
    HostMatrix<cudafloat> hC0 = random_host(M, M, crow ? RowMajor : ColumnMajor, g);
    DeviceMatrix<cudafloat> A(hA), C(hC0);
    A.MultiplyBySelfTranspose(C, alpha, beta);
    check_device("MultiplyBySelfTranspose");
    HostMatrix<cudafloat> hC(C);
    FILE* f = open_case(1, M, M, K, arow, arow, crow, C.IsRowMajor(), alpha, beta);
    if (!f) return;
    put_logical(f, hA);
    put_logical(f, hC0);
    put_logical(f, hC);
    fclose(f);
}

void sums_case(int n, int r, int m, Lcg& g) {
    HostMatrix<cudafloat> hW = random_host(n, r, ColumnMajor, g), hH = random_host(r, m, ColumnMajor, g);
    for (int i = 0; i < n; i++) hW(i, r / 2) = 0.f;  // a column that sums to 0: clamped at eps
    for (int u = 0; u < m; u++) hH(r - 1, u) = 0.f;  // a row that sums to 0
    DeviceMatrix<cudafloat> W(hW), H(hH);
    DeviceArray<cudafloat> sw(r), sh(r);
    if (KernelSumW(256, W.Pointer(), n, r, sw.Pointer()) != hipSuccess) errors++;
    if (KernelSumH(256, H.Pointer(), r, m, sh.Pointer()) != hipSuccess) errors++;
    check_device("KernelSumW/KernelSumH");
    HostArray<cudafloat> a(sw), b(sh);
    FILE* f = open_case(2, n, m, r, 0, 0, 0, 0, 1.f, 0.f);
    if (!f) return;
    put_logical(f, hW);
    put_logical(f, hH);
    fwrite(a.Pointer(), sizeof(float), r, f);
    fwrite(b.Pointer(), sizeof(float), r, f);
    fclose(f);
}

void quality_case(int count, Lcg& g) {
    HostArray<cudafloat> v(count), wh(count);
    for (int k = 0; k < count; k++) v[k] = 4.f * g.next(), wh[k] = 4.f * g.next();
    wh[0] = 0.f;  // log10(0 + eps)
    DeviceArray<cudafloat> dv(v), dwh(wh), q(1);
    if (KernelNMFquality(256, dv.Pointer(), dwh.Pointer(), count, q.Pointer()) != hipSuccess) errors++;
    check_device("KernelNMFquality");
    HostArray<cudafloat> hq(q);
    FILE* f = open_case(3, count, 1, 1, 0, 0, 0, 0, 1.f, 0.f);
    if (!f) return;
    fwrite(v.Pointer(), sizeof(float), count, f);
    fwrite(wh.Pointer(), sizeof(float), count, f);
    fwrite(hq.Pointer(), sizeof(float), 1, f);
    fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: dropin_gemm_probe OUT_DIR\n");
        return 2;
    }
    out_dir = argv[1];
    Lcg g{12345};
    const int shapes[][3] = {{1, 1, 1}, {37, 53, 29}, {64, 64, 64}, {130, 7, 257}, {200, 150, 48}};
    for (const auto& s : shapes)
        for (int arow = 0; arow < 2; arow++)
            for (int brow = 0; brow < 2; brow++) {
                multiply_case(s[0], s[1], s[2], arow, brow, true, 1.f, 0.f, g, false);      // C row-major, beta 0
                multiply_case(s[0], s[1], s[2], arow, brow, false, 0.75f, 0.5f, g, false);  // C column-major, beta
            }
    multiply_case(64, 300, 500, true, false, false, 1.f, 0.f, g, true);  // W^T V with W^T = W.ReplaceByTranspose()
    multiply_case(33, 65, 129, true, true, true, 2.f, 0.f, g, true);
    for (const auto& s : shapes)
        for (int arow = 0; arow < 2; arow++) {
            self_transpose_case(s[0], s[2], arow, true, 1.f, 0.f, g);
            self_transpose_case(s[0], s[2], arow, false, 1.5f, -0.25f, g);
        }
    sums_case(1000, 40, 777, g);
    sums_case(300, 300, 500, g);
    quality_case(100003, g);
    printf("{\"cases\": %d, \"errors\": %d}\n", case_no, errors);
    return errors ? 1 : 0;
}
