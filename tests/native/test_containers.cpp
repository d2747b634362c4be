// Checks of the host + device containers (CudaArray, CudaMatrix, HostMatrix3D,
// DeviceMatrix3D, CudaMatrix3D) against the semantics of the reference's
// src/memory headers: element layout, deep copies, UpdateDevice / UpdateHost round
// trips through HBM, ownership transfer, resizing and the Dispose* variants.
// Prints one line per failed check and a summary; exit status 1 on any failure.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <utility>
#include <vector>

#include "../../gpumlib_amd/src/memory/CudaArray.h"
#include "../../gpumlib_amd/src/memory/CudaMatrix.h"
#include "../../gpumlib_amd/src/memory/CudaMatrix3D.h"

using namespace GPUMLib;

static int checks = 0, failed = 0;
#define CHECK(cond)                                                  \
    do {                                                             \
        checks++;                                                    \
        if (!(cond)) {                                               \
            failed++;                                                \
            printf("FAILED line %d: %s\n", __LINE__, #cond);         \
        }                                                            \
    } while (0)

template <class T>
static std::vector<T> device_contents(const T* p, size_t n) {
    std::vector<T> v(n);
    if (n) (void)hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost);
    return v;
}

static void cuda_array() {
    CudaArray<float> a(1000);
    CHECK(a.Length() == 1000);
    for (int i = 0; i < 1000; i++) a[i] = 0.5f * i;
    a.UpdateDevice();
    std::vector<float> d = device_contents(a.DevicePointer(), 1000);
    CHECK(d[0] == 0.f && d[999] == 499.5f);
    for (int i = 0; i < 1000; i++) a[i] = -1.f;  // host changed, device not
    CHECK(device_contents(a.DevicePointer(), 1)[0] == 0.f);
    a.UpdateHost();
    CHECK(a[10] == 5.f);
    CudaArray<float> b(a);  // deep copy of both sides
    CHECK(b.Length() == 1000 && b.DevicePointer() != a.DevicePointer() && b.HostPointer() != a.HostPointer());
    CHECK(b[998] == 499.f);
    b[0] = 7.f;
    CHECK(a[0] == 0.f);
    CudaArray<float> c(std::move(b));  // ownership moves
    CHECK(c.Length() == 1000 && b.Length() == 0 && c[0] == 7.f);
    HostArray<float> h(3);
    h[0] = 1.f, h[1] = 2.f, h[2] = 3.f;
    c = h;  // both sides become copies of h
    CHECK(c.Length() == 3 && c[2] == 3.f && device_contents(c.DevicePointer(), 3)[1] == 2.f);
    DeviceArray<float> dv(h);
    CudaArray<float> e;
    e = dv;
    CHECK(e.Length() == 3 && e[0] == 1.f);
    CHECK(e.ResizeWithoutPreservingData(50) == 50 && e.Length() == 50);
    e[49] = 4.f;
    e.DisposeHost();  // device keeps the data
    CHECK(e.HostPointer() == nullptr && device_contents(e.DevicePointer(), 50)[49] == 4.f);
    CudaArray<int> f(4);
    for (int i = 0; i < 4; i++) f[i] = i * i;
    f.UpdateDevice();
    for (int i = 0; i < 4; i++) f[i] = 0;
    f.DisposeDevice();  // the device data is copied back to the host first
    CHECK(f.DevicePointer() == nullptr && f.GetHostArray().Length() == 4 && f[3] == 9);
    CudaArray<float> g(8);
    g.Dispose();
    CHECK(g.Length() == 0 && g.HostPointer() == nullptr && g.DevicePointer() == nullptr);
}

static void cuda_matrix() {
    CudaMatrix<float> m(3, 4, ColumnMajor);
    CHECK(m.Rows() == 3 && m.Columns() == 4 && m.Elements() == 12 && !m.IsRowMajor());
    for (size_t r = 0; r < 3; r++)
        for (size_t c = 0; c < 4; c++) m(r, c) = (float)(10 * r + c);
    CHECK(m.HostPointer()[1] == 10.f && m.HostPointer()[3] == 1.f);  // column-major layout
    m.UpdateDevice();
    std::vector<float> d = device_contents(m.DevicePointer(), 12);
    CHECK(d[5] == 21.f);  // (row 2, col 1)
    m.ReplaceByTranspose();  // metadata only
    CHECK(m.Rows() == 4 && m.Columns() == 3 && m.IsRowMajor() && m(1, 2) == 21.f);
    CHECK(m.GetDeviceMatrix().Rows() == 4 && m.GetDeviceMatrix().IsRowMajor());
    CudaMatrix<float> t = m.Transpose();
    CHECK(t.Rows() == 3 && t.Columns() == 4 && t(2, 1) == 21.f);
    CHECK(device_contents(t.DevicePointer(), 12)[5] == 21.f);
    CudaMatrix<float> rm(2, 2);  // RowMajor by default
    CHECK(rm.IsRowMajor());
    rm(0, 1) = 5.f;
    CHECK(rm.HostPointer()[1] == 5.f);
    HostMatrix<float> hm(2, 3, ColumnMajor);
    hm(1, 2) = 9.f;
    CudaMatrix<float> fromHost(hm);
    CHECK(fromHost(1, 2) == 9.f && device_contents(fromHost.DevicePointer(), 6)[5] == 9.f);
    CudaMatrix<float> moved(std::move(fromHost));
    CHECK(moved.Elements() == 6 && fromHost.Elements() == 0);
    CHECK(moved.ResizeWithoutPreservingData(5, 7) == 35 && moved.Rows() == 5 && moved.Columns() == 7);
    moved.Dispose();
    CHECK(moved.Elements() == 0 && moved.Rows() == 0);
}

static void matrices_3d() {
    HostMatrix3D<float> h(4, 3, 2);
    CHECK(h.DimX() == 4 && h.DimY() == 3 && h.DimZ() == 2 && h.Elements() == 24);
    for (size_t z = 0; z < 2; z++)
        for (size_t y = 0; y < 3; y++)
            for (size_t x = 0; x < 4; x++) h(x, y, z) = (float)(100 * z + 10 * y + x);
    CHECK(h.Pointer()[(1 * 3 + 2) * 4 + 3] == 123.f);  // x fastest, then y, then z
    DeviceMatrix3D<float> d(h);
    CHECK(d.DimX() == 4 && d.DimZ() == 2 && device_contents(d.Pointer(), 24)[23] == 123.f);
    HostMatrix3D<float> back(d);
    CHECK(back(2, 1, 1) == 112.f);
    CudaMatrix3D<float> c(h);
    CHECK(c.DimY() == 3 && c.Elements() == 24 && c(3, 2, 0) == 23.f);
    c(0, 0, 0) = -5.f;
    c.UpdateDevice();
    CHECK(device_contents(c.DevicePointer(), 1)[0] == -5.f);
    c(0, 0, 0) = 1.f;
    c.UpdateHost();
    CHECK(c(0, 0, 0) == -5.f);
    CudaMatrix3D<float> c2(c);
    CHECK(c2.DevicePointer() != c.DevicePointer() && c2(3, 2, 1) == 123.f);
    CudaMatrix3D<float> c3(std::move(c2));
    CHECK(c3.Elements() == 24 && c2.Elements() == 0 && c2.DimX() == 0);
    CHECK(c3.Get3DDeviceMatrix().DimZ() == 2 && c3.Get3DHostMatrix().DimY() == 3);
    DeviceMatrix3D<float> d2;
    d2 = h;
    HostMatrix3D<float> h2;
    h2 = d2;
    CHECK(h2.DimX() == 4 && h2(1, 2, 1) == 121.f);
    CHECK(c3.ResizeWithoutPreservingData(2, 2, 2) == 8 && c3.DimX() == 2);
    c3.DisposeHost();
    CHECK(c3.HostPointer() == nullptr && c3.DevicePointer() != nullptr);
    HostMatrix3D<float> h3(h);
    HostMatrix3D<float> h4;
    h4.TransferOwnerShipFrom(h3);
    CHECK(h4.Elements() == 24 && h3.Elements() == 0 && h3.DimX() == 0 && h4(3, 2, 1) == 123.f);
    h4.Dispose();
    CHECK(h4.Elements() == 0 && h4.DimZ() == 0);
}

int main() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        printf("no HIP device\n");
        return 2;
    }
    cuda_array();
    cuda_matrix();
    matrices_3d();
    printf("containers: %d checks, %d failed\n", checks, failed);
    return failed ? 1 : 0;
}
