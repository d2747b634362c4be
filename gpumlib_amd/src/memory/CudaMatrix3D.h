// Host + device 3-D matrix pair.  Drop-in for CudaMatrix3D<Type> (reference
// src/memory/CudaMatrix3D.h:33-234).
#ifndef GPUMLIB_CUDA_MATRIX_3D_H
#define GPUMLIB_CUDA_MATRIX_3D_H

#include "DeviceMatrix3D.h"
#include "HostMatrix3D.h"

namespace GPUMLib {

template <class Type>
class CudaMatrix3D {
  private:
    HostMatrix3D<Type> h;
    DeviceMatrix3D<Type> d;

  public:
    CudaMatrix3D() {}
    CudaMatrix3D(size_t x, size_t y, size_t z) : h(x, y, z), d(x, y, z) {}
    CudaMatrix3D(const DeviceMatrix3D<Type>& other) : h(other), d(other) {}
    CudaMatrix3D(const HostMatrix3D<Type>& other) : h(other), d(other) {}
    CudaMatrix3D(const CudaMatrix3D<Type>& other) : h(other.h), d(other.d) {}
    CudaMatrix3D(CudaMatrix3D<Type>&& other) {
        h.TransferOwnerShipFrom(other.h);
        d.TransferOwnerShipFrom(other.d);
    }

    CudaMatrix3D<Type>& operator=(const CudaMatrix3D<Type>& other) {
        if (this != &other) {
            h = other.h;
            d = other.d;
        }
        return *this;
    }
    CudaMatrix3D<Type>& operator=(const DeviceMatrix3D<Type>& other) {
        d = other;
        h = d;
        return *this;
    }
    CudaMatrix3D<Type>& operator=(const HostMatrix3D<Type>& other) {
        h = other;
        d = h;
        return *this;
    }
    CudaMatrix3D<Type>& operator=(CudaMatrix3D<Type>&& other) {
        if (this != &other) {
            h.TransferOwnerShipFrom(other.h);
            d.TransferOwnerShipFrom(other.d);
        }
        return *this;
    }

    void UpdateDevice() { d = h; }
    void UpdateHost() { h = d; }

    Type* HostPointer() const { return h.Pointer(); }
    Type* DevicePointer() const { return d.Pointer(); }
    DeviceMatrix3D<Type>& Get3DDeviceMatrix() { return d; }
    HostMatrix3D<Type>& Get3DHostMatrix() { return h; }

    Type& operator()(size_t x, size_t y, size_t z) { return h(x, y, z); }
    Type operator()(size_t x, size_t y, size_t z) const { return h(x, y, z); }

    size_t DimX() const { return h.DimX(); }
    size_t DimY() const { return h.DimY(); }
    size_t DimZ() const { return h.DimZ(); }
    // The reference returns int here (CudaMatrix3D.h:198); kept for source compatibility.
    int Elements() const { return (int)h.Elements(); }

    size_t ResizeWithoutPreservingData(size_t x, size_t y, size_t z) {
        const size_t hc = h.ResizeWithoutPreservingData(x, y, z);
        const size_t dc = d.ResizeWithoutPreservingData(x, y, z);
        return hc < dc ? hc : dc;
    }

    void Dispose() {
        d.Dispose();
        h.Dispose();
    }
    void DisposeDevice() {
        UpdateHost();
        d.Dispose();
    }
    void DisposeHost() {
        UpdateDevice();
        h.Dispose();
    }
};

}  // namespace GPUMLib

#endif
