// Host + device 1-D array pair.  Drop-in for CudaArray<Type> (reference
// src/memory/CudaArray.h:34-199): a HostArray and a DeviceArray of the same length;
// element access reads and writes the host copy; UpdateDevice / UpdateHost copy one
// side onto the other (the two may differ in between, as in the reference).
#ifndef GPUMLIB_CUDA_ARRAY_H
#define GPUMLIB_CUDA_ARRAY_H

#include "DeviceArray.h"
#include "HostArray.h"

namespace GPUMLib {

template <class Type>
class CudaArray {
  private:
    HostArray<Type> h;
    DeviceArray<Type> d;

  public:
    CudaArray() {}
    explicit CudaArray(size_t count) : h(count), d(count) {}
    CudaArray(const DeviceArray<Type>& other) : h(other), d(other) {}
    CudaArray(const HostArray<Type>& other) : h(other), d(other) {}
    CudaArray(const CudaArray<Type>& other) : h(other.h), d(other.d) {}
    CudaArray(CudaArray<Type>&& other) {
        h.TransferOwnerShipFrom(other.h);
        d.TransferOwnerShipFrom(other.d);
    }

    CudaArray<Type>& operator=(const CudaArray<Type>& other) {
        if (this != &other) {
            h = other.h;
            d = other.d;
        }
        return *this;
    }
    // Both sides become copies of the device array.
    CudaArray<Type>& operator=(const DeviceArray<Type>& other) {
        d = other;
        h = d;
        return *this;
    }
    // Both sides become copies of the host array.
    CudaArray<Type>& operator=(const HostArray<Type>& other) {
        h = other;
        d = h;
        return *this;
    }
    CudaArray<Type>& operator=(CudaArray<Type>&& other) {
        if (this != &other) {
            h.TransferOwnerShipFrom(other.h);
            d.TransferOwnerShipFrom(other.d);
        }
        return *this;
    }

    Type& operator[](size_t element) { return h[element]; }
    Type operator[](size_t element) const { return h[element]; }

    size_t Length() const { return d.Length(); }
    Type* HostPointer() const { return h.Pointer(); }
    Type* DevicePointer() const { return d.Pointer(); }

    // Returns the smaller of the two new lengths (0 if either allocation failed).
    size_t ResizeWithoutPreservingData(size_t count) {
        const size_t hc = h.ResizeWithoutPreservingData(count);
        const size_t dc = d.ResizeWithoutPreservingData(count);
        return hc < dc ? hc : dc;
    }

    void UpdateDevice() { d = h; }
    void UpdateHost() { h = d; }

    DeviceArray<Type>& GetDeviceArray() { return d; }
    HostArray<Type>& GetHostArray() { return h; }

    void Dispose() {
        d.Dispose();
        h.Dispose();
    }
    // Keeps the data on the host only (copied back first).
    void DisposeDevice() {
        UpdateHost();
        d.Dispose();
    }
    // Keeps the data on the device only (copied over first).
    void DisposeHost() {
        UpdateDevice();
        h.Dispose();
    }
};

}  // namespace GPUMLib

#endif
