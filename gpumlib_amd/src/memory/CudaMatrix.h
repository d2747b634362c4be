// Host + device matrix pair.  Drop-in for CudaMatrix<Type> (reference
// src/memory/CudaMatrix.h:33-247): a HostMatrix and a DeviceMatrix of the same
// shape and storing order; element access goes to the host copy; UpdateDevice /
// UpdateHost copy one side onto the other.
#ifndef GPUMLIB_CUDA_MATRIX_H
#define GPUMLIB_CUDA_MATRIX_H

#include "DeviceMatrix.h"
#include "HostMatrix.h"

namespace GPUMLib {

template <class Type>
class CudaMatrix {
  private:
    HostMatrix<Type> h;
    DeviceMatrix<Type> d;

  public:
    explicit CudaMatrix(StoringOrder order = RowMajor) : h(order), d(order) {}
    CudaMatrix(size_t rows, size_t columns, StoringOrder order = RowMajor)
        : h(rows, columns, order), d(rows, columns, order) {}
    CudaMatrix(const DeviceMatrix<Type>& other) : h(other), d(other) {}
    CudaMatrix(const HostMatrix<Type>& other) : h(other), d(other) {}
    CudaMatrix(const CudaMatrix<Type>& other) : h(other.h), d(other.d) {}
    CudaMatrix(CudaMatrix<Type>&& other) {
        h.TransferOwnerShipFrom(other.h);
        d.TransferOwnerShipFrom(other.d);
    }

    CudaMatrix<Type>& operator=(const CudaMatrix<Type>& other) {
        if (this != &other) {
            h = other.h;
            d = other.d;
        }
        return *this;
    }
    CudaMatrix<Type>& operator=(const DeviceMatrix<Type>& other) {
        d = other;
        h = d;
        return *this;
    }
    CudaMatrix<Type>& operator=(const HostMatrix<Type>& other) {
        h = other;
        d = h;
        return *this;
    }
    CudaMatrix<Type>& operator=(CudaMatrix<Type>&& other) {
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
    DeviceMatrix<Type>& GetDeviceMatrix() { return d; }
    HostMatrix<Type>& GetHostMatrix() { return h; }

    Type& operator()(size_t row, size_t column) { return h(row, column); }
    Type operator()(size_t row, size_t column) const { return h(row, column); }

    size_t Rows() const { return h.Rows(); }
    size_t Columns() const { return h.Columns(); }
    size_t Elements() const { return h.Elements(); }
    bool IsRowMajor() const { return h.IsRowMajor(); }

    // Metadata-only transpose of both sides.
    void ReplaceByTranspose() {
        h.ReplaceByTranspose();
        d.ReplaceByTranspose();
    }
    // A transposed copy built from the host side (both sides hold it).
    CudaMatrix<Type> Transpose() { return CudaMatrix<Type>(h.Transpose()); }

    size_t ResizeWithoutPreservingData(size_t rows, size_t columns) {
        const size_t hc = h.ResizeWithoutPreservingData(rows, columns);
        const size_t dc = d.ResizeWithoutPreservingData(rows, columns);
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
