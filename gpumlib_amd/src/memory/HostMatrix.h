// Host matrix.  Drop-in for HostMatrix<Type> (reference src/memory/HostMatrix.h:37-132).
#ifndef GPUMLIB_HOST_MATRIX_H
#define GPUMLIB_HOST_MATRIX_H

#include <cassert>

#include "BaseMatrix.h"
#include "HostMemoryManager.h"

namespace GPUMLib {

template <class Type>
class DeviceMatrix;

template <class Type>
class HostMatrix : public BaseMatrix<Type> {
  private:
    HostMemoryManager<Type> hostMem;

    size_t Index(size_t row, size_t column) const {
        assert(row < this->rows && column < this->columns);
        return this->IsRowMajor() ? row * this->columns + column : column * this->rows + row;
    }

  public:
    explicit HostMatrix(StoringOrder order = RowMajor) : BaseMatrix<Type>(hostMem, order) {}
    HostMatrix(size_t rows, size_t columns, StoringOrder order = RowMajor) : BaseMatrix<Type>(hostMem, order) {
        this->ResizeWithoutPreservingData(rows, columns);
    }
    HostMatrix(const HostMatrix<Type>& other) : BaseMatrix<Type>(hostMem) { this->AssignHostMatrix(other); }
    HostMatrix(const DeviceMatrix<Type>& other) : BaseMatrix<Type>(hostMem) { this->AssignDeviceMatrix(other); }
    HostMatrix(HostMatrix<Type>&& other) : BaseMatrix<Type>(hostMem) { this->TransferOwnerShipFrom(other); }

    HostMatrix<Type>& operator=(const HostMatrix<Type>& other) {
        if (this != &other) this->AssignHostMatrix(other);
        return *this;
    }
    HostMatrix<Type>& operator=(const DeviceMatrix<Type>& other) {
        this->AssignDeviceMatrix(other);
        return *this;
    }
    HostMatrix<Type>& operator=(HostMatrix<Type>&& other) {
        this->TransferOwnerShipFrom(other);
        return *this;
    }

    // A transposed copy (metadata flip of a deep copy).
    HostMatrix<Type> Transpose() {
        HostMatrix<Type> t(*this);
        t.ReplaceByTranspose();
        return t;
    }

    Type& operator()(size_t row, size_t column) { return this->Pointer()[Index(row, column)]; }
    Type operator()(size_t row, size_t column) const { return this->Pointer()[Index(row, column)]; }
};

}  // namespace GPUMLib

#endif
