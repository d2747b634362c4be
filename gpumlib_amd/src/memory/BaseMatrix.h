// 2-D container base.  Drop-in for BaseMatrix<Type> and StoringOrder (reference
// src/memory/BaseMatrix.h:33-159): row- or column-major storage, metadata-only
// transposition, deep-copy assignment from host or device matrices.
#ifndef GPUMLIB_BASE_MATRIX_H
#define GPUMLIB_BASE_MATRIX_H

#include "MemoryManager.h"

namespace GPUMLib {

typedef enum { RowMajor, ColumnMajor } StoringOrder;

template <class Type>
class BaseMatrix {
  protected:
    StoringOrder storingOrder;
    size_t rows;
    size_t columns;

  private:
    MemoryManager<Type>* mem;

    // After a copy, take the other matrix's shape if the copy succeeded.
    void AdoptShape(const BaseMatrix<Type>& other) {
        if (mem->Size() == other.Elements()) {
            rows = other.rows;
            columns = other.columns;
            storingOrder = other.storingOrder;
        } else {
            rows = columns = 0;
        }
    }

  protected:
    BaseMatrix(MemoryManager<Type>& storage, StoringOrder order = RowMajor)
        : storingOrder(order), rows(0), columns(0), mem(&storage) {}

    void AssignHostMatrix(const BaseMatrix<Type>& other) {
        mem->CopyDataFromHost(other.Pointer(), other.Elements());
        AdoptShape(other);
    }
    void AssignDeviceMatrix(const BaseMatrix<Type>& other) {
        mem->CopyDataFromDevice(other.Pointer(), other.Elements());
        AdoptShape(other);
    }

  public:
    // Returns the new element count; 0 (and a 0 x 0 shape) if allocation failed.
    size_t ResizeWithoutPreservingData(size_t newRows, size_t newColumns) {
        const size_t count = newRows * newColumns;
        if (mem->ResizeWithoutPreservingData(count) == count) {
            rows = newRows;
            columns = newColumns;
        } else {
            rows = columns = 0;
        }
        return Elements();
    }

    void Dispose() {
        mem->Dispose();
        rows = columns = 0;
    }

    size_t Rows() const { return rows; }
    size_t Columns() const { return columns; }
    Type* Pointer() const { return mem->Pointer(); }
    size_t Elements() const { return mem->Size(); }
    bool IsRowMajor() const { return storingOrder == RowMajor; }
    StoringOrder Order() const { return storingOrder; }

    // Transpose by swapping the shape and the storing order; the data does not move.
    void ReplaceByTranspose() {
        const size_t r = rows;
        rows = columns;
        columns = r;
        storingOrder = IsRowMajor() ? ColumnMajor : RowMajor;
    }

    void TransferOwnerShipFrom(BaseMatrix<Type>& other) {
        if (this == &other) return;
        mem->TransferOwnerShipFrom(*other.mem);
        storingOrder = other.storingOrder;
        rows = other.rows;
        columns = other.columns;
        other.rows = other.columns = 0;
    }
};

}  // namespace GPUMLib

#endif
