// Device matrix.  Drop-in for DeviceMatrix<Type> (reference src/memory/DeviceMatrix.h:38-228).
// The float specialisation adds Multiply and MultiplyBySelfTranspose; the reference
// calls legacy cublasSgemm there (:204, :226), this calls the library's own MFMA
// GEMM (gpumlib_nmf_sgemm) on the null stream with the same operand semantics:
// a row-major operand is passed as transposed, C is column-major.
#ifndef GPUMLIB_DEVICE_MATRIX_H
#define GPUMLIB_DEVICE_MATRIX_H

#include <cassert>
#include <climits>

#include "../../../include/gpumlib_nmf.h"
#include "../cuda/definitions.h"
#include "DeviceMemoryManager.h"
#include "HostMatrix.h"

namespace GPUMLib {

template <class Type>
class DeviceMatrix : public BaseMatrix<Type> {
  private:
    DeviceMemoryManager<Type> deviceMem;

  public:
    explicit DeviceMatrix(StoringOrder order = RowMajor) : BaseMatrix<Type>(deviceMem, order) {}
    DeviceMatrix(size_t rows, size_t columns, StoringOrder order = RowMajor) : BaseMatrix<Type>(deviceMem, order) {
        this->ResizeWithoutPreservingData(rows, columns);
    }
    DeviceMatrix(const DeviceMatrix<Type>& other) : BaseMatrix<Type>(deviceMem) { this->AssignDeviceMatrix(other); }
    DeviceMatrix(const HostMatrix<Type>& other) : BaseMatrix<Type>(deviceMem) { this->AssignHostMatrix(other); }
    DeviceMatrix(DeviceMatrix<Type>&& other) : BaseMatrix<Type>(deviceMem) { this->TransferOwnerShipFrom(other); }

    DeviceMatrix<Type>& operator=(const HostMatrix<Type>& other) {
        this->AssignHostMatrix(other);
        return *this;
    }
    DeviceMatrix<Type>& operator=(const DeviceMatrix<Type>& other) {
        if (this != &other) this->AssignDeviceMatrix(other);
        return *this;
    }
    DeviceMatrix<Type>& operator=(DeviceMatrix<Type>&& other) {
        this->TransferOwnerShipFrom(other);
        return *this;
    }

    // A transposed copy (through the host, as in the reference).
    DeviceMatrix<Type> Transpose() {
        HostMatrix<Type> t(*this);
        t.ReplaceByTranspose();
        return DeviceMatrix<Type>(t);
    }
};

template <>
class DeviceMatrix<cudafloat> : public BaseMatrix<cudafloat> {
  private:
    DeviceMemoryManager<cudafloat> deviceMem;

    static char Op(const BaseMatrix<cudafloat>& x) { return x.IsRowMajor() ? 'T' : 'N'; }
    static int Ld(const BaseMatrix<cudafloat>& x) { return (int)(x.IsRowMajor() ? x.Columns() : x.Rows()); }

  public:
    explicit DeviceMatrix(StoringOrder order = RowMajor) : BaseMatrix<cudafloat>(deviceMem, order) {}
    DeviceMatrix(size_t rows, size_t columns, StoringOrder order = RowMajor) : BaseMatrix<cudafloat>(deviceMem, order) {
        this->ResizeWithoutPreservingData(rows, columns);
    }
    DeviceMatrix(const DeviceMatrix<cudafloat>& other) : BaseMatrix<cudafloat>(deviceMem) {
        this->AssignDeviceMatrix(other);
    }
    DeviceMatrix(const HostMatrix<cudafloat>& other) : BaseMatrix<cudafloat>(deviceMem) {
        this->AssignHostMatrix(other);
    }
    DeviceMatrix(DeviceMatrix<cudafloat>&& other) : BaseMatrix<cudafloat>(deviceMem) {
        this->TransferOwnerShipFrom(other);
    }

    DeviceMatrix<cudafloat>& operator=(const HostMatrix<cudafloat>& other) {
        this->AssignHostMatrix(other);
        return *this;
    }
    DeviceMatrix<cudafloat>& operator=(const DeviceMatrix<cudafloat>& other) {
        if (this != &other) this->AssignDeviceMatrix(other);
        return *this;
    }
    DeviceMatrix<cudafloat>& operator=(DeviceMatrix<cudafloat>&& other) {
        this->TransferOwnerShipFrom(other);
        return *this;
    }

    DeviceMatrix<cudafloat> Transpose() {
        HostMatrix<cudafloat> t(*this);
        t.ReplaceByTranspose();
        return DeviceMatrix<cudafloat>(t);
    }

    // C = alpha * this * this^T + beta * C (C is rows x rows, made column-major).
    void MultiplyBySelfTranspose(DeviceMatrix<cudafloat>& C, cudafloat alpha = 1, cudafloat beta = 0) {
        assert(C.rows <= INT_MAX && C.rows == rows && C.columns == rows);
        if (C.IsRowMajor()) {
            assert(beta == 0);
            C.storingOrder = ColumnMajor;
        }
        const char opA = IsRowMajor() ? 'T' : 'N', opB = IsRowMajor() ? 'N' : 'T';
        (void)gpumlib_nmf_sgemm(opA, opB, (int)C.rows, (int)C.columns, (int)columns, alpha, Pointer(), Ld(*this),
                                Pointer(), Ld(*this), beta, C.Pointer(), (int)C.rows, nullptr);
    }

    // C = alpha * A * B + beta * C (C made column-major).
    static void Multiply(DeviceMatrix<cudafloat>& A, DeviceMatrix<cudafloat>& B, DeviceMatrix<cudafloat>& C,
                         cudafloat alpha = 1, cudafloat beta = 0) {
        assert(A.columns <= INT_MAX && C.rows <= INT_MAX && C.columns <= INT_MAX);
        assert(A.columns == B.rows && C.rows == A.rows && C.columns == B.columns);
        if (C.IsRowMajor()) {
            assert(beta == 0);
            C.storingOrder = ColumnMajor;
        }
        (void)gpumlib_nmf_sgemm(Op(A), Op(B), (int)C.rows, (int)C.columns, (int)A.columns, alpha, A.Pointer(), Ld(A),
                                B.Pointer(), Ld(B), beta, C.Pointer(), (int)C.rows, nullptr);
    }
};

}  // namespace GPUMLib

#endif
