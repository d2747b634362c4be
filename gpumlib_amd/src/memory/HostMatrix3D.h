// Host 3-D matrix.  Drop-in for HostMatrix3D<Type> (reference src/memory/HostMatrix3D.h:33-125).
#ifndef GPUMLIB_HOST_MATRIX_3D_H
#define GPUMLIB_HOST_MATRIX_3D_H

#include <cassert>

#include "BaseMatrix3D.h"
#include "HostMemoryManager.h"

namespace GPUMLib {

template <class Type>
class DeviceMatrix3D;

template <class Type>
class HostMatrix3D : public BaseMatrix3D<Type> {
  private:
    HostMemoryManager<Type> hostMem;

    size_t Index(size_t x, size_t y, size_t z) const {
        assert(x < this->dimX && y < this->dimY && z < this->dimZ);
        return (z * this->dimY + y) * this->dimX + x;
    }

  public:
    HostMatrix3D() : BaseMatrix3D<Type>(hostMem) {}
    HostMatrix3D(size_t x, size_t y, size_t z) : BaseMatrix3D<Type>(hostMem) {
        this->ResizeWithoutPreservingData(x, y, z);
    }
    HostMatrix3D(const HostMatrix3D<Type>& other) : BaseMatrix3D<Type>(hostMem) { this->AssignHostMatrix(other); }
    HostMatrix3D(const DeviceMatrix3D<Type>& other) : BaseMatrix3D<Type>(hostMem) { this->AssignDeviceMatrix(other); }
    HostMatrix3D(HostMatrix3D<Type>&& other) : BaseMatrix3D<Type>(hostMem) { this->TransferOwnerShipFrom(other); }

    HostMatrix3D<Type>& operator=(const HostMatrix3D<Type>& other) {
        if (this != &other) this->AssignHostMatrix(other);
        return *this;
    }
    HostMatrix3D<Type>& operator=(const DeviceMatrix3D<Type>& other) {
        this->AssignDeviceMatrix(other);
        return *this;
    }
    HostMatrix3D<Type>& operator=(HostMatrix3D<Type>&& other) {
        this->TransferOwnerShipFrom(other);
        return *this;
    }

    Type& operator()(size_t x, size_t y, size_t z) { return this->Pointer()[Index(x, y, z)]; }
    Type operator()(size_t x, size_t y, size_t z) const { return this->Pointer()[Index(x, y, z)]; }
};

}  // namespace GPUMLib

#endif
