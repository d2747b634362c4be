// Device 3-D matrix.  Drop-in for DeviceMatrix3D<Type> (reference src/memory/DeviceMatrix3D.h:33-101).
#ifndef GPUMLIB_DEVICE_MATRIX_3D_H
#define GPUMLIB_DEVICE_MATRIX_3D_H

#include "DeviceMemoryManager.h"
#include "HostMatrix3D.h"

namespace GPUMLib {

template <class Type>
class DeviceMatrix3D : public BaseMatrix3D<Type> {
  private:
    DeviceMemoryManager<Type> deviceMem;

  public:
    DeviceMatrix3D() : BaseMatrix3D<Type>(deviceMem) {}
    DeviceMatrix3D(size_t x, size_t y, size_t z) : BaseMatrix3D<Type>(deviceMem) {
        this->ResizeWithoutPreservingData(x, y, z);
    }
    DeviceMatrix3D(const DeviceMatrix3D<Type>& other) : BaseMatrix3D<Type>(deviceMem) {
        this->AssignDeviceMatrix(other);
    }
    DeviceMatrix3D(const HostMatrix3D<Type>& other) : BaseMatrix3D<Type>(deviceMem) { this->AssignHostMatrix(other); }
    DeviceMatrix3D(DeviceMatrix3D<Type>&& other) : BaseMatrix3D<Type>(deviceMem) { this->TransferOwnerShipFrom(other); }

    DeviceMatrix3D<Type>& operator=(const HostMatrix3D<Type>& other) {
        this->AssignHostMatrix(other);
        return *this;
    }
    DeviceMatrix3D<Type>& operator=(const DeviceMatrix3D<Type>& other) {
        if (this != &other) this->AssignDeviceMatrix(other);
        return *this;
    }
    DeviceMatrix3D<Type>& operator=(DeviceMatrix3D<Type>&& other) {
        this->TransferOwnerShipFrom(other);
        return *this;
    }
};

}  // namespace GPUMLib

#endif
