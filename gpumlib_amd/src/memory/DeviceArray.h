// Device 1-D array.  Drop-in for DeviceArray<Type> (reference src/memory/DeviceArray.h:33-115).
#ifndef GPUMLIB_DEVICE_ARRAY_H
#define GPUMLIB_DEVICE_ARRAY_H

#include "DeviceMemoryManager.h"
#include "HostArray.h"

namespace GPUMLib {

template <class Type>
class DeviceArray : public BaseArray<Type> {
  private:
    DeviceMemoryManager<Type> deviceMem;

  public:
    DeviceArray() : BaseArray<Type>(deviceMem) {}
    explicit DeviceArray(size_t count) : BaseArray<Type>(deviceMem) { deviceMem.Alloc(count); }
    DeviceArray(const HostArray<Type>& other) : BaseArray<Type>(deviceMem) {
        deviceMem.CopyDataFromHost(other.Pointer(), other.Length());
    }
    DeviceArray(const DeviceArray<Type>& other) : BaseArray<Type>(deviceMem) {
        deviceMem.CopyDataFromDevice(other.Pointer(), other.Length());
    }
    DeviceArray(const Type* host, size_t count) : BaseArray<Type>(deviceMem) {
        deviceMem.CopyDataFromHost(const_cast<Type*>(host), count);
    }
    DeviceArray(DeviceArray<Type>&& other) : BaseArray<Type>(deviceMem) {
        deviceMem.TransferOwnerShipFrom(other.deviceMem);
    }

    DeviceArray<Type>& operator=(const DeviceArray<Type>& other) {
        if (this != &other) deviceMem.CopyDataFromDevice(other.Pointer(), other.Length());
        return *this;
    }
    DeviceArray<Type>& operator=(const HostArray<Type>& other) {
        deviceMem.CopyDataFromHost(other.Pointer(), other.Length());
        return *this;
    }
    DeviceArray<Type>& operator=(DeviceArray<Type>&& other) {
        deviceMem.TransferOwnerShipFrom(other.deviceMem);
        return *this;
    }
    void TransferOwnerShipFrom(DeviceArray<Type>& other) {
        if (this != &other) deviceMem.TransferOwnerShipFrom(other.deviceMem);
    }
};

}  // namespace GPUMLib

#endif
