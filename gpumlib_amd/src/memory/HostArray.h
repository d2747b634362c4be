// Host 1-D array.  Drop-in for HostArray<Type> (reference src/memory/HostArray.h:33-125).
#ifndef GPUMLIB_HOST_ARRAY_H
#define GPUMLIB_HOST_ARRAY_H

#include <cassert>

#include "BaseArray.h"
#include "HostMemoryManager.h"

namespace GPUMLib {

template <class Type>
class DeviceArray;

template <class Type>
class HostArray : public BaseArray<Type> {
  private:
    HostMemoryManager<Type> hostMem;

  public:
    HostArray() : BaseArray<Type>(hostMem) {}
    explicit HostArray(size_t count) : BaseArray<Type>(hostMem) { hostMem.Alloc(count); }
    HostArray(const DeviceArray<Type>& other) : BaseArray<Type>(hostMem) {
        hostMem.CopyDataFromDevice(other.Pointer(), other.Length());
    }
    HostArray(const HostArray<Type>& other) : BaseArray<Type>(hostMem) {
        hostMem.CopyDataFromHost(other.Pointer(), other.Length());
    }
    HostArray(HostArray<Type>&& other) : BaseArray<Type>(hostMem) { hostMem.TransferOwnerShipFrom(other.hostMem); }

    HostArray<Type>& operator=(const DeviceArray<Type>& other) {
        hostMem.CopyDataFromDevice(other.Pointer(), other.Length());
        return *this;
    }
    HostArray<Type>& operator=(const HostArray<Type>& other) {
        if (this != &other) hostMem.CopyDataFromHost(other.Pointer(), other.Length());
        return *this;
    }
    HostArray<Type>& operator=(HostArray<Type>&& other) {
        hostMem.TransferOwnerShipFrom(other.hostMem);
        return *this;
    }
    void TransferOwnerShipFrom(HostArray<Type>& other) {
        if (this != &other) hostMem.TransferOwnerShipFrom(other.hostMem);
    }

    Type& operator[](size_t element) {
        assert(element < this->Length());
        return this->Pointer()[element];
    }
    Type operator[](size_t element) const {
        assert(element < this->Length());
        return this->Pointer()[element];
    }
};

}  // namespace GPUMLib

#endif
