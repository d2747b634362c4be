// Device (HBM) storage.  Drop-in for DeviceMemoryManager<Type>
// (reference src/memory/DeviceMemoryManager.h:34-68): hipMalloc/hipFree, blocking
// host-to-device and device-to-device copies.
#ifndef GPUMLIB_DEVICE_MEMORY_MANAGER_H
#define GPUMLIB_DEVICE_MEMORY_MANAGER_H

#include <hip/hip_runtime.h>

#include "MemoryManager.h"

namespace GPUMLib {

template <class Type>
class DeviceMemoryManager : public MemoryManager<Type> {
  public:
    void Alloc(size_t count) override {
        this->Reset();
        if (count == 0) return;
        void* p = nullptr;
        if (hipMalloc(&p, count * sizeof(Type)) == hipSuccess) {
            this->data = static_cast<Type*>(p);
            this->size = count;
        }
    }
    void Dispose() override {
        if (this->data) (void)hipFree(this->data);
        this->Reset();
    }
    void CopyDataFromDevice(Type* source, size_t count) override {
        if (this->ResizeWithoutPreservingData(count) > 0)
            (void)hipMemcpy(this->data, source, this->SizeInBytes(), hipMemcpyDeviceToDevice);
    }
    void CopyDataFromHost(Type* source, size_t count) override {
        if (this->ResizeWithoutPreservingData(count) > 0)
            (void)hipMemcpy(this->data, source, this->SizeInBytes(), hipMemcpyHostToDevice);
    }
    ~DeviceMemoryManager() { Dispose(); }
};

}  // namespace GPUMLib

#endif
