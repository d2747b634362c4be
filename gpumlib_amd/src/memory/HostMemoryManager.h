// Host storage (new[]/delete[]).  Drop-in for HostMemoryManager<Type>
// (reference src/memory/HostMemoryManager.h:37-72).  Device-to-host copies are
// blocking hipMemcpy, which is the read-back synchronisation point of GetW/GetH.
#ifndef GPUMLIB_HOST_MEMORY_MANAGER_H
#define GPUMLIB_HOST_MEMORY_MANAGER_H

#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "MemoryManager.h"

namespace GPUMLib {

template <class Type>
class HostMemoryManager : public MemoryManager<Type> {
  public:
    void Alloc(size_t count) override {
        this->Reset();
        if (count == 0) return;
        this->data = new (std::nothrow) Type[count];
        if (this->data) this->size = count;
    }
    void Dispose() override {
        if (this->data) delete[] this->data;
        this->Reset();
    }
    void CopyDataFromDevice(Type* source, size_t count) override {
        if (this->ResizeWithoutPreservingData(count) > 0)
            (void)hipMemcpy(this->data, source, this->SizeInBytes(), hipMemcpyDeviceToHost);
    }
    void CopyDataFromHost(Type* source, size_t count) override {
        if (this->ResizeWithoutPreservingData(count) > 0) std::memcpy(this->data, source, this->SizeInBytes());
    }
    ~HostMemoryManager() { Dispose(); }
};

}  // namespace GPUMLib

#endif
