// Storage policy base of the GPUMLib containers.  Drop-in for MemoryManager<Type>
// (reference src/memory/MemoryManager.h:33-110): owns a buffer of `size` elements;
// allocation failure leaves size 0 (the reference's contract, no exceptions).
#ifndef GPUMLIB_MEMORY_MANAGER_H
#define GPUMLIB_MEMORY_MANAGER_H

#include <cstddef>

namespace GPUMLib {

template <class Type>
class MemoryManager {
  protected:
    Type* data;
    size_t size;

    void Reset() {
        data = nullptr;
        size = 0;
    }
    MemoryManager() { Reset(); }

  public:
    virtual ~MemoryManager() {}
    virtual void Alloc(size_t size) = 0;
    virtual void Dispose() = 0;
    virtual void CopyDataFromDevice(Type* source, size_t count) = 0;
    virtual void CopyDataFromHost(Type* source, size_t count) = 0;

    size_t Size() const { return size; }
    size_t SizeInBytes() const { return size * sizeof(Type); }
    Type* Pointer() const { return data; }

    // Reallocates only when the element count changes; returns the new count (0 on failure).
    size_t ResizeWithoutPreservingData(size_t count) {
        if (count != size) {
            Dispose();
            Alloc(count);
        }
        return size;
    }

    void TransferOwnerShipFrom(MemoryManager<Type>& other) {
        if (this == &other) return;
        Dispose();
        data = other.data;
        size = other.size;
        other.Reset();
    }
};

}  // namespace GPUMLib

#endif
