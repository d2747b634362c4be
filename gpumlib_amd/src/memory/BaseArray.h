// 1-D container base.  Drop-in for BaseArray<Type> (reference src/memory/BaseArray.h:33-70).
#ifndef GPUMLIB_BASE_ARRAY_H
#define GPUMLIB_BASE_ARRAY_H

#include "MemoryManager.h"

namespace GPUMLib {

template <class Type>
class BaseArray {
  private:
    MemoryManager<Type>* mem;

  protected:
    explicit BaseArray(MemoryManager<Type>& storage) : mem(&storage) {}

  public:
    void Dispose() { mem->Dispose(); }
    size_t Length() const { return mem->Size(); }
    Type* Pointer() const { return mem->Pointer(); }
    size_t ResizeWithoutPreservingData(size_t count) { return mem->ResizeWithoutPreservingData(count); }
};

}  // namespace GPUMLib

#endif
