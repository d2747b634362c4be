// 3-D container base.  Drop-in for BaseMatrix3D<Type> (reference
// src/memory/BaseMatrix3D.h:30-145): dimX x dimY x dimZ elements, x fastest
// (element (x, y, z) at z*dimY*dimX + y*dimX + x), deep-copy assignment from host or
// device 3-D matrices, ownership transfer.
#ifndef GPUMLIB_BASE_MATRIX_3D_H
#define GPUMLIB_BASE_MATRIX_3D_H

#include "MemoryManager.h"

namespace GPUMLib {

template <class Type>
class BaseMatrix3D {
  protected:
    size_t dimX, dimY, dimZ;

  private:
    MemoryManager<Type>* mem;

    void AdoptShape(const BaseMatrix3D<Type>& other) {
        if (mem->Size() == other.Elements()) {
            dimX = other.dimX;
            dimY = other.dimY;
            dimZ = other.dimZ;
        } else {
            dimX = dimY = dimZ = 0;
        }
    }

  protected:
    explicit BaseMatrix3D(MemoryManager<Type>& storage) : dimX(0), dimY(0), dimZ(0), mem(&storage) {}

    void AssignHostMatrix(const BaseMatrix3D<Type>& other) {
        mem->CopyDataFromHost(other.Pointer(), other.Elements());
        AdoptShape(other);
    }
    void AssignDeviceMatrix(const BaseMatrix3D<Type>& other) {
        mem->CopyDataFromDevice(other.Pointer(), other.Elements());
        AdoptShape(other);
    }

  public:
    // Returns the new element count; 0 (and a 0 x 0 x 0 shape) if allocation failed.
    size_t ResizeWithoutPreservingData(size_t x, size_t y, size_t z) {
        const size_t count = x * y * z;
        if (mem->ResizeWithoutPreservingData(count) == count) {
            dimX = x;
            dimY = y;
            dimZ = z;
        } else {
            dimX = dimY = dimZ = 0;
        }
        return Elements();
    }

    void Dispose() {
        mem->Dispose();
        dimX = dimY = dimZ = 0;
    }

    size_t DimX() const { return dimX; }
    size_t DimY() const { return dimY; }
    size_t DimZ() const { return dimZ; }
    Type* Pointer() const { return mem->Pointer(); }
    size_t Elements() const { return mem->Size(); }

    void TransferOwnerShipFrom(BaseMatrix3D<Type>& other) {
        if (this == &other) return;
        mem->TransferOwnerShipFrom(*other.mem);
        dimX = other.dimX;
        dimY = other.dimY;
        dimZ = other.dimZ;
        other.dimX = other.dimY = other.dimZ = 0;
    }
};

}  // namespace GPUMLib

#endif
