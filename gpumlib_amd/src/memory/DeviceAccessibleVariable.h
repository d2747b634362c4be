// A pinned host scalar that the device can copy into asynchronously.  Drop-in for
// DeviceAccessibleVariable<Type> (reference src/memory/DeviceAccessibleVariable.h:31-77).
#ifndef GPUMLIB_DEVICE_ACCESSIBLE_VARIABLE_H
#define GPUMLIB_DEVICE_ACCESSIBLE_VARIABLE_H

#include <hip/hip_runtime.h>

namespace GPUMLib {

template <class Type>
class DeviceAccessibleVariable {
  private:
    Type* value;

  public:
    DeviceAccessibleVariable() : value(nullptr) {
        void* p = nullptr;
        if (hipHostMalloc(&p, sizeof(Type), hipHostMallocDefault) == hipSuccess) value = static_cast<Type*>(p);
    }
    explicit DeviceAccessibleVariable(const Type initialValue) : DeviceAccessibleVariable() {
        if (value) *value = initialValue;
    }
    ~DeviceAccessibleVariable() {
        if (value) (void)hipHostFree(value);
    }
    DeviceAccessibleVariable(const DeviceAccessibleVariable&) = delete;
    DeviceAccessibleVariable& operator=(const DeviceAccessibleVariable&) = delete;

    Type& Value() { return *value; }
    const Type& Value() const { return *value; }
    Type* Pointer() { return value; }
    void UpdateValue(Type* deviceValue) { (void)hipMemcpy(value, deviceValue, sizeof(Type), hipMemcpyDeviceToHost); }
    void UpdateValueAsync(Type* deviceValue, hipStream_t stream) {
        (void)hipMemcpyAsync(value, deviceValue, sizeof(Type), hipMemcpyDeviceToHost, stream);
    }
};

}  // namespace GPUMLib

#endif
