// RAII HIP stream.  Drop-in for class CudaStream (reference src/common/CudaStreams.h:33-52):
// created on construction, destroyed on destruction, converts to the native stream handle.
#ifndef GPUMLib_CudaStreams_h
#define GPUMLib_CudaStreams_h

#include <hip/hip_runtime.h>

namespace GPUMLib {

class CudaStream {
  public:
    CudaStream() {
        if (hipStreamCreate(&handle) != hipSuccess) handle = nullptr;
    }
    ~CudaStream() {
        if (handle) (void)hipStreamDestroy(handle);
    }
    CudaStream(const CudaStream&) = delete;
    CudaStream& operator=(const CudaStream&) = delete;
    operator hipStream_t() const { return handle; }

  private:
    hipStream_t handle;
};

}  // namespace GPUMLib

#endif
