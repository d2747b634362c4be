// gpumlib_nmf_iterate_noalloc: the fused iteration with its workspace in a
// file-scope device array instead of caller- or library-allocated memory, for
// callers that cannot allocate device memory (code under graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The array lives in this code object: every GPU the library runs on holds its own
// copy of kScratchBytes for the life of the process.  Its device address is looked
// up once per device (hipGetSymbolAddress with that device current) and cached; host
// code never hands the array's own name to a kernel (in host code that name is a
// host-side stand-in, not a device address).
//
// Every call on a GPU shares that GPU's array and leaves its contents behind; the
// iteration itself clears what must start at zero (the gram arrival counters, by a
// fill kernel) and rebuilds the narrow operand copies before the first pass, as
// gpumlib_nmf_iterate does with a fresh workspace.  Calls that overlap on one GPU
// (two streams, two host threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gpumlib_nmf.h"

namespace {

constexpr size_t kScratchBytes = (size_t)GPUMLIB_NMF_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` on each device (null: not looked up yet)

// Device address of this device's copy of the array, looked up on the first call with
// the device current.
hipError_t device_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" hipError_t gpumlib_nmf_iterate_noalloc(int rule, int precision, int n, int m, int r, const void* V,
                                                  float* W, float* H, float eps, int iterations, int update_w,
                                                  int objective, double* d_objective, hipStream_t stream) {
    const size_t need = gpumlib_nmf_workspace_bytes(rule, precision, n, m, r);
    if (need == 0) return hipErrorInvalidValue;     // not a problem the fused iteration plans
    if (need > kScratchBytes) return hipErrorOutOfMemory;  // larger than the static workspace
    void* ws = nullptr;
    const hipError_t e = device_scratch(&ws);
    if (e != hipSuccess) return e;
    return gpumlib_nmf_iterate(rule, precision, n, m, r, V, W, H, eps, iterations, update_w, objective, d_objective,
                               ws, kScratchBytes, stream);
}
