// Self-check of the reduction framework with the semantics of GPUMLib's reduction
// example (reference src/examples/reduction/reductionapp.cpp:30-109):
//
//   reduction_app <array size> <number tests> [<random generator>]
//
// x[i] = -1 + 2 rand() / RAND_MAX (srand(seed); seed = time(0) when not given),
// held in a CudaArray; Sum, Average, MinIndex and MaxIndex run <number tests> times
// on the device; each result is printed as "name: value ok" or
// "name: value error (expected e)" against a host loop, then the elapsed time.
// Differences from the reference app: the host expectations are accumulated in
// double and sums are compared with a tolerance relative to their magnitude
// (1e-5 * max(1, |expected|)), indices exactly (the library reports the lowest index
// of equal extremes, as the host loop does); the exit status is 1 if any check failed.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <chrono>
#include <cmath>
#include <ctime>

#include "../src/cuda/reduction/reduction.h"
#include "../src/memory/CudaArray.h"

using namespace GPUMLib;

namespace {

constexpr double kTolerance = 1e-5;
int failures = 0;

void Display(const char* name, double value, double expected, double scale) {
    printf("%s: %.9g ", name, value);
    if (std::fabs(value - expected) <= kTolerance * scale) {
        printf("ok\n");
    } else {
        printf("error (expected %.9g)\n", expected);
        failures++;
    }
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc != 3 && argc != 4) {
        printf("usage: reduction_app <array size> <number tests> [<random generator>]\n");
        return 0;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        printf("No HIP device\n");
        return 2;
    }
    const int N = atoi(argv[1]);
    const int tests = atoi(argv[2]);
    const unsigned seed = (argc == 4) ? (unsigned)strtoul(argv[3], nullptr, 10) : (unsigned)time(nullptr);
    if (N < 1 || tests < 1) {
        printf("array size and number of tests must be positive\n");
        return 2;
    }
    srand(seed);
    printf("Random Generator.........: %u\n", seed);

    CudaArray<cudafloat> x(N);
    CudaArray<cudafloat> mn(1), mx(1), sum(1), avg(1);
    CudaArray<int> mnIdx(1), mxIdx(1);
    if (x.Length() != (size_t)N) {
        printf("allocation failed\n");
        return 2;
    }
    for (int i = 0; i < N; i++) x[i] = CUDA_VALUE(-1.0) + (CUDA_VALUE(2.0) * rand()) / RAND_MAX;
    x.UpdateDevice();

    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < tests; t++) {
        Reduction::Sum(x.GetDeviceArray(), sum.GetDeviceArray());
        Reduction::Average(x.GetDeviceArray(), avg.GetDeviceArray());
        Reduction::MinIndex(x.GetDeviceArray(), mn.GetDeviceArray(), mnIdx.GetDeviceArray());
        Reduction::MaxIndex(x.GetDeviceArray(), mx.GetDeviceArray(), mxIdx.GetDeviceArray());
    }
    (void)hipDeviceSynchronize();
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    sum.UpdateHost();
    avg.UpdateHost();
    mn.UpdateHost();
    mx.UpdateHost();
    mnIdx.UpdateHost();
    mxIdx.UpdateHost();

    double hsum = 0.0;
    float hmin = x[0], hmax = x[0];
    int hposmin = 0, hposmax = 0;
    for (int i = 0; i < N; i++) {
        hsum += x[i];
        if (x[i] < hmin) {
            hmin = x[i];
            hposmin = i;
        }
        if (x[i] > hmax) {
            hmax = x[i];
            hposmax = i;
        }
    }
    const double havg = hsum / N;
    Display("sum", sum[0], hsum, std::fmax(1.0, std::fabs(hsum)));
    Display("avg", avg[0], havg, std::fmax(1.0, std::fabs(havg)));
    Display("min", mn[0], hmin, 0.0);
    Display("max", mx[0], hmax, 0.0);
    Display("min index", mnIdx[0], hposmin, 0.0);
    if (mnIdx[0] != hposmin && mnIdx[0] >= 0 && mnIdx[0] < N)
        printf("x[%d] = %.9g; x[%d] = %.9g\n", mnIdx[0], x[mnIdx[0]], hposmin, x[hposmin]);
    Display("max index", mxIdx[0], hposmax, 0.0);
    if (mxIdx[0] != hposmax && mxIdx[0] >= 0 && mxIdx[0] < N)
        printf("x[%d] = %.9g; x[%d] = %.9g\n", mxIdx[0], x[mxIdx[0]], hposmax, x[hposmax]);
    printf("%.6f s\n", seconds);
    return failures ? 1 : 0;
}
