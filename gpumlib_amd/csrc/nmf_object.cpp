// C-ABI handle over the C++ class API (NMF_MultiplicativeEuclidianDistance,
// NMF_MultiplicativeDivergence, NMF_AdditiveEuclidian, NMF_AdditiveDivergence), for
// callers that bind C rather than C++.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/gpumlib_nmf.h"
#include "../src/algorithms/nmf/nmf_additive_divergence.h"
#include "../src/algorithms/nmf/nmf_additive_euclidian.h"
#include "../src/algorithms/nmf/nmf_multiplicative_divergence.h"
#include "../src/algorithms/nmf/nmf_multiplicative_euclidian.h"

using namespace GPUMLib;

struct gpumlib_nmf_object {
    NMF* algo;
    int n, m, r;
};

static HostMatrix<cudafloat> host_matrix(const float* data, int rows, int cols) {
    HostMatrix<cudafloat> x((size_t)rows, (size_t)cols, ColumnMajor);
    if (x.Elements() == (size_t)rows * cols) std::memcpy(x.Pointer(), data, sizeof(float) * (size_t)rows * cols);
    return x;
}

static hipError_t copy_out(const HostMatrix<cudafloat>& x, float* out, size_t count) {
    if (x.Elements() != count) return hipErrorUnknown;
    std::memcpy(out, x.Pointer(), sizeof(float) * count);
    return hipSuccess;
}

extern "C" {

hipError_t gpumlib_nmf_object_create(int rule, int precision, int n, int m, int r, const float* host_V,
                                     const float* host_W, const float* host_H, gpumlib_nmf_handle* out) {
    if (!out || !host_V || !host_W || !host_H || n < 1 || m < 1 || r < 1) return hipErrorInvalidValue;
    // the C++ classes run the reference's FP32 arithmetic
    if (precision != GPUMLIB_NMF_FP32 || gpumlib_nmf_workspace_bytes(rule, precision, n, m, r) == 0)
        return hipErrorInvalidValue;
    *out = nullptr;
    HostMatrix<cudafloat> v = host_matrix(host_V, n, m);
    HostMatrix<cudafloat> w = host_matrix(host_W, n, r);
    HostMatrix<cudafloat> h = host_matrix(host_H, r, m);
    if (v.Elements() == 0 || w.Elements() == 0 || h.Elements() == 0) return hipErrorOutOfMemory;
    gpumlib_nmf_object* obj = new (std::nothrow) gpumlib_nmf_object;
    if (!obj) return hipErrorOutOfMemory;
    obj->n = n;
    obj->m = m;
    obj->r = r;
    switch (rule) {
        case GPUMLIB_NMF_EUCLIDEAN: obj->algo = new (std::nothrow) NMF_MultiplicativeEuclidianDistance(v, w, h); break;
        case GPUMLIB_NMF_KL: obj->algo = new (std::nothrow) NMF_MultiplicativeDivergence(v, w, h); break;
        case GPUMLIB_NMF_ADDITIVE_KL: obj->algo = new (std::nothrow) NMF_AdditiveDivergence(v, w, h); break;
        default: {
            NMF_AdditiveEuclidian* a = new (std::nothrow) NMF_AdditiveEuclidian(v, w, h);
            if (a) a->SetUpdateH(rule == GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_WITH_H);
            obj->algo = a;
        }
    }
    if (!obj->algo) {
        delete obj;
        return hipErrorOutOfMemory;
    }
    if (obj->algo->LastError() != hipSuccess) {
        hipError_t e = obj->algo->LastError();
        delete obj->algo;
        delete obj;
        return e;
    }
    *out = obj;
    return hipSuccess;
}

hipError_t gpumlib_nmf_object_set_epsilon(gpumlib_nmf_handle h, float eps) {
    if (!h) return hipErrorInvalidValue;
    h->algo->SetEpsilon(eps);
    return hipSuccess;
}

hipError_t gpumlib_nmf_object_set_quality_interval(gpumlib_nmf_handle h, int interval) {
    if (!h) return hipErrorInvalidValue;
    h->algo->SetQualityInterval(interval);
    return hipSuccess;
}

hipError_t gpumlib_nmf_object_do_iteration(gpumlib_nmf_handle h, int update_w) {
    if (!h) return hipErrorInvalidValue;
    h->algo->DoIteration(update_w != 0);
    return h->algo->LastError();
}

hipError_t gpumlib_nmf_object_do_iterations(gpumlib_nmf_handle h, int count, int update_w) {
    if (!h || count < 0) return hipErrorInvalidValue;
    for (int i = 0; i < count; i++) h->algo->DoIteration(update_w != 0);
    return h->algo->LastError();
}

float gpumlib_nmf_object_quality_improvement(gpumlib_nmf_handle h) { return h ? h->algo->QualityImprovement() : 0.f; }

hipError_t gpumlib_nmf_object_get_w(gpumlib_nmf_handle h, float* host_W) {
    if (!h || !host_W) return hipErrorInvalidValue;
    return copy_out(h->algo->GetW(), host_W, (size_t)h->n * h->r);
}

hipError_t gpumlib_nmf_object_get_h(gpumlib_nmf_handle h, float* host_H) {
    if (!h || !host_H) return hipErrorInvalidValue;
    return copy_out(h->algo->GetH(), host_H, (size_t)h->r * h->m);
}

hipError_t gpumlib_nmf_object_get_wh(gpumlib_nmf_handle h, float* host_WH) {
    if (!h || !host_WH) return hipErrorInvalidValue;
    return copy_out(h->algo->GetWH(), host_WH, (size_t)h->n * h->m);
}

hipError_t gpumlib_nmf_object_objective(gpumlib_nmf_handle h, int objective, double* host_value) {
    if (!h || !host_value) return hipErrorInvalidValue;
    *host_value = h->algo->Objective(objective);
    return h->algo->LastError();
}

void gpumlib_nmf_object_destroy(gpumlib_nmf_handle h) {
    if (!h) return;
    delete h->algo;
    delete h;
}

}  // extern "C"
