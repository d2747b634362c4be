// Base class of the NMF algorithms (multiplicative and additive updates).  Drop-in for
// GPUMLib::NMF (reference src/algorithms/nmf/nmf_base.h:45-151).
//
// Given a non-negative V (n x m, column-major), find W (n x r) and H (r x m) with
// V ~ W H.  V, W and H live in HBM for the life of the object; one iteration is
// a single call into the fused C-ABI (gpumlib_nmf_iterate) on the object's own
// stream, using a workspace sized once at construction.
//
// Quality: the reference computes sum V log10(WH + eps) - WH on a side stream
// whenever that stream is idle (nmf_base.h:70-89), so its cadence depends on
// timing.  Here it is computed, in FP64 and fused into the iteration's pass over
// V, every QualityInterval() iterations (default 1), and read back one quality
// point later; the bookkeeping of QualityImprovement() is the reference's
// (nmf_base.h:72-80: no improvement figure before the third reading).
#ifndef GPUMLIB_NMF_BASE_H
#define GPUMLIB_NMF_BASE_H

#include <hip/hip_runtime.h>

#include <cassert>
#include <cstdlib>

#include "../../common/CudaStreams.h"
#include "../../common/Utilities.h"
#include "../../cuda/definitions.h"
#include "../../cuda/nmf/nmf_kernels.h"
#include "../../memory/DeviceAccessibleVariable.h"
#include "../../memory/DeviceArray.h"
#include "../../memory/DeviceMatrix.h"

#define SIZE_BLOCKS_NMF (512)

namespace GPUMLib {

class NMF {
  private:
    CudaStream stream;
    int timesQualityWasObtained;
    cudafloat lastValueQuality;
    cudafloat deltaImprovement;
    DeviceAccessibleVariable<double> vquality;
    DeviceArray<double> d_quality;
    DeviceArray<unsigned char> workspace;
    hipEvent_t qualityReady;
    bool qualityPending;
    int rule;
    int precision;
    cudafloat epsilon;
    int qualityInterval;
    long iterationsDone;
    hipError_t lastError;

    // W and H drawn as rand()/RAND_MAX + eps, as in the reference (nmf_base.h:56-61).
    static void Randomize(HostMatrix<cudafloat>& x) {
        cudafloat* p = x.Pointer();
        for (size_t e = 0; e < x.Elements(); e++)
            p[e] = ((cudafloat)rand()) / RAND_MAX + SMALL_VALUE_TO_ADD_DENOMINATOR;
    }

    void Setup(int ruleKind) {
        rule = ruleKind;
        precision = GPUMLIB_NMF_FP32;
        epsilon = SMALL_VALUE_TO_ADD_DENOMINATOR;
        qualityInterval = 1;
        iterationsDone = 0;
        timesQualityWasObtained = 0;
        deltaImprovement = CUDA_VALUE(1.0);
        lastValueQuality = CUDA_VALUE(-1.0);
        vquality.Value() = 0.0;
        qualityPending = false;
        lastError = hipSuccess;
        if (hipEventCreateWithFlags(&qualityReady, hipEventDisableTiming) != hipSuccess) qualityReady = nullptr;
        const size_t bytes =
            gpumlib_nmf_workspace_bytes(rule, precision, (int)V.Rows(), (int)V.Columns(), (int)W.Columns());
        // any rank plans (ranks above 128 run the unfused GEMM sequence); 0 means an empty or
        // out-of-range shape, which the reference rejects by assert as well
        assert(bytes > 0);
        if (bytes == 0) {
            RecordError(hipErrorInvalidValue);
            return;
        }
        workspace.ResizeWithoutPreservingData(bytes);
        if (workspace.Length() != bytes) RecordError(hipErrorOutOfMemory);
    }

    void RecordError(hipError_t e) {
        if (e != hipSuccess && lastError == hipSuccess) lastError = e;
    }

    // Reads the quality computed at the previous quality point (0 before the first).
    void ConsumeQuality() {
        if (qualityPending) {
            (void)hipEventSynchronize(qualityReady);
            qualityPending = false;
        }
        const cudafloat quality = (cudafloat)vquality.Value();
        if (timesQualityWasObtained++ > 2) {
            deltaImprovement = quality - lastValueQuality;
        } else {
            timesQualityWasObtained++;
        }
        lastValueQuality = quality;
    }

  protected:
    DeviceMatrix<cudafloat> V;
    DeviceMatrix<cudafloat> W;
    DeviceMatrix<cudafloat> H;
    DeviceMatrix<cudafloat> WH;  // allocated on the first GetWH()

    NMF(HostMatrix<cudafloat>& v, int r, int ruleKind) : d_quality(1), V(v) {
        assert(!v.IsRowMajor());
        HostMatrix<cudafloat> aux(v.Rows(), r, ColumnMajor);
        Randomize(aux);
        W = aux;
        aux.ResizeWithoutPreservingData(r, v.Columns());
        Randomize(aux);
        H = aux;
        Setup(ruleKind);
    }

    NMF(HostMatrix<cudafloat>& v, HostMatrix<cudafloat>& w, HostMatrix<cudafloat>& h, int ruleKind)
        : d_quality(1), V(v), W(w), H(h) {
        assert(!v.IsRowMajor());
        assert(!w.IsRowMajor());
        assert(!h.IsRowMajor());
        assert(v.Rows() == w.Rows() && v.Columns() == h.Columns());
        assert(w.Columns() == h.Rows());
        Setup(ruleKind);
    }

    // Switches between update rules that run the same passes (hence the same workspace).
    void SetRule(int ruleKind) { rule = ruleKind; }

    // One iteration of the object's update rule (H, then W unless updateW is false).
    void Iterate(bool updateW) {
        if (workspace.Length() == 0) return;  // construction failed: LastError() says why
        const bool qualityPoint = qualityInterval > 0 && (iterationsDone % qualityInterval) == 0;
        if (qualityPoint) ConsumeQuality();
        hipError_t e = gpumlib_nmf_iterate(rule, precision, (int)V.Rows(), (int)V.Columns(), (int)W.Columns(),
                                           V.Pointer(), W.Pointer(), H.Pointer(), epsilon, 1, updateW ? 1 : 0,
                                           qualityPoint ? GPUMLIB_NMF_OBJ_QUALITY : GPUMLIB_NMF_OBJ_NONE,
                                           qualityPoint ? d_quality.Pointer() : nullptr, workspace.Pointer(),
                                           workspace.Length(), stream);
        if (e == hipSuccess && qualityPoint) {
            vquality.UpdateValueAsync(d_quality.Pointer(), stream);
            if (qualityReady && hipEventRecord(qualityReady, stream) == hipSuccess) qualityPending = true;
        }
        RecordError(e);
        iterationsDone++;
    }

  public:
    virtual ~NMF() {
        (void)hipStreamSynchronize(stream);
        if (qualityReady) (void)hipEventDestroy(qualityReady);
    }

    //! Gets the W matrix (blocking read-back)
    HostMatrix<cudafloat> GetW() {
        (void)hipStreamSynchronize(stream);
        return HostMatrix<cudafloat>(W);
    }

    //! Gets the H matrix (blocking read-back)
    HostMatrix<cudafloat> GetH() {
        (void)hipStreamSynchronize(stream);
        return HostMatrix<cudafloat>(H);
    }

    //! Gets the approximation W H of V (blocking read-back)
    HostMatrix<cudafloat> GetWH() {
        (void)hipStreamSynchronize(stream);
        if (WH.Rows() != V.Rows() || WH.Columns() != V.Columns() || WH.IsRowMajor())
            WH = DeviceMatrix<cudafloat>(V.Rows(), V.Columns(), ColumnMajor);
        // DeviceMatrix<cudafloat>::Multiply(W, H, WH) with W, H and WH column-major, on the
        // object's stream and with its status recorded
        RecordError(gpumlib_nmf_sgemm('N', 'N', (int)WH.Rows(), (int)WH.Columns(), (int)W.Columns(), CUDA_VALUE(1.0),
                                      W.Pointer(), (int)W.Rows(), H.Pointer(), (int)H.Rows(), CUDA_VALUE(0.0),
                                      WH.Pointer(), (int)WH.Rows(), stream));
        RecordError(hipStreamSynchronize(stream));
        return HostMatrix<cudafloat>(WH);
    }

    //! Does one algorithm iteration, adjusting H and (if updateW) W.
    virtual void DoIteration(bool updateW = true) = 0;

    //! Quality improvement measured between the last two quality readings.
    cudafloat QualityImprovement() const { return deltaImprovement; }

    // ---- additions (no reference counterpart) ----
    //! eps of the update ratios (reference: SMALL_VALUE_TO_ADD_DENOMINATOR = 1e-8).
    void SetEpsilon(cudafloat eps) { epsilon = eps; }
    cudafloat Epsilon() const { return epsilon; }
    //! Iterations between quality computations; 0 disables them.
    void SetQualityInterval(int interval) { qualityInterval = interval < 0 ? 0 : interval; }
    int QualityInterval() const { return qualityInterval; }
    //! FP64 objective (GPUMLIB_NMF_OBJ_*) of the current W, H; blocking.
    double Objective(int kind) {
        double value = 0.0;
        hipError_t e = gpumlib_nmf_objective(rule, precision, (int)V.Rows(), (int)V.Columns(), (int)W.Columns(),
                                             V.Pointer(), W.Pointer(), H.Pointer(), epsilon, kind,
                                             d_quality.Pointer(), workspace.Pointer(), workspace.Length(), stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&value, d_quality.Pointer(), sizeof(double), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        RecordError(e);
        return value;
    }
    //! First error reported by the library for this object (hipSuccess if none).
    hipError_t LastError() const { return lastError; }
    hipStream_t Stream() { return stream; }
    long Iterations() const { return iterationsDone; }
};

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_BASE_H
