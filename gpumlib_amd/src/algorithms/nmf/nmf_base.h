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
// timing.  Here it is computed (terms in FP32 as there, summed in FP64), fused into
// the iteration's pass over V, every QualityInterval() iterations (default 1), and
// read back one quality point later; the bookkeeping of QualityImprovement() is the
// reference's (nmf_base.h:72-80: no improvement figure before the third reading).
// DoIteration() does not wait for a quality: each one is copied into one of a few
// pinned slots behind its iteration, and the bookkeeping readings are applied in
// order when QualityImprovement() asks for them (or when a slot is needed again),
// so the host keeps the stream fed while the figures stay those of the in-order
// bookkeeping.
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
    static constexpr int kQualitySlots = 4;

    CudaStream stream;
    mutable int timesQualityWasObtained;
    mutable cudafloat lastValueQuality;
    mutable cudafloat deltaImprovement;
    DeviceAccessibleVariable<double> vquality[kQualitySlots];  // quality of point k in slot k % kQualitySlots
    DeviceArray<double> d_quality;                              // the same slots on the device, + one for Objective()
    DeviceArray<unsigned char> workspace;
    hipEvent_t qualityReady[kQualitySlots];
    long qualityPoints;            // quality points reached
    mutable long readingsApplied;  // bookkeeping readings applied (reading k is taken at point k)
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
        qualityPoints = 0;
        readingsApplied = 0;
        lastError = hipSuccess;
        for (int k = 0; k < kQualitySlots; k++) {
            vquality[k].Value() = 0.0;
            if (hipEventCreateWithFlags(&qualityReady[k], hipEventDisableTiming) != hipSuccess) qualityReady[k] = nullptr;
        }
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

    // Applies the bookkeeping readings before reading `end`.  Reading k reads the quality
    // computed at the previous quality point k - 1 (0 at the first point).
    void ApplyReadings(long end) const {
        for (; readingsApplied < end; readingsApplied++) {
            double value = 0.0;
            if (readingsApplied > 0) {
                const int slot = (int)((readingsApplied - 1) % kQualitySlots);
                if (qualityReady[slot])
                    (void)hipEventSynchronize(qualityReady[slot]);
                else
                    (void)hipStreamSynchronize(stream);
                value = vquality[slot].Value();
            }
            const cudafloat quality = (cudafloat)value;
            if (timesQualityWasObtained++ > 2) {
                deltaImprovement = quality - lastValueQuality;
            } else {
                timesQualityWasObtained++;
            }
            lastValueQuality = quality;
        }
    }

  protected:
    DeviceMatrix<cudafloat> V;
    DeviceMatrix<cudafloat> W;
    DeviceMatrix<cudafloat> H;
    DeviceMatrix<cudafloat> WH;  // allocated on the first GetWH()

    NMF(HostMatrix<cudafloat>& v, int r, int ruleKind) : d_quality(kQualitySlots + 1), V(v) {
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
        : d_quality(kQualitySlots + 1), V(v), W(w), H(h) {
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
        const int slot = (int)(qualityPoints % kQualitySlots);
        // the slot's previous value (point qualityPoints - kQualitySlots) is read by the reading after it
        if (qualityPoint) ApplyReadings(qualityPoints - kQualitySlots + 2);
        double* dq = d_quality.Pointer() + slot;
        hipError_t e = gpumlib_nmf_iterate(rule, precision, (int)V.Rows(), (int)V.Columns(), (int)W.Columns(),
                                           V.Pointer(), W.Pointer(), H.Pointer(), epsilon, 1, updateW ? 1 : 0,
                                           qualityPoint ? GPUMLIB_NMF_OBJ_QUALITY : GPUMLIB_NMF_OBJ_NONE,
                                           qualityPoint ? dq : nullptr, workspace.Pointer(), workspace.Length(), stream);
        if (qualityPoint) {
            if (e == hipSuccess) {
                vquality[slot].UpdateValueAsync(dq, stream);
                if (qualityReady[slot]) e = hipEventRecord(qualityReady[slot], stream);
            } else {
                vquality[slot].Value() = 0.0;
            }
            qualityPoints++;
        }
        RecordError(e);
        iterationsDone++;
    }

  public:
    virtual ~NMF() {
        (void)hipStreamSynchronize(stream);
        for (int k = 0; k < kQualitySlots; k++)
            if (qualityReady[k]) (void)hipEventDestroy(qualityReady[k]);
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

    //! Quality improvement measured between the last two quality readings (waits, if it must,
    //! for the quality of the quality point before the last).
    cudafloat QualityImprovement() const {
        ApplyReadings(qualityPoints);
        return deltaImprovement;
    }

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
                                             d_quality.Pointer() + kQualitySlots, workspace.Pointer(),
                                             workspace.Length(), stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&value, d_quality.Pointer() + kQualitySlots, sizeof(double), hipMemcpyDeviceToHost, stream);
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
