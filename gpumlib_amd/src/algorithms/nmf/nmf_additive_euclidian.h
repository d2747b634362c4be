// Additive-update NMF under the Euclidean distance.  Drop-in for
// GPUMLib::NMF_AdditiveEuclidian (reference src/algorithms/nmf/nmf_additive_euclidian.h:34-70,
// .cpp:31-56), with its behaviour:
//   H is left unchanged: the reference's H step computes deltaH = W^T V and
//   deltaH2 = (W^T W) H, then passes (H, deltaH, deltaH2) to the multiplicative
//   launcher, whose arguments are (numerator, denominator, updated matrix)
//   (.cpp:42), so only the scratch deltaH2 is rewritten;
//   then (if updateW) W <- max(0, W + (W / WHHt) (VHt - WHHt)), VHt = V H^T,
//   WHHt = W (H H^T)  (kernel nmf_kernels.cu:31-39, no eps).
// SetUpdateH(true) (no reference counterpart) also applies the additive rule to H,
//   H <- max(0, H + (H / WtWH) (WtV - WtWH)), the step the scratch buffers were for.
#ifndef GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_H
#define GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_H

#include "../../common/Utilities.h"
#include "nmf_base.h"

namespace GPUMLib {

class NMF_AdditiveEuclidian : public NMF {
  public:
    //! V (n x m, column-major) with W (n x r) and H (r x m) drawn at random.
    NMF_AdditiveEuclidian(HostMatrix<cudafloat>& v, int r) : NMF(v, r, GPUMLIB_NMF_ADDITIVE_EUCLIDEAN) {}
    //! V (n x m), W (n x r), H (r x m), all column-major.
    NMF_AdditiveEuclidian(HostMatrix<cudafloat>& v, HostMatrix<cudafloat>& w, HostMatrix<cudafloat>& h)
        : NMF(v, w, h, GPUMLIB_NMF_ADDITIVE_EUCLIDEAN) {}

    void DoIteration(bool updateW = true) override { Iterate(updateW); }

    // ---- addition (no reference counterpart) ----
    //! true: H is updated by the additive rule too; false (default): H is kept, as in the reference.
    void SetUpdateH(bool on) { SetRule(on ? GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_WITH_H : GPUMLIB_NMF_ADDITIVE_EUCLIDEAN); }
};

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_ADDITIVE_EUCLIDEAN_H
