// Additive-update NMF under the (Kullback-Leibler) divergence.  Drop-in for
// GPUMLib::NMF_AdditiveDivergence (reference src/algorithms/nmf/nmf_additive_divergence.h:35-68,
// .cpp:30-51; kernels src/cuda/nmf/nmf_kernels_additive_divergence.cu:57-193):
//   H(a,u) <- max(0, H + (H / max(sum_i W(i,a), eps)) sum_i W(i,a) (V(i,u)/((WH)(i,u) + eps) - 1)),
//   then (if updateW) with WH recomputed from the new H
//   W(i,a) <- max(0, W + (W / max(sum_u H(a,u), eps)) sum_u H(a,u) (V(i,u)/((WH)(i,u) + eps) - 1)).
// WH is never stored: the fused KL passes form it tile by tile, as for
// NMF_MultiplicativeDivergence.
#ifndef GPUMLIB_NMF_ADDITIVE_DIVERGENCE_H
#define GPUMLIB_NMF_ADDITIVE_DIVERGENCE_H

#include "../../common/Utilities.h"
#include "../../memory/DeviceArray.h"
#include "nmf_base.h"

namespace GPUMLib {

class NMF_AdditiveDivergence : public NMF {
  public:
    //! V (n x m, column-major) with W (n x r) and H (r x m) drawn at random.
    NMF_AdditiveDivergence(HostMatrix<cudafloat>& v, int r) : NMF(v, r, GPUMLIB_NMF_ADDITIVE_KL) {}
    //! V (n x m), W (n x r), H (r x m), all column-major.
    NMF_AdditiveDivergence(HostMatrix<cudafloat>& v, HostMatrix<cudafloat>& w, HostMatrix<cudafloat>& h)
        : NMF(v, w, h, GPUMLIB_NMF_ADDITIVE_KL) {}

    void DoIteration(bool updateW = true) override { Iterate(updateW); }
};

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_ADDITIVE_DIVERGENCE_H
