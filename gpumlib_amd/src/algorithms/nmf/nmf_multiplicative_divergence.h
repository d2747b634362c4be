// Multiplicative-update NMF under the (Kullback-Leibler) divergence.  Drop-in for
// GPUMLib::NMF_MultiplicativeDivergence (reference
// src/algorithms/nmf/nmf_multiplicative_divergence.h:35-63, .cpp:29-50):
//   H(a,u) <- H(a,u) sum_i W(i,a) V(i,u)/((WH)(i,u) + eps) / max(sum_i W(i,a), eps),
//   then (if updateW) with WH recomputed from the new H
//   W(i,a) <- W(i,a) sum_u H(a,u) V(i,u)/((WH)(i,u) + eps) / max(sum_u H(a,u), eps).
#ifndef GPUMLIB_NMF_MULTIPLICATIVE_DIVERGENCE_H
#define GPUMLIB_NMF_MULTIPLICATIVE_DIVERGENCE_H

#include "../../common/Utilities.h"
#include "../../memory/DeviceArray.h"
#include "nmf_base.h"

namespace GPUMLib {

class NMF_MultiplicativeDivergence : public NMF {
  public:
    //! V (n x m, column-major) with W (n x r) and H (r x m) drawn at random.
    NMF_MultiplicativeDivergence(HostMatrix<cudafloat>& v, int r) : NMF(v, r, GPUMLIB_NMF_KL) {}
    //! V (n x m), W (n x r), H (r x m), all column-major.
    NMF_MultiplicativeDivergence(HostMatrix<cudafloat>& v, HostMatrix<cudafloat>& w, HostMatrix<cudafloat>& h)
        : NMF(v, w, h, GPUMLIB_NMF_KL) {}

    void DoIteration(bool updateW = true) override { Iterate(updateW); }
};

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_MULTIPLICATIVE_DIVERGENCE_H
