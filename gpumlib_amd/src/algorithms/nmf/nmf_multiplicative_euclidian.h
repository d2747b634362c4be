// Multiplicative-update NMF under the Euclidean distance.  Drop-in for
// GPUMLib::NMF_MultiplicativeEuclidianDistance (reference
// src/algorithms/nmf/nmf_multiplicative_euclidian.h:34-70, .cpp:29-67):
//   H <- H .* (W^T V) ./ (W^T W H + eps), then (if updateW)
//   W <- W .* (V H^T) ./ (W H H^T + eps) with the new H.
#ifndef GPUMLIB_NMF_MULTIPLICATIVE_EUCLIDEAN_H
#define GPUMLIB_NMF_MULTIPLICATIVE_EUCLIDEAN_H

#include "../../common/Utilities.h"
#include "nmf_base.h"

namespace GPUMLib {

class NMF_MultiplicativeEuclidianDistance : public NMF {
  public:
    //! V (n x m, column-major) with W (n x r) and H (r x m) drawn at random.
    NMF_MultiplicativeEuclidianDistance(HostMatrix<cudafloat>& v, int r) : NMF(v, r, GPUMLIB_NMF_EUCLIDEAN) {}
    //! V (n x m), W (n x r), H (r x m), all column-major.
    NMF_MultiplicativeEuclidianDistance(HostMatrix<cudafloat>& v, HostMatrix<cudafloat>& w, HostMatrix<cudafloat>& h)
        : NMF(v, w, h, GPUMLIB_NMF_EUCLIDEAN) {}

    void DoIteration(bool updateW = true) override { Iterate(updateW); }
};

}  // namespace GPUMLib

#endif  // GPUMLIB_NMF_MULTIPLICATIVE_EUCLIDEAN_H
