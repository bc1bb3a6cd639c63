// FDK reconstruction driver: the computation of FDK_reconstruct_3d
// (xray/FDK_reconstruct_3d.cpp:128-161) -- FBP mult_MH, forward projection of
// the result, least-squares rescale f *= <Af,p>/<Af,Af>.
#pragma once

#include <memory>
#include <vector>

#include "gtt/operators.h"

namespace gtt {

struct FDKResult {
  std::shared_ptr<hoCuNDArray<float>> image;
  double scale = 0.0;
};

// acquisition angles are shifted in place by the operator setup, as in the
// reference.  frequency_filter: nu+1 one-sided taps (nullptr = the default
// filter of compute_default_frequency_filter).
FDKResult FDK_reconstruct_3d(std::shared_ptr<CBCT_acquisition> acquisition,
                             std::shared_ptr<CBCT_binning> binning,
                             const std::vector<size_t> &is_dims, floatd3 is_dims_in_mm,
                             std::shared_ptr<std::vector<float>> frequency_filter = nullptr,
                             float samples_per_pixel = 1.5f);

}  // namespace gtt
