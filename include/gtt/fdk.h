// FDK reconstruction drivers:
//   FDK_reconstruct_3d (xray/FDK_reconstruct_3d.cpp:128-161) -- FBP mult_MH,
//     forward projection of the result, least-squares rescale
//     f *= <Af,p>/<Af,Af>.
//   FDK_reconstruct_4d (xray/FDK_reconstruct_4d.cpp:100-170) -- McKinnon-Bates
//     4-D FDK: f = FBP over the union of the bins, s = <Af,p>/<Af,Af>,
//     image_b = (FBP_b(p - s A f) + f) * s for every bin b.
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
// filter of compute_default_frequency_filter).  One FBP volume per bin
// ((nx, ny, nz, nbins) when nbins > 1); the projections are only read.
// Runs device-resident (the reference's hoCu operator streams through host
// memory and forms the inner products on the host); same results.
FDKResult FDK_reconstruct_3d(std::shared_ptr<CBCT_acquisition> acquisition,
                             std::shared_ptr<CBCT_binning> binning,
                             const std::vector<size_t> &is_dims, floatd3 is_dims_in_mm,
                             std::shared_ptr<std::vector<float>> frequency_filter = nullptr,
                             float samples_per_pixel = 1.5f);

struct FDK4DResult {
  std::shared_ptr<hoCuNDArray<float>> image;    // (nx, ny, nz, nbins)
  std::shared_ptr<hoCuNDArray<float>> fdk_3d;   // (nx, ny, nz): the FBP over the union of the bins
  double scale = 0.0;
};

// is_dims: the 3-D image dimensions; the result has one volume per bin.
// Views in no bin take no part.  An empty bin's volume is f * s.  Angles are
// shifted in place as by the operator setup.
FDK4DResult FDK_reconstruct_4d(std::shared_ptr<CBCT_acquisition> acquisition,
                               std::shared_ptr<CBCT_binning> binning,
                               const std::vector<size_t> &is_dims, floatd3 is_dims_in_mm,
                               std::shared_ptr<std::vector<float>> frequency_filter = nullptr,
                               float samples_per_pixel = 1.5f);

}  // namespace gtt
