// Internal to libgtt: acquisition + binning objects from flat C arguments.
#pragma once

#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "cbct.h"
#include "gtt/CBCT_acquisition.h"

namespace gtt {

struct Scan {
  std::shared_ptr<CBCT_acquisition> acq;
  std::shared_ptr<CBCT_binning> binning;
  floatd3 dims_mm;
  std::vector<size_t> vol_dims;
};

// bin_index holds the concatenated view indices of nbins bins with sizes
// bin_sizes; nbins <= 0 selects one bin of all views.  proj (nullable) is
// copied into the acquisition's host projection array, or, with borrow, the
// array wraps the caller's buffer (then nothing may write to it: the subset
// operators are set up without the host permutation).
inline Scan make_scan(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
                      const unsigned *bin_index, const int *bin_sizes, int nbins,
                      const float *proj, bool borrow = false) {
  if (!g || !angles || nviews <= 0) throw std::runtime_error("illegal geometry or angle array provided");
  Scan s;
  auto geom = std::make_shared<CBCT_geometry>();
  geom->set_SAD(g->sad);
  geom->set_SDD(g->sdd);
  geom->set_FOV(floatd2{{g->fov_u, g->fov_v}});
  geom->set_angles(std::vector<float>(angles, angles + nviews));
  std::vector<floatd2> o(nviews, floatd2{{0, 0}});
  if (offsets)
    for (int i = 0; i < nviews; ++i) o[i] = floatd2{{offsets[2 * i], offsets[2 * i + 1]}};
  geom->set_offsets(o);
  std::shared_ptr<hoCuNDArray<float>> p;
  if (proj) {
    const std::vector<size_t> dims{(size_t)g->nu, (size_t)g->nv, (size_t)nviews};
    if (borrow) {
      p = std::make_shared<hoCuNDArray<float>>(dims, const_cast<float *>(proj));
    } else {
      p = std::make_shared<hoCuNDArray<float>>(dims, no_init);
      std::memcpy(p->get_data_ptr(), proj, p->get_number_of_bytes());
    }
  }
  s.acq = std::make_shared<CBCT_acquisition>(p, geom);
  s.binning = std::make_shared<CBCT_binning>();
  if (nbins <= 0) {
    s.binning->set_as_default_3d_bin((unsigned)nviews);
    nbins = 1;
  } else {
    std::vector<std::vector<unsigned int>> bins(nbins);
    size_t k = 0;
    for (int b = 0; b < nbins; ++b)
      for (int i = 0; i < bin_sizes[b]; ++i) bins[b].push_back(bin_index[k++]);
    s.binning = std::make_shared<CBCT_binning>(bins);
  }
  s.dims_mm = floatd3{{g->lx, g->ly, g->lz}};
  s.vol_dims = {(size_t)g->nx, (size_t)g->ny, (size_t)g->nz};
  if (nbins > 1) s.vol_dims.push_back((size_t)nbins);
  return s;
}

}  // namespace gtt
