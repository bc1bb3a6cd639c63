// C entry points of libgtt (include/gtt_c.h).
#include "gtt_c.h"

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "gtt/fdk.h"
#include "gtt/operators.h"
#include "gtt/solvers.h"
#include "scan.h"

namespace gtt {
int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return 0;
  return d;
}
}  // namespace gtt

using namespace gtt;

namespace {

thread_local std::string g_err;

// Runs an entry point; device blocks the call cached for reuse are freed on
// the way out.
template <class F>
int guarded(F &&f) {
  int rc = 0;
  try {
    f();
  } catch (const std::exception &e) {
    g_err = e.what();
    rc = -1;
  }
  trim_device_cache();
  return rc;
}

// CBSubsetOperator, or CBSubsetWeightOperator when weights (host, the
// projections' shape) are given, set up on the scan (cuCBOS_reconstruct.cpp:481-503).
// Unweighted scans may borrow the caller's projections: they stay in
// acquisition order on the host (see subset_data).
std::shared_ptr<CBSubsetOperator> make_subset_operator(int nsubsets, const Scan &s, const float *weights) {
  if (!weights) {
    auto E = std::make_shared<CBSubsetOperator>(nsubsets);
    E->setup(s.acq, s.binning, s.dims_mm, false);
    return E;
  }
  auto proj = s.acq->get_projections();
  if (!proj) throw std::runtime_error("Weight dimensions must match that of the projection data");
  auto w = std::make_shared<hoCuNDArray<float>>(proj->dims());
  std::memcpy(w->get_data_ptr(), weights, w->get_number_of_bytes());
  auto E = std::make_shared<CBSubsetWeightOperator>(nsubsets);
  E->setup(s.acq, s.binning, s.dims_mm, w);
  return E;
}

// The solver's data in subset order on the device: the weighted operator
// permuted (and weighted) the host array at setup; otherwise the host array
// is still in acquisition order and is permuted on the device.
cuNDArray<float> subset_data(const CBSubsetOperator &E, const Scan &s, const float *weights) {
  if (weights) return cuNDArray<float>(*s.acq->get_projections());
  return upload_in_subset_order(*s.acq->get_projections(), E.get_permutation());
}

}  // namespace

struct gtt_subset_op_s {
  std::shared_ptr<CBSubsetOperator> op;
};

struct gtt_cu_op_s {
  std::shared_ptr<cuConebeamProjectionOperator> op;
  std::vector<size_t> dom;          // nx, ny, nz[, nbins]
  std::vector<size_t> codom;        // nu, nv, views in the bins
};

struct gtt_ho_op_s {
  std::shared_ptr<hoCuConebeamProjectionOperator> op;
  std::vector<size_t> dom;
  std::vector<size_t> codom;        // nu, nv, nviews (acquisition order)
};

namespace {
size_t views_in_bins(const CBCT_binning &b) {
  size_t n = 0;
  for (unsigned i = 0; i < b.get_number_of_bins(); ++i) n += b.bin(i).size();
  return n;
}
}  // namespace

extern "C" {

const char *gtt_last_error(void) { return g_err.c_str(); }

int gtt_set_default_storage(int dtype) {
  if (dtype != CBCT_F32 && dtype != CBCT_BF16) {
    g_err = "gtt_set_default_storage: storage must be CBCT_F32 or CBCT_BF16";
    return -1;
  }
  set_default_storage(dtype);
  return 0;
}

int gtt_set_autotune(int on) {
  set_default_autotune(on != 0);
  return 0;
}

int gtt_set_graphs(int on) {
  set_default_graphs(on != 0);
  return 0;
}

long long gtt_graph_launches(void) { return graph_launches(); }

int gtt_fdk_reconstruct_3d(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
                           const unsigned *bin_index, const int *bin_sizes, int nbins,
                           const float *projections, const float *filter, float spp,
                           float *volume_out, double *scale_out) {
  return guarded([&] {
    if (!projections || !volume_out) throw std::runtime_error("illegal array pointer provided");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections, true);
    std::shared_ptr<std::vector<float>> F;
    if (filter) F = std::make_shared<std::vector<float>>(filter, filter + g->nu + 1);
    FDKResult r = FDK_reconstruct_3d(s.acq, s.binning, s.vol_dims, s.dims_mm, F, spp);
    std::memcpy(volume_out, r.image->get_data_ptr(), r.image->get_number_of_bytes());
    if (scale_out) *scale_out = r.scale;
  });
}

int gtt_fdk_reconstruct_4d(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
                           const unsigned *bin_index, const int *bin_sizes, int nbins,
                           const float *projections, const float *filter, float spp,
                           float *volume_out, float *fdk3d_out, double *scale_out) {
  return guarded([&] {
    if (!projections || !volume_out) throw std::runtime_error("illegal array pointer provided");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections, true);
    std::shared_ptr<std::vector<float>> F;
    if (filter) F = std::make_shared<std::vector<float>>(filter, filter + g->nu + 1);
    const std::vector<size_t> dims3(s.vol_dims.begin(), s.vol_dims.begin() + 3);
    FDK4DResult r = FDK_reconstruct_4d(s.acq, s.binning, dims3, s.dims_mm, F, spp > 0 ? spp : 1.5f);
    std::memcpy(volume_out, r.image->get_data_ptr(), r.image->get_number_of_bytes());
    if (fdk3d_out) std::memcpy(fdk3d_out, r.fdk_3d->get_data_ptr(), r.fdk_3d->get_number_of_bytes());
    if (scale_out) *scale_out = r.scale;
  });
}

int gtt_os_sart(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
                const unsigned *bin_index, const int *bin_sizes, int nbins, const float *projections,
                int nsubsets, int iterations, float beta, float gamma, int nonneg, float *volume_out,
                double *objective_out) {
  return gtt_os_sart_weighted(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections,
                              nullptr, nsubsets, iterations, beta, gamma, nonneg, volume_out,
                              objective_out);
}

int gtt_os_sart_weighted(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
                         const unsigned *bin_index, const int *bin_sizes, int nbins,
                         const float *projections, const float *weights, int nsubsets,
                         int iterations, float beta, float gamma, int nonneg, float *volume_out,
                         double *objective_out) {
  return guarded([&] {
    if (!projections || !volume_out) throw std::runtime_error("illegal array pointer provided");
    if (nsubsets <= 0) throw std::runtime_error("number of subsets must be positive");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections,
                       weights == nullptr);
    auto E = make_subset_operator(nsubsets, s, weights);
    E->set_domain_dimensions(&s.vol_dims);
    if (default_autotune()) E->autotune();
    cuNDArray<float> b = subset_data(*E, s, weights);
    osSARTSolver<cuNDArray<float>> solver;
    solver.set_encoding_operator(E);
    solver.set_max_iterations(iterations);
    solver.set_beta(beta);
    solver.set_gamma(gamma);
    solver.set_non_negativity_constraint(nonneg != 0);
    solver.set_graphs(default_graphs());
    if (objective_out)
      solver.set_objective_callback([&](int it, double v) { objective_out[it] = v; });
    auto x = solver.solve(&b);
    copy_d2h(volume_out, x->get_data_ptr(), x->get_number_of_bytes());
  });
}

int gtt_os_mom(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
               const unsigned *bin_index, const int *bin_sizes, int nbins, const float *projections,
               int nsubsets, int iterations, float tv_weight, float tau, int reg_steps, float beta,
               float gamma, int nonneg, int fuse_tv, float *volume_out) {
  return gtt_os_mom_weighted(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections,
                             nullptr, nsubsets, iterations, tv_weight, tau, reg_steps, beta, gamma,
                             nonneg, fuse_tv, volume_out);
}

int gtt_os_mom_weighted(const cbct_geom *g, const float *angles, const float *offsets, int nviews,
                        const unsigned *bin_index, const int *bin_sizes, int nbins,
                        const float *projections, const float *weights, int nsubsets,
                        int iterations, float tv_weight, float tau, int reg_steps, float beta,
                        float gamma, int nonneg, int fuse_tv, float *volume_out) {
  return guarded([&] {
    if (!projections || !volume_out) throw std::runtime_error("illegal array pointer provided");
    if (nsubsets <= 0) throw std::runtime_error("number of subsets must be positive");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections,
                       weights == nullptr);
    auto E = make_subset_operator(nsubsets, s, weights);
    E->set_domain_dimensions(&s.vol_dims);
    if (default_autotune()) E->autotune();
    cuNDArray<float> b = subset_data(*E, s, weights);
    osMOMSolverD<cuNDArray<float>> solver;
    solver.set_dump(false);
    solver.set_encoding_operator(E);
    solver.set_max_iterations(iterations);
    solver.set_tau(tau);
    solver.set_non_negativity_constraint(nonneg != 0);
    solver.set_reg_steps((unsigned)reg_steps);
    solver.set_beta(beta);
    solver.set_gamma(gamma);
    solver.set_fuse_tv(fuse_tv != 0);
    if (tv_weight > 0) {
      auto TV = std::make_shared<cuTVPrimalDualOperator>(0.0f);
      TV->set_weight(tv_weight);
      solver.add_regularization_operator(TV);
    }
    auto z = solver.solve(&b);
    copy_d2h(volume_out, z->get_data_ptr(), z->get_number_of_bytes());
  });
}

int gtt_subset_op_create(int nsubsets, gtt_subset_op *out) {
  return guarded([&] {
    if (!out || nsubsets <= 0) throw std::runtime_error("gtt_subset_op_create: bad arguments");
    auto *h = new gtt_subset_op_s();
    h->op = std::make_shared<CBSubsetOperator>(nsubsets);
    *out = h;
  });
}

int gtt_subset_op_destroy(gtt_subset_op op) {
  delete op;
  return 0;
}

int gtt_subset_op_setup(gtt_subset_op op, const cbct_geom *g, const float *angles,
                        const float *offsets, int nviews, const unsigned *bin_index,
                        const int *bin_sizes, int nbins, const float *projections,
                        float *permuted_out, unsigned *permutation_out) {
  return gtt_subset_op_setup_weighted(op, g, angles, offsets, nviews, bin_index, bin_sizes, nbins,
                                      projections, nullptr, permuted_out, permutation_out);
}

int gtt_subset_op_setup_weighted(gtt_subset_op op, const cbct_geom *g, const float *angles,
                                 const float *offsets, int nviews, const unsigned *bin_index,
                                 const int *bin_sizes, int nbins, const float *projections,
                                 const float *weights, float *permuted_out,
                                 unsigned *permutation_out) {
  return guarded([&] {
    if (!op) throw std::runtime_error("null operator");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, projections);
    if (!projections)  // projection dims still come from the geometry
      s.acq->set_projections(std::make_shared<hoCuNDArray<float>>(
          std::vector<size_t>{(size_t)g->nu, (size_t)g->nv, (size_t)nviews}));
    if (weights) {
      op->op = make_subset_operator(op->op->get_number_of_subsets(), s, weights);
    } else {
      if (std::dynamic_pointer_cast<CBSubsetWeightOperator>(op->op))
        op->op = std::make_shared<CBSubsetOperator>(op->op->get_number_of_subsets());
      op->op->setup(s.acq, s.binning, s.dims_mm);
    }
    op->op->set_domain_dimensions(&s.vol_dims);
    auto p = s.acq->get_projections();
    if (permuted_out && projections) std::memcpy(permuted_out, p->get_data_ptr(), p->get_number_of_bytes());
    if (permutation_out) {
      const auto &perm = op->op->get_permutation();
      for (size_t i = 0; i < perm.size(); ++i) permutation_out[i] = perm[i];
    }
  });
}

int gtt_subset_op_views(gtt_subset_op op, int subset) {
  if (!op) return -1;
  try {
    if (subset < 0) return (int)op->op->get_codomain_dimensions()->back();
    return (int)op->op->get_codomain_dimensions(subset)->back();
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int gtt_subset_op_mult_M(gtt_subset_op op, float *vol, float *proj, int subset, int accumulate) {
  return guarded([&] {
    auto dom = *op->op->get_domain_dimensions();
    cuNDArray<float> x(dom, vol);
    if (subset < 0) {
      cuNDArray<float> y(*op->op->get_codomain_dimensions(), proj);
      op->op->mult_M(&x, &y, accumulate != 0);
    } else {
      cuNDArray<float> y(*op->op->get_codomain_dimensions(subset), proj);
      op->op->mult_M(&x, &y, subset, accumulate != 0);
    }
  });
}

int gtt_subset_op_mult_MH(gtt_subset_op op, float *proj, float *vol, int subset, int accumulate) {
  return guarded([&] {
    auto dom = *op->op->get_domain_dimensions();
    cuNDArray<float> x(dom, vol);
    if (subset < 0) {
      cuNDArray<float> y(*op->op->get_codomain_dimensions(), proj);
      op->op->mult_MH(&y, &x, accumulate != 0);
    } else {
      cuNDArray<float> y(*op->op->get_codomain_dimensions(subset), proj);
      op->op->mult_MH(&y, &x, subset, accumulate != 0);
    }
  });
}

int gtt_cu_op_create(gtt_cu_op *out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("gtt_cu_op_create: null output");
    auto *h = new gtt_cu_op_s();
    h->op = std::make_shared<cuConebeamProjectionOperator>();
    *out = h;
  });
}

int gtt_cu_op_destroy(gtt_cu_op op) {
  return guarded([&] { delete op; });
}

int gtt_cu_op_setup(gtt_cu_op op, const cbct_geom *g, const float *angles, const float *offsets,
                    int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                    int transform_angles, int use_offset_correction) {
  return guarded([&] {
    if (!op) throw std::runtime_error("null operator");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, nullptr);
    op->op = std::make_shared<cuConebeamProjectionOperator>();
    if (use_offset_correction >= 0) op->op->set_use_offset_correction(use_offset_correction != 0);
    op->op->setup(s.acq, s.binning, s.dims_mm, transform_angles != 0);
    op->op->set_domain_dimensions(&s.vol_dims);
    op->dom = s.vol_dims;
    op->codom = {(size_t)g->nu, (size_t)g->nv, views_in_bins(*s.binning)};
    op->op->set_codomain_dimensions(&op->codom);
  });
}

int gtt_cu_op_mult_M(gtt_cu_op op, const float *vol, float *proj, int accumulate) {
  return guarded([&] {
    if (!op || op->dom.empty()) throw std::runtime_error("Error: cuConebeamProjectionOperator::mult_M: setup not performed");
    cuNDArray<float> x(op->dom, const_cast<float *>(vol));
    cuNDArray<float> y(op->codom, proj);
    op->op->mult_M(&x, &y, accumulate != 0);
  });
}

int gtt_cu_op_mult_MH(gtt_cu_op op, const float *proj, float *vol, int accumulate) {
  return guarded([&] {
    if (!op || op->dom.empty()) throw std::runtime_error("Error: cuConebeamProjectionOperator::mult_MH: setup not performed");
    cuNDArray<float> y(op->codom, const_cast<float *>(proj));
    cuNDArray<float> x(op->dom, vol);
    op->op->mult_MH(&y, &x, accumulate != 0);
  });
}

int gtt_cu_op_calculate_mask(gtt_cu_op op, const float *proj, float limit, unsigned char *mask) {
  return guarded([&] {
    if (!op || op->dom.empty()) throw std::runtime_error("calculate_mask: setup not performed");
    if (!proj || !mask) throw std::runtime_error("calculate_mask: illegal array pointer provided");
    cuNDArray<float> y(op->codom, const_cast<float *>(proj));
    auto m = op->op->calculate_mask(&y, limit);
    copy_d2d(mask, m->get_data_ptr(), m->get_number_of_bytes());
  });
}

int gtt_cu_op_set_mask(gtt_cu_op op, const unsigned char *mask) {
  return guarded([&] {
    if (!op || op->dom.empty()) throw std::runtime_error("set_mask: setup not performed");
    if (!mask) {
      op->op->set_mask(nullptr);
      return;
    }
    auto m = std::make_shared<cuNDArray<unsigned char>>(op->dom);
    copy_d2d(m->get_data_ptr(), mask, m->get_number_of_bytes());
    op->op->set_mask(m);
  });
}

int gtt_cu_op_flags(gtt_cu_op op, int *short_scan, int *use_offset_correction) {
  return guarded([&] {
    if (!op) throw std::runtime_error("null operator");
    if (short_scan) *short_scan = op->op->is_short_scan() ? 1 : 0;
    if (use_offset_correction) *use_offset_correction = op->op->get_use_offset_correction() ? 1 : 0;
  });
}

int gtt_ho_op_create(gtt_ho_op *out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("gtt_ho_op_create: null output");
    auto *h = new gtt_ho_op_s();
    h->op = std::make_shared<hoCuConebeamProjectionOperator>();
    *out = h;
  });
}

int gtt_ho_op_destroy(gtt_ho_op op) {
  return guarded([&] { delete op; });
}

int gtt_ho_op_setup(gtt_ho_op op, const cbct_geom *g, const float *angles, const float *offsets,
                    int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                    int use_fbp, int use_offset_correction, const float *filter, float spp) {
  return guarded([&] {
    if (!op) throw std::runtime_error("null operator");
    Scan s = make_scan(g, angles, offsets, nviews, bin_index, bin_sizes, nbins, nullptr);
    op->op = std::make_shared<hoCuConebeamProjectionOperator>();
    if (use_offset_correction >= 0) op->op->set_use_offset_correction(use_offset_correction != 0);
    op->op->set_use_filtered_backprojection(use_fbp != 0);
    if (spp > 0) op->op->set_num_samples_per_pixel(spp);
    if (filter)
      op->op->set_frequency_filter(std::make_shared<std::vector<float>>(filter, filter + g->nu + 1));
    op->op->setup(s.acq, s.binning, s.dims_mm, true);
    op->op->set_domain_dimensions(&s.vol_dims);
    op->dom = s.vol_dims;
    op->codom = {(size_t)g->nu, (size_t)g->nv, (size_t)nviews};
    op->op->set_codomain_dimensions(&op->codom);
  });
}

int gtt_ho_op_mult_M(gtt_ho_op op, const float *image, float *proj, int accumulate) {
  return guarded([&] {
    if (!op || op->dom.empty()) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: setup not performed");
    if (!image || !proj) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: illegal array pointer provided");
    hoCuNDArray<float> x(op->dom, const_cast<float *>(image));
    hoCuNDArray<float> y(op->codom, proj);
    op->op->mult_M(&x, &y, accumulate != 0);
  });
}

int gtt_ho_op_mult_MH(gtt_ho_op op, const float *proj, float *image, int accumulate) {
  return guarded([&] {
    if (!op || op->dom.empty()) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: setup not performed");
    if (!image || !proj) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH:: illegal array pointer provided");
    hoCuNDArray<float> y(op->codom, const_cast<float *>(proj));
    hoCuNDArray<float> x(op->dom, image);
    op->op->mult_MH(&y, &x, accumulate != 0);
  });
}

int gtt_ho_op_flags(gtt_ho_op op, int *short_scan, int *use_offset_correction) {
  return guarded([&] {
    if (!op) throw std::runtime_error("null operator");
    if (short_scan) *short_scan = op->op->is_short_scan() ? 1 : 0;
    if (use_offset_correction) *use_offset_correction = op->op->get_use_offset_correction() ? 1 : 0;
  });
}

int gtt_write_real(const char *path, int ndims, const size_t *dims, const float *data) {
  return guarded([&] {
    if (!path || ndims <= 0 || ndims > 16 || !dims || !data)
      throw std::runtime_error("gtt_write_real: illegal arguments");
    std::vector<size_t> d(dims, dims + ndims);
    hoCuNDArray<float> a(d, const_cast<float *>(data));
    write_nd_array(&a, path);
  });
}

int gtt_read_real_header(const char *path, int *ndims, size_t *dims) {
  return guarded([&] {
    if (!path || !ndims || !dims) throw std::runtime_error("gtt_read_real_header: illegal arguments");
    std::vector<size_t> d = read_nd_array_dims(path);
    *ndims = (int)d.size();
    for (size_t i = 0; i < d.size(); ++i) dims[i] = d[i];
  });
}

int gtt_read_real(const char *path, float *data, size_t n) {
  return guarded([&] {
    if (!path || !data) throw std::runtime_error("gtt_read_real: illegal arguments");
    hoCuNDArray<float> a = read_nd_array(path);
    if (a.get_number_of_elements() != n)
      throw std::runtime_error("gtt_read_real: the file holds " + std::to_string(a.get_number_of_elements()) +
                               " elements, not " + std::to_string(n));
    std::memcpy(data, a.get_data_ptr(), a.get_number_of_bytes());
  });
}

int gtt_cbct_forwards_projection(const cbct_geom *g, float start_angle, float angular_spacing,
                                 int nviews, const float *offsets, const unsigned *bin_index,
                                 const int *bin_sizes, int nbins, float spp, const float *image,
                                 float *proj_out, float *angles_out) {
  return guarded([&] {
    if (!g || nviews <= 0 || !image || !proj_out)
      throw std::runtime_error("CBCT_forwards_projection: illegal arguments");
    std::vector<float> angles(nviews);
    for (int i = 0; i < nviews; ++i) angles[i] = start_angle + (float)(unsigned)i * angular_spacing;
    if (angles_out) std::memcpy(angles_out, angles.data(), sizeof(float) * nviews);
    Scan s = make_scan(g, angles.data(), offsets, nviews, bin_index, bin_sizes, nbins, nullptr);
    if (s.binning->get_maximum_projection_index() >= nviews)
      throw std::runtime_error("Maximum projection index in binning file exceeds the number of projections requested");
    std::vector<size_t> ps_dims{(size_t)g->nu, (size_t)g->nv, (size_t)nviews};
    hoCuNDArray<float> projections(ps_dims, proj_out);
    std::memset(proj_out, 0, projections.get_number_of_bytes());   // the binning may not write every view
    hoCuConebeamProjectionOperator E;
    if (spp > 0) E.set_num_samples_per_pixel(spp);
    E.setup(s.acq, s.binning, s.dims_mm);
    std::vector<size_t> idims = s.vol_dims;
    if (idims.size() == 3) idims.push_back(1);
    hoCuNDArray<float> img(idims, const_cast<float *>(image));
    E.mult_M(&img, &projections);
  });
}

}  // extern "C"
