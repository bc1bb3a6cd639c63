// cuConebeamProjectionOperator, CBSubsetOperator, cuTVPrimalDualOperator and
// the osMOMSolverD<cuNDArray<float>> element-wise specialisations.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <map>

#include "gtt/operators.h"
#include "gtt/solvers.h"

namespace gtt {

namespace {

const float kPi = 3.14159265358979f;

// hoCuConebeamProjectionOperator.h:57-65 / cuConebeamProjectionOperator.cpp:166-174
bool detect_short_scan(const std::vector<float> &angles, floatd2 fov, float sdd) {
  if (angles.empty()) return false;
  const float span = *std::max_element(angles.begin(), angles.end());
  const float delta = std::atan(fov[0] / (2.0f * sdd));
  return !(span * kPi / 180.0f > kPi + 3.0f * delta);
}

floatd2 mean_offset(const std::vector<floatd2> &o) {
  floatd2 m{{0.0f, 0.0f}};
  for (auto &v : o) {
    m[0] += v[0];
    m[1] += v[1];
  }
  if (!o.empty()) {
    m[0] /= float(o.size());
    m[1] /= float(o.size());
  }
  return m;
}

void upload_views(cbct_ctx ctx, const std::vector<float> &a, const std::vector<floatd2> &o) {
  if (a.empty()) return;
  std::vector<float> off(2 * a.size());
  for (size_t i = 0; i < a.size(); ++i) {
    off[2 * i] = o[i][0];
    off[2 * i + 1] = o[i][1];
  }
  check(cbct_views_set(ctx, a.data(), off.data(), (int)a.size()), ctx);
}

}  // namespace

// ---------------------------------------------------------------------------
cuConebeamProjectionOperator::cuConebeamProjectionOperator() {
  ctx_ = acquire_context();
  check(cbct_storage(ctx_, default_storage()), ctx_);
}

cuConebeamProjectionOperator::~cuConebeamProjectionOperator() { release_context(ctx_); }

cbct_geom cuConebeamProjectionOperator::geom_for(size_t nx, size_t ny, size_t nz, size_t nu,
                                                 size_t nv) const {
  auto g = acquisition_->get_geometry();
  cbct_geom cg;
  cg.nx = (int)nx;
  cg.ny = (int)ny;
  cg.nz = (int)nz;
  cg.lx = is_dims_in_mm_[0];
  cg.ly = is_dims_in_mm_[1];
  cg.lz = is_dims_in_mm_[2];
  cg.nu = (int)nu;
  cg.nv = (int)nv;
  cg.fov_u = g->get_FOV()[0];
  cg.fov_v = g->get_FOV()[1];
  cg.sad = g->get_SAD();
  cg.sdd = g->get_SDD();
  return cg;
}

void cuConebeamProjectionOperator::setup(std::shared_ptr<CBCT_acquisition> acquisition,
                                         floatd3 is_dims_in_mm, bool transform_angles) {
  acquisition_ = acquisition;
  is_dims_in_mm_ = is_dims_in_mm;
  auto geom = acquisition->get_geometry();
  std::vector<float> &all_angles = geom->get_angles();
  if (transform_angles && !all_angles.empty()) {
    const float mn = *std::min_element(all_angles.begin(), all_angles.end());
    for (auto &a : all_angles) a -= mn;
  }
  short_scan_ = detect_short_scan(all_angles, geom->get_FOV(), geom->get_SDD());
  std::vector<floatd2> all_offsets = geom->get_offsets();
  if (all_offsets.size() < all_angles.size()) all_offsets.resize(all_angles.size(), floatd2{{0, 0}});
  const floatd2 mo = mean_offset(all_offsets);
  if (allow_offset_correction_override_ && std::abs(mo[0]) > geom->get_FOV()[0] * 0.1f)
    use_offset_correction_ = true;
  preprocessed_ = true;
  // The reference sizes this default bin by its still-empty per-bin angle
  // list (an empty bin, :184); it is sized by the number of views here.
  if (!binning_) {
    std::vector<unsigned int> bins(all_angles.size());
    for (size_t i = 0; i < bins.size(); ++i) bins[i] = (unsigned int)i;
    binning_ = std::make_shared<CBCT_binning>(std::vector<std::vector<unsigned int>>(1, bins));
  }
  angles_.clear();
  offsets_.clear();
  std::vector<float> table_a;
  std::vector<floatd2> table_o;
  bin_view0_.clear();
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    angles_.emplace_back();
    offsets_.emplace_back();
    bin_view0_.push_back((int)table_a.size());
    for (auto idx : binning_->bin(b)) {
      if (idx >= all_angles.size())
        throw std::runtime_error("Error: conebeam_forwards_projection: illegal index in bin");
      angles_.back().push_back(all_angles[idx]);
      offsets_.back().push_back(all_offsets[idx]);
      table_a.push_back(all_angles[idx]);
      table_o.push_back(all_offsets[idx]);
    }
  }
  all_view0_ = (int)table_a.size();
  table_a.insert(table_a.end(), all_angles.begin(), all_angles.end());
  table_o.insert(table_o.end(), all_offsets.begin(), all_offsets.begin() + all_angles.size());
  upload_views(ctx_, table_a, table_o);
  // permute the projections into bin order (:201-204)
  auto proj = acquisition->get_projections();
  if (proj && proj->get_number_of_elements()) {
    std::vector<unsigned int> order;
    for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b)
      for (auto p : binning_->bin(b)) order.push_back(p);
    const size_t plane = proj->get_size(0) * proj->get_size(1);
    auto out = std::make_shared<hoCuNDArray<float>>(
        std::vector<size_t>{proj->get_size(0), proj->get_size(1), order.size()});
    for (size_t i = 0; i < order.size(); ++i)
      std::copy(proj->get_data_ptr() + plane * order[i], proj->get_data_ptr() + plane * (order[i] + 1),
                out->get_data_ptr() + plane * i);
    *proj = *out;
  }
}

void cuConebeamProjectionOperator::setup(std::shared_ptr<CBCT_acquisition> acquisition,
                                         std::shared_ptr<CBCT_binning> binning,
                                         floatd3 is_dims_in_mm, bool transform_angles) {
  binning_ = binning;
  setup(acquisition, is_dims_in_mm, transform_angles);
}

void cuConebeamProjectionOperator::mult_M(cuNDArray<float> *input, cuNDArray<float> *output,
                                          bool accumulate) {
  if (!input || !output)
    throw std::runtime_error("Error: cuConebeamProjectionOperator::mult_M: illegal array pointer provided");
  if (!preprocessed_)
    throw std::runtime_error("Error: cuConebeamProjectionOperator::mult_M: setup not performed");
  std::vector<size_t> dims3d = input->dims();
  if (dims3d.size() == 4) dims3d.pop_back();
  if (dims3d.size() != 3 || output->get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: conebeam_forwards_projection: image array must be three-dimensional");
  const size_t n3 = prod(dims3d);
  const size_t plane = output->get_size(0) * output->get_size(1);
  float *in_ptr = input->get_data_ptr();
  float *out_ptr = output->get_data_ptr();
  const cbct_geom g = geom_for(dims3d[0], dims3d[1], dims3d[2], output->get_size(0), output->get_size(1));
  size_t total = 0;
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) total += angles_[b].size();
  if (total != output->get_size(2))
    throw std::runtime_error("Error: conebeam_forwards_projection: inconsistent sizes of input arrays/vectors");
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    if (binning_->bin(b).empty()) continue;
    const int nb = (int)angles_[b].size();
    check(cbct_fp(ctx_, &g, bin_view0_[b], nb, in_ptr, CBCT_F32, out_ptr, samples_per_pixel_,
                  accumulate ? 1 : 0, current_stream()), ctx_);
    // the volume pointer advances only past non-empty bins, as in the reference
    in_ptr += n3;
    out_ptr += plane * nb;
  }
}

void cuConebeamProjectionOperator::autotune(size_t nu, size_t nv) {
  if (!preprocessed_) throw std::runtime_error("Error: cuConebeamProjectionOperator::autotune: setup not performed");
  std::vector<size_t> dims3d = domain_;
  if (dims3d.size() == 4) dims3d.pop_back();
  if (dims3d.size() != 3)
    throw std::runtime_error("Error: cuConebeamProjectionOperator::autotune: domain dimensions not set");
  size_t best = 0;
  for (unsigned int b = 1; b < binning_->get_number_of_bins(); ++b)
    if (angles_[b].size() > angles_[best].size()) best = b;
  if (angles_.empty() || angles_[best].empty()) return;
  const cbct_geom g = geom_for(dims3d[0], dims3d[1], dims3d[2], nu, nv);
  check(cbct_fp_autotune(ctx_, &g, bin_view0_[best], (int)angles_[best].size(), CBCT_F32,
                         samples_per_pixel_, 0, current_stream(), nullptr), ctx_);
}

void cuConebeamProjectionOperator::mult_MH(cuNDArray<float> *input, cuNDArray<float> *output,
                                           bool accumulate) {
  if (!input || !output)
    throw std::runtime_error("Error: cuConebeamProjectionOperator::mult_MH: illegal array pointer provided");
  if (!preprocessed_)
    throw std::runtime_error("Error: cuConebeamProjectionOperator::mult_MH: setup not performed");
  std::vector<size_t> dims3d = output->dims();
  if (dims3d.size() == 4) dims3d.pop_back();
  if (dims3d.size() != 3 || input->get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: conebeam_backwards_projection: image array must be three-dimensional");
  const size_t n3 = prod(dims3d);
  const size_t plane = input->get_size(0) * input->get_size(1);
  float *in_ptr = input->get_data_ptr();
  float *out_ptr = output->get_data_ptr();
  const cbct_geom g = geom_for(dims3d[0], dims3d[1], dims3d[2], input->get_size(0), input->get_size(1));
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    if (binning_->bin(b).empty()) continue;
    const int nb = (int)angles_[b].size();
    const float *src = in_ptr;
    cuNDArray<float> weighted;
    if (use_offset_correction_) {
      // full half-fan weights on a copy (conebeam_projection.cu:1242-1244)
      weighted = cuNDArray<float>(std::vector<size_t>{input->get_size(0), input->get_size(1), (size_t)nb});
      copy_d2d(weighted.get_data_ptr(), in_ptr, plane * nb * sizeof(float));
      check(cbct_offset_weights(ctx_, &g, bin_view0_[b], nb, weighted.get_data_ptr(), 0, current_stream()), ctx_);
      src = weighted.get_data_ptr();
    }
    check(cbct_bp(ctx_, &g, bin_view0_[b], nb, src, CBCT_F32, out_ptr, 0, accumulate ? 1 : 0, current_stream()), ctx_);
    in_ptr += plane * nb;
    out_ptr += n3;
  }
  if (mask_) {
    cbct_ctx c = device_context();
    check(cbct_apply_mask(c, output->get_number_of_elements(), output->get_data_ptr(),
                          mask_->get_data_ptr(), current_stream()), c);
  }
}

void cuConebeamProjectionOperator::offset_correct(cuNDArray<float> *projections) {
  if (!preprocessed_)
    throw std::runtime_error("Error: cuConebeamProjectionOperator::offset_correct: setup not performed");
  const size_t nv = projections->get_size(2);
  if ((int)nv > (int)acquisition_->get_geometry()->get_angles().size())
    throw std::runtime_error("Error: offset_correct: inconsistent sizes of input arrays/vectors");
  const cbct_geom g = geom_for(1, 1, 1, projections->get_size(0), projections->get_size(1));
  check(cbct_offset_weights(ctx_, &g, all_view0_, (int)nv, projections->get_data_ptr(), 1, current_stream()), ctx_);
}

std::shared_ptr<cuNDArray<unsigned char>> cuConebeamProjectionOperator::calculate_mask(
    cuNDArray<float> *projections, float limit) {
  auto dims = *get_domain_dimensions();
  if (dims.size() < 3) throw std::runtime_error("calculate_mask: domain dimensions not set");
  std::vector<size_t> dims3d(dims.begin(), dims.begin() + 3);
  auto mask = std::make_shared<cuNDArray<unsigned char>>(dims);
  unsigned char *mp = mask->get_data_ptr();
  const float *ip = projections->get_data_ptr();
  const size_t plane = projections->get_size(0) * projections->get_size(1);
  const cbct_geom g = geom_for(dims3d[0], dims3d[1], dims3d[2], projections->get_size(0),
                               projections->get_size(1));
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    if (binning_->bin(b).empty()) continue;
    const int nb = (int)angles_[b].size();
    check(cbct_spacecarver(ctx_, &g, bin_view0_[b], nb, ip, mp, limit, current_stream()), ctx_);
    ip += plane * nb;
    mp += prod(dims3d);
  }
  return mask;
}

// ---------------------------------------------------------------------------
CBSubsetOperator::CBSubsetOperator(int subsets) : subsetOperator<cuNDArray<float>>(subsets) {
  for (int i = 0; i < subsets; ++i) operators_.push_back(std::make_shared<cuConebeamProjectionOperator>());
}

void CBSubsetOperator::mult_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool acc) {
  operators_.at(subset)->mult_M(in, out, acc);
}
void CBSubsetOperator::mult_MH(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool acc) {
  operators_.at(subset)->mult_MH(in, out, acc);
}
void CBSubsetOperator::mult_MH_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool acc) {
  cuNDArray<float> tmp(get_codomain_dimensions(subset));
  mult_M(in, &tmp, subset, false);
  mult_MH(&tmp, out, subset, acc);
}

std::shared_ptr<std::vector<size_t>> CBSubsetOperator::get_codomain_dimensions(int subset) {
  return projection_dims_.at(subset);
}

void CBSubsetOperator::setup(std::shared_ptr<CBCT_acquisition> acq,
                             std::shared_ptr<CBCT_binning> binning, floatd3 is_dims_in_mm,
                             bool permute_projections) {
  auto geometry = acq->get_geometry();
  std::vector<float> all_angles = geometry->get_angles();
  std::vector<floatd2> all_offsets = geometry->get_offsets();
  if (all_offsets.size() < all_angles.size()) all_offsets.resize(all_angles.size(), floatd2{{0, 0}});
  if (all_angles.empty()) throw std::runtime_error("CBSubsetOperator::setup: no angles");
  const float mn = *std::min_element(all_angles.begin(), all_angles.end());
  for (auto &a : all_angles) a -= mn;

  const int S = number_of_subsets;
  std::vector<std::vector<float>> subset_angles(S);
  std::vector<std::vector<floatd2>> subset_offsets(S);
  std::vector<std::vector<unsigned int>> subset_projections(S);
  std::map<unsigned int, unsigned int> bins;
  auto bin_vector = binning->get_bins();
  for (unsigned int b = 0; b < bin_vector.size(); ++b)
    for (auto p : bin_vector[b]) bins[p] = b;
  int k = 0;
  for (unsigned int i = 0; i < all_angles.size(); ++i)
    if (bins.count(i)) {
      subset_projections[k % S].push_back(i);
      ++k;
    }
  // Sort subsets by bin.  The reference's std::sort is unstable with equal
  // keys; a stable sort keeps ascending view order within a bin.
  for (auto &s : subset_projections)
    std::stable_sort(s.begin(), s.end(), [&](unsigned a, unsigned b) { return bins[a] < bins[b]; });
  for (int i = 0; i < S; ++i)
    for (auto p : subset_projections[i]) {
      subset_angles[i].push_back(all_angles[p]);
      subset_offsets[i].push_back(all_offsets[p]);
    }
  permutations_.clear();
  for (auto &s : subset_projections)
    for (auto p : s) permutations_.push_back(p);

  // permute the projection array (:106, :192-212)
  auto proj = acq->get_projections();
  if (proj && proj->get_number_of_elements() && !permute_projections) {
    for (auto p : permutations_)
      if (p >= proj->get_size(2)) throw std::runtime_error("Error: CBSubsetOperator: illegal index in bin");
  } else if (proj && proj->get_number_of_elements()) {
    auto dims = proj->dims();
    dims.back() = permutations_.size();
    auto res = std::make_shared<hoCuNDArray<float>>(dims, no_init);
    const size_t plane = proj->get_size(0) * proj->get_size(1);
    for (size_t i = 0; i < permutations_.size(); ++i) {
      if (permutations_[i] >= proj->get_size(2))
        throw std::runtime_error("Error: CBSubsetOperator: illegal index in bin");
      std::copy(proj->get_data_ptr() + plane * permutations_[i],
                proj->get_data_ptr() + plane * (permutations_[i] + 1), res->get_data_ptr() + plane * i);
    }
    acq->set_projections(res);
  }
  std::vector<float> new_angles;
  std::vector<floatd2> new_offsets;
  for (auto p : permutations_) {
    new_angles.push_back(all_angles[p]);
    new_offsets.push_back(all_offsets[p]);
  }
  geometry->set_angles(new_angles);
  geometry->set_offsets(new_offsets);

  for (int s = 0; s < S; ++s) {
    auto sg = std::make_shared<CBCT_geometry>(*geometry);
    sg->set_angles(subset_angles[s]);
    sg->set_offsets(subset_offsets[s]);
    auto sa = std::make_shared<CBCT_acquisition>();
    sa->set_geometry(sg);
    std::vector<std::vector<unsigned int>> sb(binning->get_number_of_bins());
    for (unsigned int i = 0; i < subset_projections[s].size(); ++i)
      sb[bins[subset_projections[s][i]]].push_back(i);
    operators_[s]->setup(sa, std::make_shared<CBCT_binning>(sb), is_dims_in_mm, false);
  }
  size_t nu = 0, nv = 0;
  if (proj) {
    nu = acq->get_projections()->get_size(0);
    nv = acq->get_projections()->get_size(1);
  }
  projection_dims_.clear();
  for (int s = 0; s < S; ++s)
    projection_dims_.push_back(std::make_shared<std::vector<size_t>>(std::vector<size_t>{nu, nv, subset_angles[s].size()}));
}

void CBSubsetOperator::setup(std::shared_ptr<CBCT_acquisition> acq, floatd3 is_dims_in_mm) {
  auto b = std::make_shared<CBCT_binning>();
  b->set_as_default_3d_bin((unsigned int)acq->get_projections()->get_size(2));
  setup(acq, b, is_dims_in_mm);
}

void CBSubsetOperator::offset_correct(cuNDArray<float> *projections) {
  auto sets = projection_subsets(projections);
  for (int s = 0; s < number_of_subsets; ++s)
    if (operators_[s]->get_use_offset_correction()) operators_[s]->offset_correct(sets[s].get());
}

cuNDArray<float> upload_in_subset_order(const hoCuNDArray<float> &projections,
                                        const std::vector<unsigned int> &permutation) {
  const size_t nu = projections.get_size(0), nv = projections.get_size(1);
  const size_t n = projections.get_size(2), plane = nu * nv;
  for (auto p : permutation)
    if (p >= n) throw std::runtime_error("Error: CBSubsetOperator: illegal index in bin");
  cuNDArray<float> raw(projections.dims());
  copy_h2d(raw.get_data_ptr(), projections.get_data_ptr(), projections.get_number_of_bytes());
  cuNDArray<float> out(std::vector<size_t>{nu, nv, permutation.size()});
  for (size_t i = 0; i < permutation.size();) {
    size_t run = 1;
    while (i + run < permutation.size() && permutation[i + run] == permutation[i + run - 1] + 1) ++run;
    copy_d2d(out.get_data_ptr() + plane * i, raw.get_data_ptr() + plane * permutation[i],
             run * plane * sizeof(float));
    i += run;
  }
  return out;
}

// ---------------------------------------------------------------------------
void CBSubsetWeightOperator::setup(std::shared_ptr<CBCT_acquisition> acq,
                                   std::shared_ptr<CBCT_binning> binning, floatd3 is_dims_in_mm,
                                   std::shared_ptr<hoCuNDArray<float>> projection_weights) {
  auto proj = acq->get_projections();
  if (!proj || !projection_weights || proj->dims() != projection_weights->dims())
    throw std::runtime_error("Weight dimensions must match that of the projection data");
  {
    float *p = proj->get_data_ptr();
    const float *w = projection_weights->get_data_ptr();
    for (size_t i = 0, n = proj->get_number_of_elements(); i < n; ++i) p[i] *= w[i];
  }
  CBSubsetOperator::setup(acq, binning, is_dims_in_mm);
  // the weights in the projections' permuted order (permute_projections)
  auto dims = projection_weights->dims();
  dims.back() = permutations_.size();
  hoCuNDArray<float> permuted(dims, no_init);
  const size_t plane = projection_weights->get_size(0) * projection_weights->get_size(1);
  for (size_t i = 0; i < permutations_.size(); ++i)
    std::copy(projection_weights->get_data_ptr() + plane * permutations_[i],
              projection_weights->get_data_ptr() + plane * (permutations_[i] + 1),
              permuted.get_data_ptr() + plane * i);
  weights_ = cuNDArray<float>(permuted);
  subset_weights_ = projection_subsets(&weights_);
}

void CBSubsetWeightOperator::mult_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool acc) {
  if (subset < 0 || subset >= (int)subset_weights_.size())
    throw std::runtime_error("CBSubsetWeightOperator::mult_M: setup not performed or subset out of range");
  if (acc) {
    cuNDArray<float> tmp(out->dims());
    CBSubsetOperator::mult_M(in, &tmp, subset, false);
    tmp *= *subset_weights_[subset];
    *out += tmp;
  } else {
    CBSubsetOperator::mult_M(in, out, subset, false);
    *out *= *subset_weights_[subset];
  }
}

void CBSubsetWeightOperator::mult_MH(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool acc) {
  if (subset < 0 || subset >= (int)subset_weights_.size())
    throw std::runtime_error("CBSubsetWeightOperator::mult_MH: setup not performed or subset out of range");
  cuNDArray<float> tmp_in(*in);
  tmp_in *= *subset_weights_[subset];
  CBSubsetOperator::mult_MH(&tmp_in, out, subset, acc);
}

void CBSubsetWeightOperator::mult_MH_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool acc) {
  cuNDArray<float> tmp(get_codomain_dimensions(subset));
  mult_M(in, &tmp, subset, false);
  mult_MH(&tmp, out, subset, acc);
}

// ---------------------------------------------------------------------------
void cuTVPrimalDualOperator::primalDual(cuNDArray<float> *in, cuNDArray<float> *out, float sigma,
                                        bool accumulate) {
  if (!in->dimensions_equal(out)) throw std::runtime_error("Input and reference dimensions must agree");
  if (in->get_number_of_dimensions() < 3) throw std::runtime_error("cuTVPrimalDualOperator: need 3-D input");
  size_t remaining = 1;
  for (size_t i = 3; i < in->get_number_of_dimensions(); ++i) remaining *= in->get_size(i);
  cbct_ctx c = device_context();
  check(cbct_tv_primal_dual(c, (int)in->get_size(0), (int)in->get_size(1), (int)in->get_size(2),
                            (int)remaining, in->get_data_ptr(), out->get_data_ptr(), sigma, alpha,
                            weight, accumulate ? 1 : 0, current_stream()), c);
}

template <>
void osMOMSolverD<cuNDArray<float>>::updateG(cuNDArray<float> &g, cuNDArray<float> &x,
                                             cuNDArray<float> &s, cuNDArray<float> &precon,
                                             float tau, float avg_lambda) {
  cbct_ctx c = device_context();
  check(cbct_update_g(c, x.get_number_of_elements(), g.get_data_ptr(), x.get_data_ptr(),
                      s.get_data_ptr(), precon.get_data_ptr(), tau, avg_lambda, current_stream()), c);
}

template <>
void osMOMSolverD<cuNDArray<float>>::fused_tv_update(cuTVPrimalDualOperator *tv,
                                                     cuNDArray<float> *x, cuNDArray<float> *g,
                                                     cuNDArray<float> &s,
                                                     cuNDArray<float> &precon, float sigma,
                                                     float tau, float avg_lambda) {
  size_t remaining = 1;
  for (size_t i = 3; i < x->get_number_of_dimensions(); ++i) remaining *= x->get_size(i);
  cbct_ctx c = device_context();
  check(cbct_tv_update_g(c, (int)x->get_size(0), (int)x->get_size(1), (int)x->get_size(2),
                         (int)remaining, x->get_data_ptr(), g->get_data_ptr(), s.get_data_ptr(),
                         precon.get_data_ptr(), sigma, tv->get_alpha(), tv->get_weight(), tau,
                         avg_lambda, current_stream()), c);
  x->swap(*g);
}

}  // namespace gtt
