// hoCuConebeamProjectionOperator (xray/hoCuConebeamProjectionOperator.{h,cpp}).
// Host arrays in and out; each call stages the bin's data in HBM, runs the
// gfx950 kernels and copies the result back.  The reference streams 20-view
// batches through two small device buffers; with 288 GB of HBM a whole bin is
// staged at once.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "gtt/fdk.h"
#include "gtt/operators.h"

namespace gtt {

namespace {
const float kPi = 3.14159265358979f;
}



hoCuConebeamProjectionOperator::hoCuConebeamProjectionOperator() {
  ctx_ = acquire_context();
  check(cbct_storage(ctx_, default_storage()), ctx_);
}

hoCuConebeamProjectionOperator::~hoCuConebeamProjectionOperator() { release_context(ctx_); }

void hoCuConebeamProjectionOperator::set_binning(std::shared_ptr<CBCT_binning> b) {
  binning_ = std::move(b);
  bin_view0_.clear();
}

cbct_geom hoCuConebeamProjectionOperator::geom_for(size_t nx, size_t ny, size_t nz, size_t nu,
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

void hoCuConebeamProjectionOperator::setup(std::shared_ptr<CBCT_acquisition> acquisition,
                                           floatd3 is_dims_in_mm, bool transform_angles) {
  acquisition_ = acquisition;
  is_dims_in_mm_ = is_dims_in_mm;
  auto geom = acquisition->get_geometry();
  std::vector<float> &angles = geom->get_angles();
  if (angles.empty()) throw std::runtime_error("hoCuConebeamProjectionOperator::setup: no angles");
  if (transform_angles) {
    const float mn = *std::min_element(angles.begin(), angles.end());
    for (auto &a : angles) a -= mn;
  }
  const float span = *std::max_element(angles.begin(), angles.end());
  const float delta = std::atan(geom->get_FOV()[0] / (2.0f * geom->get_SDD()));
  short_scan_ = !(span * kPi / 180.0f > kPi + 3.0f * delta);
  auto &offs = geom->get_offsets();
  if (offs.size() < angles.size()) offs.resize(angles.size(), floatd2{{0, 0}});
  float mo = 0.0f;
  for (auto &o : offs) mo += o[0];
  mo /= float(offs.size());
  if (allow_offset_correction_override_ && std::abs(mo) > geom->get_FOV()[0] * 0.1f)
    use_offset_correction_ = true;
  preprocessed_ = true;
  std::vector<unsigned int> bins(angles.size());
  for (size_t i = 0; i < bins.size(); ++i) bins[i] = (unsigned int)i;
  set_binning(std::make_shared<CBCT_binning>(std::vector<std::vector<unsigned int>>(1, bins)));
}

void hoCuConebeamProjectionOperator::setup(std::shared_ptr<CBCT_acquisition> acquisition,
                                           std::shared_ptr<CBCT_binning> binning,
                                           floatd3 is_dims_in_mm, bool transform_angles) {
  setup(acquisition, is_dims_in_mm, transform_angles);
  set_binning(binning);
}

// View table: the acquisition's views in order, then each bin's views.
void hoCuConebeamProjectionOperator::build_view_table() {
  auto geom = acquisition_->get_geometry();
  const auto &angles = geom->get_angles();
  const auto &offs = geom->get_offsets();
  std::vector<float> a(angles.begin(), angles.end());
  std::vector<float> o;
  for (size_t i = 0; i < angles.size(); ++i) {
    o.push_back(offs[i][0]);
    o.push_back(offs[i][1]);
  }
  bin_view0_.clear();
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    bin_view0_.push_back((int)a.size());
    for (auto idx : binning_->bin(b)) {
      if (idx >= angles.size())
        throw std::runtime_error("Error: conebeam_backwards_projection: illegal index in bin");
      a.push_back(angles[idx]);
      o.push_back(offs[idx][0]);
      o.push_back(offs[idx][1]);
    }
  }
  check(cbct_views_set(ctx_, a.data(), o.data(), (int)a.size()), ctx_);
}

void hoCuConebeamProjectionOperator::mult_M(hoCuNDArray<float> *image,
                                            hoCuNDArray<float> *projections, bool accumulate) {
  if (!image || !projections)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: illegal array pointer provided");
  if (image->get_number_of_dimensions() != 4 && image->get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: image array must be four or three -dimensional");
  if (projections->get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: projections array must be three-dimensional");
  if (!preprocessed_) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: setup not performed");
  if (!binning_) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: binning not provided");
  auto geom = acquisition_->get_geometry();
  if (projections->get_size(2) != geom->get_angles().size() ||
      projections->get_size(2) != geom->get_offsets().size())
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_M: inconsistent sizes of input arrays/vectors");
  if (bin_view0_.empty()) build_view_table();
  const size_t nu = projections->get_size(0), nv = projections->get_size(1);
  const size_t plane = nu * nv;
  std::vector<size_t> dims3d(image->dims().begin(), image->dims().begin() + 3);
  const size_t n3 = prod(dims3d);
  hoCuNDArray<float> tmp;
  hoCuNDArray<float> *out = projections;
  if (accumulate) {
    tmp = hoCuNDArray<float>(projections->dims());
    clear(&tmp);
    out = &tmp;
  }
  const cbct_geom g = geom_for(dims3d[0], dims3d[1], dims3d[2], nu, nv);
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    const auto &idx = binning_->bin(b);
    if (idx.empty()) continue;
    hoCuNDArray<float> vol_view(dims3d, image->get_data_ptr() + b * n3);
    cuNDArray<float> dvol(vol_view);
    cuNDArray<float> dproj(std::vector<size_t>{nu, nv, idx.size()});
    check(cbct_fp(ctx_, &g, bin_view0_[b], (int)idx.size(), dvol.get_data_ptr(), CBCT_F32,
                  dproj.get_data_ptr(), samples_per_pixel_, 0, current_stream()), ctx_);
    hoCuNDArray<float> hproj(dproj.dims());
    dproj.to_host(&hproj);
    for (size_t p = 0; p < idx.size(); ++p)
      std::memcpy(out->get_data_ptr() + plane * idx[p], hproj.get_data_ptr() + plane * p, plane * sizeof(float));
  }
  if (use_offset_correction_ && !use_fbp_) offset_correct(out);
  if (accumulate) {
    float *o = projections->get_data_ptr();
    const float *t = tmp.get_data_ptr();
    for (size_t i = 0, n = projections->get_number_of_elements(); i < n; ++i) o[i] += t[i];
  }
}

void hoCuConebeamProjectionOperator::mult_MH(hoCuNDArray<float> *projections,
                                             hoCuNDArray<float> *image, bool accumulate) {
  if (!image || !projections)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH:: illegal array pointer provided");
  if (image->get_number_of_dimensions() != 4 && image->get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: image array must be four or three -dimensional");
  if (projections->get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: projections array must be three-dimensional");
  if (!preprocessed_) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: setup not performed");
  if (!binning_) throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: binning not provided");
  auto geom = acquisition_->get_geometry();
  if (projections->get_size(2) != geom->get_angles().size() ||
      projections->get_size(2) != geom->get_offsets().size())
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: inconsistent sizes of input arrays/vectors");
  if (bin_view0_.empty()) build_view_table();
  const size_t nu = projections->get_size(0), nv = projections->get_size(1);
  const size_t plane = nu * nv;
  std::vector<size_t> dims3d(image->dims().begin(), image->dims().begin() + 3);
  const size_t n3 = prod(dims3d);
  const cbct_geom g = geom_for(dims3d[0], dims3d[1], dims3d[2], nu, nv);
  if (use_fbp_ && !filter_uploaded_) {
    if (!frequency_filter_) {
      // compute_default_frequency_filter (.cpp:26-62)
      frequency_filter_ = std::make_shared<std::vector<float>>(nu + 1);
      check(cbct_default_filter((int)nu, frequency_filter_->data()), ctx_);
    }
    if (frequency_filter_->size() != nu + 1)
      throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: frequency filter must hold nu+1 taps");
    check(cbct_filter_set(ctx_, (int)nu, frequency_filter_->data()), ctx_);
    filter_uploaded_ = true;
  }
  for (unsigned int b = 0; b < binning_->get_number_of_bins(); ++b) {
    const auto &idx = binning_->bin(b);
    float *img3 = image->get_data_ptr() + b * n3;
    if (idx.empty()) {
      // conebeam_projection.cu:1343-1350.  The reference then scales by
      // 2/0 in FBP mode (NaN); an empty bin is left at zero here.
      if (!accumulate) std::fill(img3, img3 + n3, 0.0f);
      continue;
    }
    const int nb = (int)idx.size();
    // stage the bin's projections (runs of consecutive indices in one copy)
    cuNDArray<float> dproj(std::vector<size_t>{nu, nv, idx.size()});
    size_t p = 0;
    while (p < idx.size()) {
      size_t run = 1;
      while (p + run < idx.size() && idx[p + run] == idx[p + run - 1] + 1) ++run;
      copy_h2d(dproj.get_data_ptr() + plane * p, projections->get_data_ptr() + plane * idx[p],
               run * plane * sizeof(float));
      p += run;
    }
    cuNDArray<float> dimg(dims3d);
    if (accumulate) copy_h2d(dimg.get_data_ptr(), img3, n3 * sizeof(float));
    if (use_fbp_) {
      check(cbct_fdk_prefilter(ctx_, &g, bin_view0_[b], nb, dproj.get_data_ptr(), dproj.get_data_ptr(),
                               short_scan_ ? 1 : 0, use_offset_correction_ ? 1 : 0, current_stream()), ctx_);
      check(cbct_bp(ctx_, &g, bin_view0_[b], nb, dproj.get_data_ptr(), CBCT_F32, dimg.get_data_ptr(), 1,
                    accumulate ? 1 : 0, current_stream()), ctx_);
      dimg *= 2.0f / float(nb);  // .cpp:253
    } else {
      if (use_offset_correction_)
        check(cbct_offset_weights(ctx_, &g, bin_view0_[b], nb, dproj.get_data_ptr(), 1, current_stream()), ctx_);
      check(cbct_bp(ctx_, &g, bin_view0_[b], nb, dproj.get_data_ptr(), CBCT_F32, dimg.get_data_ptr(), 0,
                    accumulate ? 1 : 0, current_stream()), ctx_);
    }
    copy_d2h(img3, dimg.get_data_ptr(), n3 * sizeof(float));
  }
}

void hoCuConebeamProjectionOperator::offset_correct(hoCuNDArray<float> *projections) {
  if (!preprocessed_)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::offset_correct: setup not performed");
  if (bin_view0_.empty()) build_view_table();
  const size_t nv = projections->get_size(2);
  if (nv > acquisition_->get_geometry()->get_angles().size())
    throw std::runtime_error("Error: offset_correct: inconsistent sizes of input arrays/vectors");
  cuNDArray<float> d(*projections);
  const cbct_geom g = geom_for(1, 1, 1, projections->get_size(0), projections->get_size(1));
  check(cbct_offset_weights(ctx_, &g, 0, (int)nv, d.get_data_ptr(), 1, current_stream()), ctx_);
  d.to_host(projections);
}

}  // namespace gtt
