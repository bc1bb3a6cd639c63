// FDK drivers, device-resident: FDK_reconstruct_3d
// (xray/FDK_reconstruct_3d.cpp:128-161) and the McKinnon-Bates
// FDK_reconstruct_4d (xray/FDK_reconstruct_4d.cpp:100-170).
//
// The reference runs both on host-resident hoCuConebeamProjectionOperators,
// which stream each bin through host memory and leave the projection of the
// result, and the inner products of the least-squares scale, on the host.
// Here each program stays in HBM: the caller's projections are uploaded once
// (no host copy), every filter, projection and reduction runs on the device
// (FP64 inner products), and only the result comes back.
//
// 4-D: f = FBP over the union of the bins; A f is one FP launch over the
// union's views (every bin holds the same copy of f, and a view's projection
// does not depend on the other views of a launch, so this is the per-bin
// result); s = <A f, p> / <A f, A f>; the residual views of all bins are
// gathered in bin order and pre-filtered in one launch (the FDK weights and
// the ramp filter act on each view alone); x_b = (FBP_b(p - s A f) + f) s.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "gtt/fdk.h"

namespace gtt {

namespace {

const float kPi = 3.14159265358979f;

struct CtxGuard {
  cbct_ctx c = nullptr;
  ~CtxGuard() { release_context(c); }
};

struct ScanFlags {
  bool short_scan = false, offset_correct = false;
};

// Operator setup (xray/hoCuConebeamProjectionOperator.h:39-93): angles
// shifted to start at zero (in place, as the reference does), short scan when
// the span does not exceed pi + 3 half fan angles, offset correction when the
// mean u offset exceeds a tenth of the detector width.  All of a driver's
// operators see the same acquisition and so the same flags.
ScanFlags operator_setup(CBCT_acquisition &acquisition) {
  auto geom = acquisition.get_geometry();
  std::vector<float> &angles = geom->get_angles();
  const float mn = *std::min_element(angles.begin(), angles.end());
  for (auto &a : angles) a -= mn;
  const float span = *std::max_element(angles.begin(), angles.end());
  const float delta = std::atan(geom->get_FOV()[0] / (2.0f * geom->get_SDD()));
  ScanFlags f;
  f.short_scan = !(span * kPi / 180.0f > kPi + 3.0f * delta);
  auto &offs = geom->get_offsets();
  if (offs.size() < angles.size()) offs.resize(angles.size(), floatd2{{0, 0}});
  float mo = 0.0f;
  for (auto &o : offs) mo += o[0];
  mo /= float(offs.size());
  f.offset_correct = std::abs(mo) > geom->get_FOV()[0] * 0.1f;
  return f;
}

// Context with the view table and the frequency filter
// (compute_default_frequency_filter, xray/hoCuConebeamProjectionOperator.cpp:26-62,
// when none is given).
void make_context(CtxGuard &ctx, const std::vector<float> &ta, const std::vector<float> &to, size_t nu,
                  std::shared_ptr<std::vector<float>> frequency_filter, const char *who) {
  ctx.c = acquire_context();
  check(cbct_storage(ctx.c, default_storage()), ctx.c);
  check(cbct_views_set(ctx.c, ta.data(), to.data(), (int)ta.size()), ctx.c);
  std::vector<float> F;
  if (frequency_filter) {
    F = *frequency_filter;
  } else {
    F.resize(nu + 1);
    check(cbct_default_filter((int)nu, F.data()), ctx.c);
  }
  if (F.size() != nu + 1)
    throw std::runtime_error(std::string("Error: ") + who + ": frequency filter must hold nu+1 taps");
  check(cbct_filter_set(ctx.c, (int)nu, F.data()), ctx.c);
}

cbct_geom make_geom(const CBCT_geometry &geom, const std::vector<size_t> &is_dims, floatd3 mm, size_t nu,
                    size_t nv) {
  cbct_geom g;
  g.nx = (int)is_dims[0];
  g.ny = (int)is_dims[1];
  g.nz = (int)is_dims[2];
  g.lx = mm[0];
  g.ly = mm[1];
  g.lz = mm[2];
  g.nu = (int)nu;
  g.nv = (int)nv;
  g.fov_u = geom.get_FOV()[0];
  g.fov_v = geom.get_FOV()[1];
  g.sad = geom.get_SAD();
  g.sdd = geom.get_SDD();
  return g;
}

// Planes of `src` listed in `idx` (runs of consecutive indices in one copy),
// to consecutive planes of `dst`.
void gather_planes(float *dst, const float *src, const std::vector<unsigned int> &idx, size_t plane) {
  for (size_t p = 0; p < idx.size();) {
    size_t run = 1;
    while (p + run < idx.size() && idx[p + run] == idx[p + run - 1] + 1) ++run;
    copy_d2d(dst + plane * p, src + plane * idx[p], run * plane * sizeof(float));
    p += run;
  }
}
// Consecutive planes of `src` to the planes of `dst` listed in `idx`.
void scatter_planes(float *dst, const float *src, const std::vector<unsigned int> &idx, size_t plane) {
  for (size_t p = 0; p < idx.size();) {
    size_t run = 1;
    while (p + run < idx.size() && idx[p + run] == idx[p + run - 1] + 1) ++run;
    copy_d2d(dst + plane * idx[p], src + plane * p, run * plane * sizeof(float));
    p += run;
  }
}

}  // namespace

FDKResult FDK_reconstruct_3d(std::shared_ptr<CBCT_acquisition> acquisition,
                             std::shared_ptr<CBCT_binning> binning,
                             const std::vector<size_t> &is_dims, floatd3 is_dims_in_mm,
                             std::shared_ptr<std::vector<float>> frequency_filter,
                             float samples_per_pixel) {
  if (!acquisition || !acquisition->get_projections())
    throw std::runtime_error("Error: FDK_reconstruct_3d: acquisition holds no projections");
  const hoCuNDArray<float> &proj = *acquisition->get_projections();
  if (proj.get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: projections array must be three-dimensional");
  const size_t nu = proj.get_size(0), nv = proj.get_size(1), nviews = proj.get_size(2);
  if (!binning) {
    binning = std::make_shared<CBCT_binning>();
    binning->set_as_default_3d_bin((unsigned int)nviews);
  }
  auto geom = acquisition->get_geometry();
  if (geom->get_angles().size() != nviews)
    throw std::runtime_error("Error: hoCuConebeamProjectionOperator::mult_MH: inconsistent sizes of input arrays/vectors");
  if (is_dims.size() < 3) throw std::runtime_error("Error: FDK_reconstruct_3d: image dimensions must be three or four");
  const ScanFlags fl = operator_setup(*acquisition);
  const auto &angles = geom->get_angles();
  const auto &offs = geom->get_offsets();
  const size_t nb = binning->get_number_of_bins();
  // view table: each bin's views in bin order
  std::vector<float> ta, to;
  std::vector<int> view0(nb);
  size_t maxb = 0;
  for (size_t b = 0; b < nb; ++b) {
    const auto &idx = binning->bin((unsigned)b);
    view0[b] = (int)ta.size();
    maxb = std::max(maxb, idx.size());
    for (auto i : idx) {
      if (i >= nviews) throw std::runtime_error("Error: conebeam_backwards_projection: illegal index in bin");
      ta.push_back(angles[i]);
      to.push_back(offs[i][0]);
      to.push_back(offs[i][1]);
    }
  }
  if (ta.empty()) throw std::runtime_error("Error: FDK_reconstruct_3d: the bins hold no views");
  CtxGuard ctx;
  make_context(ctx, ta, to, nu, frequency_filter, "FDK_reconstruct_3d");
  const std::vector<size_t> dims3(is_dims.begin(), is_dims.begin() + 3);
  const cbct_geom g = make_geom(*geom, dims3, is_dims_in_mm, nu, nv);
  void *st = current_stream();
  const size_t plane = nu * nv, n3 = prod(dims3);

  // the acquisition's projections, once
  cuNDArray<float> P(proj.dims());
  copy_h2d(P.get_data_ptr(), proj.get_data_ptr(), proj.get_number_of_bytes());
  auto identity = [&](const std::vector<unsigned int> &idx) {
    if (idx.size() != nviews) return false;
    for (size_t i = 0; i < idx.size(); ++i)
      if (idx[i] != i) return false;
    return true;
  };
  cuNDArray<float> Q(std::vector<size_t>{nu, nv, maxb});   // a bin's views, then their filtered copy
  cuNDArray<float> F(std::vector<size_t>{nu, nv, maxb});
  std::vector<size_t> out_dims = dims3;
  if (nb > 1) out_dims.push_back(nb);
  cuNDArray<float> img(out_dims);

  // FBP mult_MH per bin (hoCuConebeamProjectionOperator.cpp:179-266): FDK
  // pre-filter, FDK-weighted BP, image *= 2/N (:253).  An empty bin is left
  // at zero (the reference scales it by 2/0).
  for (size_t b = 0; b < nb; ++b) {
    const auto &idx = binning->bin((unsigned)b);
    float *ib = img.get_data_ptr() + b * n3;
    if (idx.empty()) {
      check(cbct_fill(ctx.c, n3, ib, 0.0f, st), ctx.c);
      continue;
    }
    const float *src = P.get_data_ptr();
    if (!identity(idx)) {
      gather_planes(Q.get_data_ptr(), P.get_data_ptr(), idx, plane);
      src = Q.get_data_ptr();
    }
    check(cbct_fdk_prefilter(ctx.c, &g, view0[b], (int)idx.size(), src, F.get_data_ptr(),
                             fl.short_scan ? 1 : 0, fl.offset_correct ? 1 : 0, st), ctx.c);
    check(cbct_bp(ctx.c, &g, view0[b], (int)idx.size(), F.get_data_ptr(), CBCT_F32, ib, 1, 0, st), ctx.c);
    check(cbct_scale(ctx.c, n3, ib, 2.0f / float(idx.size()), st), ctx.c);
  }
  // mult_M (.cpp:105-176; no offset weights in FBP mode): each bin's volume
  // into its views, in acquisition order; views in no bin stay zero, and a
  // view in several bins keeps the last bin's projection, as the host
  // operator writes them
  cuNDArray<float> A(proj.dims());
  check(cbct_fill(ctx.c, A.get_number_of_elements(), A.get_data_ptr(), 0.0f, st), ctx.c);
  for (size_t b = 0; b < nb; ++b) {
    const auto &idx = binning->bin((unsigned)b);
    if (idx.empty()) continue;
    const float *ib = img.get_data_ptr() + b * n3;
    if (identity(idx)) {
      check(cbct_fp(ctx.c, &g, view0[b], (int)idx.size(), ib, CBCT_F32, A.get_data_ptr(), samples_per_pixel,
                    0, st), ctx.c);
    } else {
      check(cbct_fp(ctx.c, &g, view0[b], (int)idx.size(), ib, CBCT_F32, F.get_data_ptr(), samples_per_pixel,
                    0, st), ctx.c);
      scatter_planes(A.get_data_ptr(), F.get_data_ptr(), idx, plane);
    }
  }
  // least-squares rescale (FDK_reconstruct_3d.cpp:150-156), FP64 reductions
  FDKResult r;
  r.scale = dot(&A, &P) / dot(&A, &A);
  check(cbct_scale(ctx.c, img.get_number_of_elements(), img.get_data_ptr(), float(r.scale), st), ctx.c);
  r.image = std::make_shared<hoCuNDArray<float>>(out_dims, no_init);
  img.to_host(r.image.get());
  return r;
}

FDK4DResult FDK_reconstruct_4d(std::shared_ptr<CBCT_acquisition> acquisition,
                               std::shared_ptr<CBCT_binning> binning,
                               const std::vector<size_t> &is_dims, floatd3 is_dims_in_mm,
                               std::shared_ptr<std::vector<float>> frequency_filter,
                               float samples_per_pixel) {
  if (!acquisition || !acquisition->get_projections())
    throw std::runtime_error("Error: FDK_reconstruct_4d: acquisition holds no projections");
  if (!binning || binning->get_number_of_bins() == 0)
    throw std::runtime_error("Error: FDK_reconstruct_4d: binning not provided");
  if (is_dims.size() != 3) throw std::runtime_error("Error: FDK_reconstruct_4d: image dimensions must be three");
  auto geom = acquisition->get_geometry();
  const hoCuNDArray<float> &proj = *acquisition->get_projections();
  if (proj.get_number_of_dimensions() != 3)
    throw std::runtime_error("Error: FDK_reconstruct_4d: projections array must be three-dimensional");
  const size_t nu = proj.get_size(0), nv = proj.get_size(1), nviews = proj.get_size(2);
  std::vector<float> &angles = geom->get_angles();
  if (angles.size() != nviews)
    throw std::runtime_error("Error: FDK_reconstruct_4d: inconsistent sizes of input arrays/vectors");

  const ScanFlags fl = operator_setup(*acquisition);
  const bool short_scan = fl.short_scan, offset_correct = fl.offset_correct;
  auto &offs = geom->get_offsets();

  // View table: the union of the bins (CBCT_binning::get_3d_binning,
  // xray/CBCT_binning.h:95-104, sorted), then every bin in its own order.
  auto b3 = binning->get_3d_binning();
  const std::vector<unsigned int> &uni = b3->bin(0);
  const size_t nb = binning->get_number_of_bins(), nun = uni.size();
  if (nun == 0) throw std::runtime_error("Error: FDK_reconstruct_4d: the bins hold no views");
  for (auto idx : uni)
    if (idx >= nviews) throw std::runtime_error("Error: conebeam_backwards_projection: illegal index in bin");
  std::vector<int> pos(nviews, -1);
  for (size_t i = 0; i < nun; ++i) pos[uni[i]] = (int)i;
  std::vector<float> ta, to;
  auto push_view = [&](unsigned int i) {
    ta.push_back(angles[i]);
    to.push_back(offs[i][0]);
    to.push_back(offs[i][1]);
  };
  for (auto i : uni) push_view(i);
  std::vector<size_t> bin_off(nb);    // first plane of each bin in the gathered array
  size_t nbv = 0;
  for (size_t b = 0; b < nb; ++b) {
    bin_off[b] = nbv;
    for (auto i : binning->bin((unsigned)b)) push_view(i);
    nbv += binning->bin((unsigned)b).size();
  }

  CtxGuard ctx;
  make_context(ctx, ta, to, nu, frequency_filter, "FDK_reconstruct_4d");
  const cbct_geom g = make_geom(*geom, is_dims, is_dims_in_mm, nu, nv);
  void *st = current_stream();
  const size_t plane = nu * nv, n3 = prod(is_dims);

  // stage the union's projections once (runs of consecutive views in one copy)
  cuNDArray<float> U(std::vector<size_t>{nu, nv, nun});
  for (size_t p = 0; p < nun;) {
    size_t run = 1;
    while (p + run < nun && uni[p + run] == uni[p + run - 1] + 1) ++run;
    copy_h2d(U.get_data_ptr() + plane * p, proj.get_data_ptr() + plane * uni[p], run * plane * sizeof(float));
    p += run;
  }
  cuNDArray<float> Q(std::vector<size_t>{nu, nv, std::max(nun, nbv)});

  // 3-D FDK over the union (FDK_reconstruct_4d.cpp:124-130): pre-filter,
  // FDK-weighted BP, image *= 2/N (hoCuConebeamProjectionOperator.cpp:253)
  FDK4DResult r;
  cuNDArray<float> f(is_dims);
  check(cbct_fdk_prefilter(ctx.c, &g, 0, (int)nun, U.get_data_ptr(), Q.get_data_ptr(), short_scan ? 1 : 0,
                           offset_correct ? 1 : 0, st), ctx.c);
  check(cbct_bp(ctx.c, &g, 0, (int)nun, Q.get_data_ptr(), CBCT_F32, f.get_data_ptr(), 1, 0, st), ctx.c);
  check(cbct_scale(ctx.c, n3, f.get_data_ptr(), 2.0f / float(nun), st), ctx.c);

  // A f over the union of the bins' views (:149-150).  The reference's
  // diff_proj is zero for views in no bin; those add nothing to the inner
  // products and are never backprojected.
  cuNDArray<float> D(std::vector<size_t>{nu, nv, nun});
  check(cbct_fp(ctx.c, &g, 0, (int)nun, f.get_data_ptr(), CBCT_F32, D.get_data_ptr(), samples_per_pixel, 0, st),
        ctx.c);
  r.scale = dot(&D, &U) / dot(&D, &D);
  const float s = float(r.scale);                        // float scaler (:151)
  check(cbct_scale(ctx.c, plane * nun, D.get_data_ptr(), s, st), ctx.c);                  // diff_proj *= scaler
  check(cbct_sub(ctx.c, plane * nun, D.get_data_ptr(), U.get_data_ptr(), st), ctx.c);     // projections -= diff_proj

  // 4-D FBP of the residual (:158-160): bins gathered in bin order, one
  // pre-filter launch over all of them, one BP per bin into its volume.
  for (size_t b = 0; b < nb; ++b) {
    const auto &idx = binning->bin((unsigned)b);
    for (size_t p = 0; p < idx.size();) {
      size_t run = 1;
      while (p + run < idx.size() && pos[idx[p + run]] == pos[idx[p + run - 1]] + 1) ++run;
      copy_d2d(Q.get_data_ptr() + plane * (bin_off[b] + p), U.get_data_ptr() + plane * pos[idx[p]],
               run * plane * sizeof(float));
      p += run;
    }
  }
  if (nbv > 0)
    check(cbct_fdk_prefilter(ctx.c, &g, (int)nun, (int)nbv, Q.get_data_ptr(), Q.get_data_ptr(),
                             short_scan ? 1 : 0, offset_correct ? 1 : 0, st), ctx.c);
  std::vector<size_t> dims4 = is_dims;
  dims4.push_back(nb);
  cuNDArray<float> R(dims4);
  for (size_t b = 0; b < nb; ++b) {
    float *rb = R.get_data_ptr() + b * n3;
    const size_t n = binning->bin((unsigned)b).size();
    if (n == 0) {
      // an empty bin's backprojection is left at zero (the reference scales
      // it by 2/0)
      check(cbct_fill(ctx.c, n3, rb, 0.0f, st), ctx.c);
    } else {
      check(cbct_bp(ctx.c, &g, (int)(nun + bin_off[b]), (int)n, Q.get_data_ptr() + plane * bin_off[b], CBCT_F32,
                    rb, 1, 0, st), ctx.c);
      check(cbct_scale(ctx.c, n3, rb, 2.0f / float(n), st), ctx.c);
    }
    check(cbct_add(ctx.c, n3, f.get_data_ptr(), rb, st), ctx.c);   // result += fdk
  }
  check(cbct_scale(ctx.c, n3 * nb, R.get_data_ptr(), s, st), ctx.c);   // result *= scaler

  r.image = std::make_shared<hoCuNDArray<float>>(dims4);
  r.fdk_3d = std::make_shared<hoCuNDArray<float>>(is_dims);
  R.to_host(r.image.get());
  f.to_host(r.fdk_3d.get());
  return r;
}

}  // namespace gtt
