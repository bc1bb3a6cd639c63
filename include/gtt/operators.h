// Operator surface of the reference's reconstruction path:
//   linearOperator / subsetOperator          (Gadgetron toolbox bases, not vendored
//                                             in the reference; semantics from the
//                                             call sites listed per method)
//   cuConebeamProjectionOperator             xray/cuConebeamProjectionOperator.{h,cpp}
//   hoCuConebeamProjectionOperator           xray/hoCuConebeamProjectionOperator.{h,cpp}
//   CBSubsetOperator<cuNDArray>              xray/CBSubsetOperator.{h,cpp}
//   primalDualOperator / cuTVPrimalDualOperator  operators/primalDualOperator.h,
//                                             operators/cuTVPrimalDualOperator.{h,cu}
// Every projector call goes to the gfx950 kernels of libcbct.
#pragma once

#include <memory>
#include <vector>

#include "gtt/CBCT_acquisition.h"
#include "gtt/ndarray.h"

namespace gtt {

template <class ARRAY>
class linearOperator {
 public:
  typedef typename ARRAY::element_type ELEMENT_TYPE;
  virtual ~linearOperator() = default;
  virtual void mult_M(ARRAY *in, ARRAY *out, bool accumulate = false) = 0;
  virtual void mult_MH(ARRAY *in, ARRAY *out, bool accumulate = false) = 0;
  virtual void mult_MH_M(ARRAY *in, ARRAY *out, bool accumulate = false) {
    ARRAY tmp(get_codomain_dimensions());
    mult_M(in, &tmp, false);
    mult_MH(&tmp, out, accumulate);
  }
  virtual void set_domain_dimensions(const std::vector<size_t> *d) { domain_ = *d; }
  virtual void set_codomain_dimensions(const std::vector<size_t> *d) { codomain_ = *d; }
  virtual std::shared_ptr<std::vector<size_t>> get_domain_dimensions() {
    return std::make_shared<std::vector<size_t>>(domain_);
  }
  virtual std::shared_ptr<std::vector<size_t>> get_codomain_dimensions() {
    return std::make_shared<std::vector<size_t>>(codomain_);
  }
  virtual void set_weight(float w) { weight_ = w; }
  virtual float get_weight() { return weight_; }

 protected:
  std::vector<size_t> domain_, codomain_;
  float weight_ = 1.0f;
};

// Ordered-subset operator base.  Full mult_M fills every subset's contiguous
// block of the codomain; full mult_MH clears the output (unless accumulating)
// and adds every subset's backprojection (call sites:
// solvers/osSARTSolver.h:62,112; solvers/osMOMSolverD.h:117-118).
template <class ARRAY>
class subsetOperator : public linearOperator<ARRAY> {
 public:
  explicit subsetOperator(int subsets = 1) : number_of_subsets(subsets) {}
  virtual void mult_M(ARRAY *in, ARRAY *out, int subset, bool accumulate) = 0;
  virtual void mult_MH(ARRAY *in, ARRAY *out, int subset, bool accumulate) = 0;
  virtual void mult_MH_M(ARRAY *in, ARRAY *out, int subset, bool accumulate) {
    ARRAY tmp(get_codomain_dimensions(subset));
    mult_M(in, &tmp, subset, false);
    mult_MH(&tmp, out, subset, accumulate);
  }
  void mult_M(ARRAY *in, ARRAY *out, bool accumulate = false) override {
    auto subs = projection_subsets(out);
    for (int i = 0; i < get_number_of_subsets(); ++i) mult_M(in, subs[i].get(), i, accumulate);
  }
  void mult_MH(ARRAY *in, ARRAY *out, bool accumulate = false) override {
    auto subs = projection_subsets(in);
    if (!accumulate) clear(out);
    for (int i = 0; i < get_number_of_subsets(); ++i) mult_MH(subs[i].get(), out, i, true);
  }
  void mult_MH_M(ARRAY *in, ARRAY *out, bool accumulate = false) override {
    ARRAY tmp(this->get_codomain_dimensions());
    mult_M(in, &tmp, false);
    mult_MH(&tmp, out, accumulate);
  }
  virtual std::shared_ptr<std::vector<size_t>> get_codomain_dimensions(int subset) = 0;
  std::shared_ptr<std::vector<size_t>> get_codomain_dimensions() override {
    // all subsets stacked along the last dimension
    auto d = *get_codomain_dimensions(0);
    size_t n = 0;
    for (int i = 0; i < get_number_of_subsets(); ++i) n += get_codomain_dimensions(i)->back();
    d.back() = n;
    return std::make_shared<std::vector<size_t>>(d);
  }
  virtual int get_number_of_subsets() { return number_of_subsets; }
  // Views onto the contiguous per-subset blocks of a projection array.
  virtual std::vector<std::shared_ptr<ARRAY>> projection_subsets(ARRAY *projections) {
    std::vector<std::shared_ptr<ARRAY>> res;
    auto *p = projections->get_data_ptr();
    for (int s = 0; s < get_number_of_subsets(); ++s) {
      auto d = *get_codomain_dimensions(s);
      res.push_back(std::make_shared<ARRAY>(d, p));
      p += prod(d);
    }
    return res;
  }

 protected:
  int number_of_subsets;
};

// ---------------------------------------------------------------------------
// Device-resident cone-beam projector (xray/cuConebeamProjectionOperator.cpp)
// ---------------------------------------------------------------------------
class cuConebeamProjectionOperator : public linearOperator<cuNDArray<float>> {
 public:
  cuConebeamProjectionOperator();
  ~cuConebeamProjectionOperator() override;
  cuConebeamProjectionOperator(const cuConebeamProjectionOperator &) = delete;
  cuConebeamProjectionOperator &operator=(const cuConebeamProjectionOperator &) = delete;

  // :38-80 -- per non-empty bin, FP of that bin's volume into its views
  void mult_M(cuNDArray<float> *in, cuNDArray<float> *out, bool accumulate = false) override;
  // :82-116 -- per non-empty bin, BP (full offset weights on a copy if enabled)
  void mult_MH(cuNDArray<float> *in, cuNDArray<float> *out, bool accumulate = false) override;

  // :148-206 -- shifts angles (unless told not to), detects short scans,
  // auto-enables offset correction, builds per-bin view tables and permutes
  // the acquisition's projections into bin order.
  void setup(std::shared_ptr<CBCT_acquisition> acquisition, floatd3 is_dims_in_mm,
             bool transform_angles = true);
  void setup(std::shared_ptr<CBCT_acquisition> acquisition, std::shared_ptr<CBCT_binning> binning,
             floatd3 is_dims_in_mm, bool transform_angles = true);
  // sqrt half-fan weights on all views (:22-32)
  void offset_correct(cuNDArray<float> *projections);
  bool get_use_offset_correction() const { return use_offset_correction_; }
  void set_use_offset_correction(bool use) {
    use_offset_correction_ = use;
    allow_offset_correction_override_ = false;
  }
  void set_num_samples_per_pixel(float s) { samples_per_pixel_ = s; }
  void set_mask(std::shared_ptr<cuNDArray<unsigned char>> m) { mask_ = std::move(m); }
  // :118-146 -- space-carving mask per bin
  std::shared_ptr<cuNDArray<unsigned char>> calculate_mask(cuNDArray<float> *projections, float limit);
  bool is_short_scan() const { return short_scan_; }
  cbct_ctx context() const { return ctx_; }
  // Working-copy storage of the projectors (CBCT_F32 / CBCT_BF16); new
  // operators take the creating thread's default_storage().
  void set_storage(int dtype) { check(cbct_storage(ctx_, dtype), ctx_); }
  // Tunes the forward projector for this operator's geometry (domain
  // dimensions set, detector nu x nv) on its largest bin (cbct_fp_autotune).
  // No reference counterpart.
  void autotune(size_t nu, size_t nv);

 protected:
  cbct_geom geom_for(size_t nx, size_t ny, size_t nz, size_t nu, size_t nv) const;
  cbct_ctx ctx_ = nullptr;
  std::shared_ptr<CBCT_acquisition> acquisition_;
  std::shared_ptr<CBCT_binning> binning_;
  std::vector<std::vector<float>> angles_;
  std::vector<std::vector<floatd2>> offsets_;
  std::vector<int> bin_view0_;   // offset of each bin in the context's view table
  int all_view0_ = 0;            // offset of the acquisition-order table
  floatd3 is_dims_in_mm_{{0, 0, 0}};
  float samples_per_pixel_ = 1.5f;
  bool preprocessed_ = false;
  bool short_scan_ = false;
  bool use_offset_correction_ = false;
  bool allow_offset_correction_override_ = true;
  std::shared_ptr<cuNDArray<unsigned char>> mask_;
};

// ---------------------------------------------------------------------------
// Host-resident cone-beam projector with FDK mode
// (xray/hoCuConebeamProjectionOperator.{h,cpp})
// ---------------------------------------------------------------------------
class hoCuConebeamProjectionOperator : public linearOperator<hoCuNDArray<float>> {
 public:
  hoCuConebeamProjectionOperator();
  ~hoCuConebeamProjectionOperator() override;
  hoCuConebeamProjectionOperator(const hoCuConebeamProjectionOperator &) = delete;
  hoCuConebeamProjectionOperator &operator=(const hoCuConebeamProjectionOperator &) = delete;

  // .cpp:105-176
  void mult_M(hoCuNDArray<float> *in, hoCuNDArray<float> *out, bool accumulate = false) override;
  // .cpp:179-266 (FBP branch: cosine weights, [Parker], ramp filter,
  // [offset weights], FDK-weighted BP, image *= 2/N)
  void mult_MH(hoCuNDArray<float> *in, hoCuNDArray<float> *out, bool accumulate = false) override;
  void offset_correct(hoCuNDArray<float> *projections);
  // .h:39-93
  void setup(std::shared_ptr<CBCT_acquisition> acquisition, floatd3 is_dims_in_mm,
             bool transform_angles = true);
  void setup(std::shared_ptr<CBCT_acquisition> acquisition, std::shared_ptr<CBCT_binning> binning,
             floatd3 is_dims_in_mm, bool transform_angles = true);
  void set_use_filtered_backprojection(bool on) { use_fbp_ = on; }
  void set_use_offset_correction(bool use) {
    use_offset_correction_ = use;
    allow_offset_correction_override_ = false;
  }
  bool get_use_offset_correction() const { return use_offset_correction_; }
  void set_num_projections_per_batch(unsigned int n) { projections_per_batch_ = n; }
  void set_num_samples_per_pixel(float s) { samples_per_pixel_ = s; }
  // One-sided frequency response, nu+1 taps (.h:117-119)
  void set_frequency_filter(std::shared_ptr<std::vector<float>> weights) {
    frequency_filter_ = std::move(weights);
    filter_uploaded_ = false;
  }
  void set_acquisition(std::shared_ptr<CBCT_acquisition> a) { acquisition_ = std::move(a); }
  std::shared_ptr<CBCT_acquisition> get_acquisition() { return acquisition_; }
  void set_binning(std::shared_ptr<CBCT_binning> b);
  std::shared_ptr<CBCT_binning> get_binning() { return binning_; }
  bool is_short_scan() const { return short_scan_; }
  // Working-copy storage of the projectors (CBCT_F32 / CBCT_BF16)
  void set_storage(int dtype) { check(cbct_storage(ctx_, dtype), ctx_); }

 protected:
  void build_view_table();
  cbct_geom geom_for(size_t nx, size_t ny, size_t nz, size_t nu, size_t nv) const;
  cbct_ctx ctx_ = nullptr;
  std::shared_ptr<CBCT_acquisition> acquisition_;
  std::shared_ptr<CBCT_binning> binning_;
  std::vector<int> bin_view0_;   // bins laid out after the acquisition-order views
  floatd3 is_dims_in_mm_{{0, 0, 0}};
  float samples_per_pixel_ = 1.5f;
  unsigned int projections_per_batch_ = 20;
  bool use_fbp_ = false;
  bool short_scan_ = false;
  bool preprocessed_ = false;
  bool use_offset_correction_ = false;
  bool allow_offset_correction_override_ = true;
  bool filter_uploaded_ = false;
  std::shared_ptr<std::vector<float>> frequency_filter_;
};

// ---------------------------------------------------------------------------
// Ordered subsets over device-resident projectors (xray/CBSubsetOperator.cpp)
// ---------------------------------------------------------------------------
class CBSubsetOperator : public subsetOperator<cuNDArray<float>> {
 public:
  explicit CBSubsetOperator(int subsets = 1);
  void mult_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool accumulate) override;
  void mult_MH(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool accumulate) override;
  void mult_MH_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool accumulate) override;
  using subsetOperator<cuNDArray<float>>::mult_M;
  using subsetOperator<cuNDArray<float>>::mult_MH;
  using subsetOperator<cuNDArray<float>>::mult_MH_M;
  std::shared_ptr<std::vector<size_t>> get_codomain_dimensions(int subset) override;
  using subsetOperator<cuNDArray<float>>::get_codomain_dimensions;
  void set_domain_dimensions(const std::vector<size_t> *d) override {
    linearOperator<cuNDArray<float>>::set_domain_dimensions(d);
    for (auto &op : operators_) op->set_domain_dimensions(d);
  }
  // :52-158 -- round-robin split, permutation of projections/angles/offsets,
  // per-subset projector setup (transform_angles = false).  With
  // permute_projections = false (no reference counterpart) the acquisition's
  // host projections are left in acquisition order, for callers that upload
  // them once and permute on the device (upload_in_subset_order).
  void setup(std::shared_ptr<CBCT_acquisition> acq, std::shared_ptr<CBCT_binning> binning,
             floatd3 is_dims_in_mm, bool permute_projections = true);
  void setup(std::shared_ptr<CBCT_acquisition> acq, floatd3 is_dims_in_mm);
  void offset_correct(cuNDArray<float> *projections);
  void set_use_offset_correction(bool use) {
    for (auto &op : operators_) op->set_use_offset_correction(use);
  }
  void set_mask(std::shared_ptr<cuNDArray<unsigned char>> mask) {
    for (auto &op : operators_) op->set_mask(mask);
  }
  void set_storage(int dtype) {
    for (auto &op : operators_) op->set_storage(dtype);
  }
  // Tunes the forward projector for the subsets' geometry (after
  // set_domain_dimensions); the choice is shared by every operator with the
  // same geometry.  No reference counterpart.
  void autotune() {
    if (!operators_.empty() && !projection_dims_.empty())
      operators_[0]->autotune(projection_dims_[0]->at(0), projection_dims_[0]->at(1));
  }
  const std::vector<unsigned int> &get_permutation() const { return permutations_; }
  std::shared_ptr<cuConebeamProjectionOperator> get_subset_operator(int s) { return operators_[s]; }

 protected:
  std::vector<std::shared_ptr<cuConebeamProjectionOperator>> operators_;
  std::vector<std::shared_ptr<std::vector<size_t>>> projection_dims_;
  std::vector<unsigned int> permutations_;
};

// Host projections in acquisition order, uploaded once and put in the
// subset operator's order (its permutation) on the device.  No reference
// counterpart: the reference permutes on the host (CBSubsetOperator.cpp:106,
// :192-212) and uploads the permuted array.
cuNDArray<float> upload_in_subset_order(const hoCuNDArray<float> &projections,
                                        const std::vector<unsigned int> &permutation);

// ---------------------------------------------------------------------------
// Per-projection statistical weights around the subset projector
// (xray/CBSubsetWeightOperator.h:11-63): M_w = W A, M_w^H = A^H W, with W
// the weights in the subsets' permuted order.
// ---------------------------------------------------------------------------
class CBSubsetWeightOperator : public CBSubsetOperator {
 public:
  explicit CBSubsetWeightOperator(int subsets = 1) : CBSubsetOperator(subsets) {}
  // :17-31 -- the acquisition's projections are multiplied by the weights
  // (in place, host), the subsets are set up, and the weights are permuted
  // like the projections and kept on the device.
  void setup(std::shared_ptr<CBCT_acquisition> acq, std::shared_ptr<CBCT_binning> binning,
             floatd3 is_dims_in_mm, std::shared_ptr<hoCuNDArray<float>> projection_weights);
  using CBSubsetOperator::setup;
  // :33-43 -- the subset's projection times its weights
  void mult_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool accumulate) override;
  // :44-49 -- the input times the subset's weights, backprojected
  void mult_MH(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool accumulate) override;
  // :50-57
  void mult_MH_M(cuNDArray<float> *in, cuNDArray<float> *out, int subset, bool accumulate) override;
  using CBSubsetOperator::mult_M;
  using CBSubsetOperator::mult_MH;
  using CBSubsetOperator::mult_MH_M;
  const cuNDArray<float> &get_weights() const { return weights_; }

 protected:
  cuNDArray<float> weights_;
  std::vector<std::shared_ptr<cuNDArray<float>>> subset_weights_;
};

// ---------------------------------------------------------------------------
// Primal-dual regularisers (operators/primalDualOperator.h,
// operators/cuTVPrimalDualOperator.cu)
// ---------------------------------------------------------------------------
template <class ARRAY>
class primalDualOperator {
 public:
  virtual ~primalDualOperator() = default;
  virtual void primalDual(ARRAY *in, ARRAY *out, float sigma = 0, bool accumulate = false) = 0;
  virtual void set_weight(float w) { weight = w; }
  virtual float get_weight() { return weight; }

 protected:
  float weight = 1.0f;
};

class cuTVPrimalDualOperator : public primalDualOperator<cuNDArray<float>> {
 public:
  cuTVPrimalDualOperator() : alpha(0.0f) {}
  explicit cuTVPrimalDualOperator(float alpha_) : alpha(alpha_) {}
  void primalDual(cuNDArray<float> *in, cuNDArray<float> *out, float sigma, bool accumulate) override;
  float get_alpha() const { return alpha; }

 private:
  float alpha;
};

}  // namespace gtt
