// Ordered-subset solvers over the cone-beam subset operator:
//   osSARTSolver   solvers/osSARTSolver.h:11-137
//   osMOMSolverD   solvers/osMOMSolverD.h:11-255 (base-class updateG)
// The element-wise passes are fused libcbct kernels; the TV proximal step
// optionally runs as one fused TV+updateG kernel (bitwise equal to the
// unfused sequence).
#pragma once

#include <cmath>
#include <cstdio>
#include <functional>
#include <iostream>
#include <memory>
#include <sstream>
#include <vector>

#include "gtt/io.h"
#include "gtt/operators.h"

namespace gtt {

template <class ARRAY>
class solverBase {
 public:
  enum { OUTPUT_SILENT = 0, OUTPUT_WARNINGS = 1, OUTPUT_VERBOSE = 2 };
  virtual ~solverBase() = default;
  void set_output_mode(int m) { output_mode_ = m; }
  void set_x0(std::shared_ptr<ARRAY> x0) { x0_ = std::move(x0); }

 protected:
  int output_mode_ = OUTPUT_SILENT;
  std::shared_ptr<ARRAY> x0_;
};

template <class ARRAY>
class osSARTSolver : public solverBase<ARRAY> {
  typedef typename ARRAY::element_type ELEMENT_TYPE;

 public:
  void set_max_iterations(int i) { iterations_ = i; }
  int get_max_iterations() const { return iterations_; }
  void set_non_negativity_constraint(bool neg = true) { non_negativity_ = neg; }
  void set_beta(float b) { beta_ = b; }
  void set_gamma(float g) { gamma_ = g; }
  // The reference always evaluates ||Ax-b||^2 and writes osSART-<i>.real
  // after every iteration (:110-120); both are optional here.
  void set_objective_callback(std::function<void(int, double)> cb) { objective_cb_ = std::move(cb); }
  void set_dump(bool d) { dump_ = d; }
  // Replay iterations 2.. as captured HIP graphs (one per subset order; the
  // first iteration runs directly and sizes every workspace).  Applies when
  // gamma = 0 (the step size is then the same every iteration) and the
  // output is not verbose (that reads a norm back per subset); runs on a
  // stream of its own when the thread has none.  Same results bit for bit.
  void set_graphs(bool g) { graphs_ = g; }
  void set_encoding_operator(std::shared_ptr<subsetOperator<ARRAY>> op) { encoding_operator_ = std::move(op); }

  std::shared_ptr<ARRAY> solve(ARRAY *in) {
    if (!encoding_operator_) throw std::runtime_error("Error: cgSolver::compute_rhs : no encoding operator is set");
    const bool graphs = graphs_ && gamma_ == 0.0f && iterations_ > 1 &&
                        this->output_mode_ < solverBase<ARRAY>::OUTPUT_VERBOSE;
    PrivateStream own_stream(graphs && current_stream() == nullptr);
    auto image_dims = encoding_operator_->get_domain_dimensions();
    if (image_dims->empty())
      throw std::runtime_error("Error: cgSolver::compute_rhs : encoding operator has not set domain dimension");
    auto x = std::make_shared<ARRAY>(image_dims);
    if (this->x0_) *x = *this->x0_;
    else clear(x.get());

    auto subsets = encoding_operator_->projection_subsets(in);
    ARRAY ones_projection(in->get_dimensions());
    ARRAY tmp_image(image_dims);
    fill(&tmp_image, ELEMENT_TYPE(1));
    encoding_operator_->mult_M(&tmp_image, &ones_projection, false);
    clamp_min(&ones_projection, ELEMENT_TYPE(1e-6));
    reciprocal_inplace(&ones_projection);
    auto ones_projections = encoding_operator_->projection_subsets(&ones_projection);

    ARRAY tmp_projection(in->get_dimensions());
    fill(&tmp_projection, ELEMENT_TYPE(1));
    auto tmp_projections = encoding_operator_->projection_subsets(&tmp_projection);

    const int S = encoding_operator_->get_number_of_subsets();
    std::vector<ARRAY> ones_images;
    ones_images.reserve(S);
    for (int i = 0; i < S; ++i) {
      ones_images.emplace_back(image_dims);
      encoding_operator_->mult_MH(tmp_projections[i].get(), &ones_images.back(), i, false);
      clamp_min(&ones_images.back(), ELEMENT_TYPE(1e-6));
      reciprocal_inplace(&ones_images.back());
    }
    if (this->output_mode_ >= solverBase<ARRAY>::OUTPUT_VERBOSE)
      std::cout << "osSART setup done, starting iterations:" << std::endl;

    std::vector<int> isubsets(S);
    for (int i = 0; i < S; ++i) isubsets[i] = i;
    auto iteration = [&](int it) {
      for (int is = 0; is < S; ++is) {
        const int s = isubsets[is];
        encoding_operator_->mult_M(x.get(), tmp_projections[s].get(), s, false);
        // tmp = -(A_s x - b_s) .* W_p   (:92-98)
        sart_residual(tmp_projections[s].get(), subsets[s].get(), ones_projections[s].get());
        if (this->output_mode_ >= solverBase<ARRAY>::OUTPUT_VERBOSE)
          std::cout << "Iteration " << it << " Subset " << s << " Update norm: "
                    << nrm2(tmp_projections[s].get()) << std::endl;
        encoding_operator_->mult_MH(tmp_projections[s].get(), &tmp_image, s, false);
        // x += beta/(1+gamma i) * tmp_image .* W_x ; clamp   (:99-105)
        sart_update(x.get(), &tmp_image, &ones_images[s], float(beta_ / (1 + gamma_ * it)),
                    non_negativity_);
      }
    };
    std::unique_ptr<StreamGraph> graph[2];   // subsets in order, reversed
    for (int it = 0; it < iterations_; ++it) {
      if (graphs && it > 0) {
        auto &gr = graph[isubsets.front() == 0 ? 0 : 1];
        if (!gr) gr = std::make_unique<StreamGraph>([&] { iteration(it); });
        gr->launch();
      } else {
        iteration(it);
      }
      std::reverse(isubsets.begin(), isubsets.end());
      if (objective_cb_) {
        ARRAY tmp_proj(in->get_dimensions());
        encoding_operator_->mult_M(x.get(), &tmp_proj, false);
        tmp_proj -= *in;
        objective_cb_(it, dot(&tmp_proj, &tmp_proj));
      }
      if (dump_) {
        std::stringstream ss;
        ss << "osSART-" << it << ".real";
        write_nd_array(x.get(), ss.str());
      }
    }
    return x;
  }

 protected:
  int iterations_ = 10;
  float beta_ = 1.0f, gamma_ = 0.0f;
  bool non_negativity_ = false;
  bool dump_ = false;
  bool graphs_ = false;
  std::function<void(int, double)> objective_cb_;
  std::shared_ptr<subsetOperator<ARRAY>> encoding_operator_;
};

template <class ARRAY>
class osMOMSolverD : public solverBase<ARRAY> {
  typedef typename ARRAY::element_type ELEMENT_TYPE;

 public:
  void set_max_iterations(int i) { iterations_ = i; }
  int get_max_iterations() const { return iterations_; }
  void set_non_negativity_constraint(bool neg = true) { non_negativity_ = neg; }
  void set_beta(float b) { beta_ = b; }
  void set_gamma(float g) { gamma_ = g; }
  void set_kappa(float k) { kappa_ = k; }
  void set_dump(bool d) { dump_ = d; }
  void set_tau(float t) { tau0_ = t; }
  void set_huber(float h) { denoise_alpha_ = h; }
  void set_reg_steps(unsigned int n) { reg_steps_ = n; }
  void set_preconditioning_image(std::shared_ptr<ARRAY> p) { preconditioning_image_ = std::move(p); }
  void set_encoding_operator(std::shared_ptr<subsetOperator<ARRAY>> op) { encoding_operator_ = std::move(op); }
  void add_regularization_operator(std::shared_ptr<primalDualOperator<ARRAY>> op) {
    regularization_operators_.push_back(std::move(op));
  }
  // Run a single TV regulariser as one fused TV+updateG kernel (same bits).
  void set_fuse_tv(bool f) { fuse_tv_ = f; }

  std::shared_ptr<ARRAY> solve(ARRAY *in) {
    if (!encoding_operator_) throw std::runtime_error("Error: cgSolver::compute_rhs : no encoding operator is set");
    auto image_dims = encoding_operator_->get_domain_dimensions();
    if (image_dims->empty())
      throw std::runtime_error("Error: cgSolver::compute_rhs : encoding operator has not set domain dimension");
    auto z = std::make_shared<ARRAY>(image_dims);
    if (this->x0_) *z = *this->x0_;
    else clear(z.get());
    ARRAY xa(*z), xb(*z);
    ARRAY *x = &xa, *xold = &xb;

    auto subsets = encoding_operator_->projection_subsets(in);
    ARRAY tmp_projection(in->get_dimensions());
    auto tmp_projections = encoding_operator_->projection_subsets(&tmp_projection);

    std::shared_ptr<ARRAY> precon;
    if (preconditioning_image_) {
      precon = preconditioning_image_;
    } else {
      precon = std::make_shared<ARRAY>(image_dims);
      fill(precon.get(), ELEMENT_TYPE(1));
      encoding_operator_->mult_M(precon.get(), &tmp_projection, false);
      encoding_operator_->mult_MH(&tmp_projection, precon.get(), false);
      clamp_min(precon.get(), ELEMENT_TYPE(1e-6));
      reciprocal_inplace(precon.get());
    }

    float avg_lambda = 0;
    for (auto &op : regularization_operators_) avg_lambda += op->get_weight() / regularization_operators_.size();
    for (auto &op : regularization_operators_) op->set_weight(op->get_weight() / avg_lambda);

    float t = 1, told = 1;
    if (this->output_mode_ >= solverBase<ARRAY>::OUTPUT_VERBOSE)
      std::cout << "osMOM setup done, starting iterations:" << std::endl;
    const int S = encoding_operator_->get_number_of_subsets();
    ARRAY tmp_image(image_dims);
    ARRAY g(image_dims);
    // gamma of the step-size rule depends only on the preconditioner (:206)
    const double asum_p = regularization_operators_.empty() ? 0.0 : asum(precon.get());
    for (int it = 0; it < iterations_; ++it) {
      for (int s = 0; s < S; ++s) {
        t = float(0.5 * (1 + std::sqrt(1 + 4 * double(t) * t)));
        encoding_operator_->mult_M(x, tmp_projections[s].get(), s, false);
        *tmp_projections[s] -= *subsets[s];
        if (this->output_mode_ >= solverBase<ARRAY>::OUTPUT_VERBOSE)
          std::cout << "Iteration " << it << " Subset " << s << " Update norm: "
                    << nrm2(tmp_projections[s].get()) << std::endl;
        encoding_operator_->mult_MH(tmp_projections[s].get(), &tmp_image, s, false);
        // z -= beta/(1+gamma i) * S * tmp_image .* precon   (:155-157)
        mom_gradient_step(z.get(), &tmp_image, precon.get(),
                          -float(beta_ / (1 + gamma_ * it)) * float(S));
        denoise(x, &g, *z, *precon, avg_lambda, asum_p);
        // nonneg clamp and momentum (:166-175)
        const float c = (told - 1) / t;
        mom_momentum(x, xold, z.get(), c, non_negativity_);
        std::swap(x, xold);
        *x = *z;
        told = t;
      }
      if (dump_) {
        std::stringstream ss;
        ss << "osMOM-" << it << ".real";
        write_nd_array(x, ss.str());
      }
    }
    for (auto &op : regularization_operators_) op->set_weight(op->get_weight() * avg_lambda);
    return z;
  }

 protected:
  // solvers/osMOMSolverD.h:205-230.  The fused path swaps the buffers of *x and *g.
  void denoise(ARRAY *x, ARRAY *g, ARRAY &s, ARRAY &precon, float avg_lambda, double asum_p) {
    if (regularization_operators_.empty()) {
      *x = s;
      return;
    }
    const float gam = float(0.35 / avg_lambda / (precon.get_number_of_elements() / asum_p));
    const float L = 4;
    float tau = tau0_;
    float sigma = 1 / (tau * L * L);
    auto *tv = regularization_operators_.size() == 1
                   ? dynamic_cast<cuTVPrimalDualOperator *>(regularization_operators_[0].get())
                   : nullptr;
    for (unsigned int k = 0; k < reg_steps_; ++k) {
      if (fuse_tv_ && tv) {
        fused_tv_update(tv, x, g, s, precon, sigma, tau, avg_lambda);
      } else {
        clear(g);
        for (auto &op : regularization_operators_) op->primalDual(x, g, sigma, true);
        updateG(*g, *x, s, precon, tau, avg_lambda);
      }
      const float theta = float(1 / std::sqrt(1 + 2 * double(gam) * tau));
      tau *= theta;
      sigma /= theta;
    }
  }
  // base-class updateG (:233-247):
  // x = (x - tau g + (tau/lambda) s/precon) / (1 + tau/(lambda precon))
  virtual void updateG(ARRAY &g, ARRAY &x, ARRAY &s, ARRAY &precon, float tau, float avg_lambda);
  void fused_tv_update(cuTVPrimalDualOperator *tv, ARRAY *x, ARRAY *g, ARRAY &s, ARRAY &precon,
                       float sigma, float tau, float avg_lambda);

  int iterations_ = 10;
  float beta_ = 1.0f, gamma_ = 0.0f, kappa_ = 1.0f, tau0_ = 1e-3f, denoise_alpha_ = 0.0f;
  bool non_negativity_ = false, dump_ = true, fuse_tv_ = true;
  unsigned int reg_steps_ = 1;
  std::vector<std::shared_ptr<primalDualOperator<ARRAY>>> regularization_operators_;
  std::shared_ptr<subsetOperator<ARRAY>> encoding_operator_;
  std::shared_ptr<ARRAY> preconditioning_image_;
};

// cuNDArray<float> specialisations live in the library.
template <>
void osMOMSolverD<cuNDArray<float>>::updateG(cuNDArray<float> &g, cuNDArray<float> &x,
                                             cuNDArray<float> &s, cuNDArray<float> &precon,
                                             float tau, float avg_lambda);
template <>
void osMOMSolverD<cuNDArray<float>>::fused_tv_update(cuTVPrimalDualOperator *tv,
                                                     cuNDArray<float> *x, cuNDArray<float> *g,
                                                     cuNDArray<float> &s,
                                                     cuNDArray<float> &precon, float sigma,
                                                     float tau, float avg_lambda);

}  // namespace gtt
