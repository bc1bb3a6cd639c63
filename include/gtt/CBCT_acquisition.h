// Cone-beam scan geometry, acquisition and binning: the data model of
// xray/CBCT_acquisition.h:24-340 and xray/CBCT_binning.h:18-172.
#pragma once

#include <algorithm>
#include <array>
#include <memory>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "gtt/ndarray.h"

namespace gtt {

typedef std::array<float, 2> floatd2;
typedef std::array<float, 3> floatd3;

class CBCT_geometry {
 public:
  void set_SAD(float v) { SAD_ = v; }
  float get_SAD() const { return SAD_; }
  void set_SDD(float v) { SDD_ = v; }
  float get_SDD() const { return SDD_; }
  void set_FOV(floatd2 v) { FOV_ = v; }
  floatd2 get_FOV() const { return FOV_; }
  void set_angles(std::vector<float> a) { angles_ = std::move(a); }
  std::vector<float> &get_angles() { return angles_; }
  void set_offsets(std::vector<floatd2> o) { offsets_ = std::move(o); }
  std::vector<floatd2> &get_offsets() { return offsets_; }

  void print(std::ostream &os) const {
    os << "------------ GEOMETRY ------------\n";
    if (angles_.empty()) os << "Angles: EMPTY\n";
    else os << "Angles: " << angles_.front() << " ... " << angles_.back() << ", number of angles: " << angles_.size() << "\n";
    os << "Offsets: contains " << offsets_.size() << " elements\n";
    os << "SDD: " << SDD_ << "mm\nSAD: " << SAD_ << "mm\nFOV: " << FOV_[0] << " " << FOV_[1] << "mm\n";
    os << "----------------------------------\n";
  }

 private:
  float SAD_ = 0.0f, SDD_ = 0.0f;
  floatd2 FOV_{{0.0f, 0.0f}};
  std::vector<float> angles_;
  std::vector<floatd2> offsets_;
};

class CBCT_acquisition {
 public:
  CBCT_acquisition() = default;
  CBCT_acquisition(std::shared_ptr<hoCuNDArray<float>> projections,
                   std::shared_ptr<CBCT_geometry> geometry)
      : geometry_(std::move(geometry)), projections_(std::move(projections)) {}
  void set_geometry(std::shared_ptr<CBCT_geometry> g) { geometry_ = std::move(g); }
  std::shared_ptr<CBCT_geometry> get_geometry() { return geometry_; }
  void set_projections(std::shared_ptr<hoCuNDArray<float>> p) { projections_ = std::move(p); }
  std::shared_ptr<hoCuNDArray<float>> get_projections() { return projections_; }
  // Air scan (I0) if the file holds one (nullptr otherwise).
  std::shared_ptr<hoCuNDArray<float>> get_I0() { return I0_; }
  void set_I0(std::shared_ptr<hoCuNDArray<float>> I0) { I0_ = std::move(I0); }

  // HDF5 acquisition files (geometry format 2, projection format 1):
  // xray/CBCT_acquisition.h:177-333.  Defined in libgtt_h5.so.
  void load(const std::string &filename);
  void save(const std::string &filename);

 private:
  std::shared_ptr<CBCT_geometry> geometry_;
  std::shared_ptr<hoCuNDArray<float>> projections_;
  std::shared_ptr<hoCuNDArray<float>> I0_;
};

class CBCT_binning {
 public:
  CBCT_binning() = default;
  explicit CBCT_binning(std::vector<std::vector<unsigned int>> b) : binning_(std::move(b)) {}
  unsigned int get_number_of_bins() const { return (unsigned int)binning_.size(); }
  unsigned int get_number_of_projections() const {
    unsigned int acc = 0;
    for (auto &b : binning_) acc += (unsigned int)b.size();
    return acc;
  }
  unsigned int get_number_of_projections(unsigned int bin) const {
    if (bin >= binning_.size())
      throw std::runtime_error("CBCT_binning::get_number_of_projections(int) : bin is out of range");
    return (unsigned int)binning_[bin].size();
  }
  int get_maximum_projection_index() const {
    int m = -1;
    for (auto &b : binning_)
      for (auto v : b) m = std::max(m, (int)v);
    return m;
  }
  void set_bins(std::vector<std::vector<unsigned int>> &b) { binning_ = b; }
  std::vector<std::vector<unsigned int>> get_bins() const { return binning_; }
  const std::vector<unsigned int> &bin(unsigned int b) const { return binning_[b]; }
  std::vector<unsigned int> get_bin(unsigned int b) const {
    if (b >= binning_.size()) throw std::runtime_error("CBCT_binning::get_bin() : bin is out of range");
    return binning_[b];
  }
  void set_bin(std::vector<unsigned int> &bin, unsigned int n) {
    if (n > binning_.size()) throw std::runtime_error("CBCT_binning::set_bin() : bin is out of range");
    if (n == binning_.size()) binning_.push_back(bin);
    else binning_[n] = bin;
  }
  void set_as_default_3d_bin(unsigned int nproj) {
    std::vector<unsigned int> b(nproj);
    for (unsigned int i = 0; i < nproj; ++i) b[i] = i;
    binning_.push_back(b);
  }
  // HDF5 binning files (format 1: /binning_dataformat_version, /numBins,
  // /bin_1.. /bin_N): xray/CBCT_binning.h:106-157; save writes the layout of
  // python/l1shroud.py:139-150.  Defined in libgtt_h5.so.
  void load(const std::string &filename);
  void save(const std::string &filename) const;
  std::shared_ptr<CBCT_binning> get_3d_binning() const {
    std::set<unsigned int> u;
    for (auto &b : binning_) u.insert(b.begin(), b.end());
    return std::make_shared<CBCT_binning>(std::vector<std::vector<unsigned int>>(1, std::vector<unsigned int>(u.begin(), u.end())));
  }

 private:
  std::vector<std::vector<unsigned int>> binning_;
};

}  // namespace gtt
