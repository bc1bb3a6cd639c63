// The reference's HDF5 file formats over the HDF5 C library (high-level "lite"
// API, as the reference uses): acquisitions, binnings and output images.
// See include/gtt_h5.h for the layouts and their reference sources.
#include "gtt_h5.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "gtt/CBCT_acquisition.h"

namespace gtt {
namespace {

std::mutex g_h5_mu;  // the HDF5 library is not built thread-safe

struct File {
  hid_t id = -1;
  File(const std::string &name, bool create) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // report through exceptions only
    id = create ? H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                : H5Fopen(name.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (id < 0) throw std::runtime_error("Error opening HDF5 file " + name);
  }
  ~File() {
    if (id >= 0) H5Fclose(id);
  }
  File(const File &) = delete;
  File &operator=(const File &) = delete;
};

bool exists(const std::string &f) { return std::ifstream(f).good(); }

// Dimensions of a dataset (any rank); throws `msg` if it cannot be read.
std::vector<hsize_t> dims_of(hid_t f, const char *name, const std::string &msg) {
  int rank = 0;
  if (H5LTget_dataset_ndims(f, name, &rank) < 0 || rank < 0) throw std::runtime_error(msg);
  std::vector<hsize_t> d(rank > 0 ? rank : 1, 1);
  if (H5LTget_dataset_info(f, name, d.data(), nullptr, nullptr) < 0) throw std::runtime_error(msg);
  d.resize(rank);
  return d;
}

size_t count(const std::vector<hsize_t> &d) {
  size_t n = 1;
  for (auto v : d) n *= (size_t)v;
  return n;
}

// A 3-D dataset [n][v][u] as an array (u, v, n), first index fastest.
std::shared_ptr<hoCuNDArray<float>> read_stack(hid_t f, const char *name, const std::string &info_msg,
                                               const std::string &read_msg) {
  std::vector<hsize_t> d = dims_of(f, name, info_msg);
  if (d.size() != 3) throw std::runtime_error(info_msg);
  std::vector<size_t> dims{(size_t)d[2], (size_t)d[1], (size_t)d[0]};
  auto a = std::make_shared<hoCuNDArray<float>>(dims);
  if (a->get_number_of_elements() &&
      H5LTread_dataset(f, name, H5T_NATIVE_FLOAT, a->get_data_ptr()) < 0)
    throw std::runtime_error(read_msg);
  return a;
}

void make_1d(hid_t f, const char *name, hid_t type, size_t n, const void *data) {
  hsize_t d[1] = {n};
  if (H5LTmake_dataset(f, name, 1, d, type, data) < 0)
    throw std::runtime_error(std::string("Error writing ") + name);
}

}  // namespace

// xray/CBCT_acquisition.h:177-318
void CBCT_acquisition::load(const std::string &filename) {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  if (!exists(filename)) throw std::runtime_error("Projection file does not exist");
  File file(filename, false);
  const hid_t f = file.id;
  unsigned int version = 0;
  if (H5LTread_dataset(f, "/geometry_dataformat_version", H5T_NATIVE_UINT, &version) < 0)
    throw std::runtime_error("Error reading /geometry_dataformat_version from file.");
  if (version != 2) {
    std::stringstream ss;
    ss << "wrong geometry data version inside hdf5 file, found: " << version << ", needed: 2"
       << std::endl;
    throw std::runtime_error(ss.str());
  }
  auto geometry = std::make_shared<CBCT_geometry>();
  std::vector<float> angles(count(dims_of(f, "/angles", "Error getting /angles dataset info from file.")));
  if (!angles.empty() && H5LTread_dataset(f, "/angles", H5T_NATIVE_FLOAT, angles.data()) < 0)
    throw std::runtime_error("Error reading /angles from file.");
  geometry->set_angles(angles);
  std::vector<float> ox(count(dims_of(f, "/offsetx", "Error getting /offsetx dataset info from file.")));
  if (!ox.empty() && H5LTread_dataset(f, "/offsetx", H5T_NATIVE_FLOAT, ox.data()) < 0)
    throw std::runtime_error("Error reading /offsetx from file.");
  std::vector<float> oy(count(dims_of(f, "/offsety", "Error getting /offsety dataset info from file.")));
  if (!oy.empty() && H5LTread_dataset(f, "/offsety", H5T_NATIVE_FLOAT, oy.data()) < 0)
    throw std::runtime_error("Error reading /offsety from file.");
  if (ox.size() != oy.size())
    throw std::runtime_error("CBCT_geometry::load : x/y offset arrays has different lengths");
  std::vector<floatd2> offsets;
  for (size_t i = 0; i < ox.size(); ++i) offsets.push_back(floatd2{{ox[i], oy[i]}});
  geometry->set_offsets(offsets);

  unsigned int pversion = 0;
  if (H5LTread_dataset(f, "/projection_dataformat_version", H5T_NATIVE_UINT, &pversion) < 0)
    throw std::runtime_error("Error reading /projection_dataformat_version from file.");
  if (pversion != 1) {
    std::stringstream ss;
    ss << "wrong projection data format version inside hdf5 file, found: " << pversion
       << ", needed: 1";
    throw std::runtime_error(ss.str());
  }
  projections_ = read_stack(f, "/projections", "Error getting /projections dataset info from file.",
                            "Error reading /projections from file.");
  I0_.reset();
  if (H5LTpath_valid(f, "/airscan", true) > 0)
    I0_ = read_stack(f, "/airscan", "Error getting /projections dataset info from file.",
                     "Error reading /projections from file.");

  float SAD = 0, SDD = 0;
  floatd2 FOV{{0, 0}};
  if (H5LTread_dataset(f, "/SAD", H5T_NATIVE_FLOAT, &SAD) < 0)
    throw std::runtime_error("Error reading /SAD from file.");
  if (H5LTread_dataset(f, "/SDD", H5T_NATIVE_FLOAT, &SDD) < 0)
    throw std::runtime_error("Error reading /SDD from file.");
  if (H5LTread_dataset(f, "/FOV", H5T_NATIVE_FLOAT, FOV.data()) < 0) {
    // older files: detector pixel spacing instead of the field of view
    floatd2 spacing{{0, 0}};
    if (H5LTread_dataset(f, "/spacing", H5T_NATIVE_FLOAT, spacing.data()) < 0)
      throw std::runtime_error("Error reading /FOV from file.");
    FOV[0] = spacing[0] * projections_->get_size(0);
    FOV[1] = spacing[1] * projections_->get_size(1);
  }
  geometry->set_SAD(SAD);
  geometry->set_SDD(SDD);
  geometry->set_FOV(FOV);
  geometry_ = geometry;
}

// xray/CBCT_acquisition.h:320-333 with CBCT_geometry::save (:76-111)
void CBCT_acquisition::save(const std::string &filename) {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  if (!projections_ || !geometry_) throw std::runtime_error("CBCT_acquisition::save: no data");
  File file(filename, true);
  const hid_t f = file.id;
  const unsigned int pversion = 1;
  make_1d(f, "/projection_dataformat_version", H5T_NATIVE_UINT, 1, &pversion);
  const size_t nd = projections_->get_number_of_dimensions();
  std::vector<hsize_t> d(nd);
  for (size_t i = 0; i < nd; ++i) d[i] = projections_->get_size(nd - 1 - i);
  if (H5LTmake_dataset(f, "/projections", (int)nd, d.data(), H5T_NATIVE_FLOAT,
                       projections_->get_data_ptr()) < 0)
    throw std::runtime_error("Error writing /projections");
  const unsigned int gversion = 2;
  make_1d(f, "/geometry_dataformat_version", H5T_NATIVE_UINT, 1, &gversion);
  const float SAD = geometry_->get_SAD(), SDD = geometry_->get_SDD();
  const floatd2 FOV = geometry_->get_FOV();
  make_1d(f, "/SAD", H5T_NATIVE_FLOAT, 1, &SAD);
  make_1d(f, "/SDD", H5T_NATIVE_FLOAT, 1, &SDD);
  make_1d(f, "/FOV", H5T_NATIVE_FLOAT, 2, FOV.data());
  const std::vector<float> &angles = geometry_->get_angles();
  make_1d(f, "/angles", H5T_NATIVE_FLOAT, angles.size(), angles.data());
  std::vector<float> ox, oy;
  for (auto &o : geometry_->get_offsets()) {
    ox.push_back(o[0]);
    oy.push_back(o[1]);
  }
  make_1d(f, "/offsetx", H5T_NATIVE_FLOAT, ox.size(), ox.data());
  make_1d(f, "/offsety", H5T_NATIVE_FLOAT, oy.size(), oy.data());
}

// xray/CBCT_binning.h:106-157.  (A wrong format version ends the reference's
// program with exit(); here it is an exception with the same message.)
void CBCT_binning::load(const std::string &filename) {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  if (!exists(filename)) throw std::runtime_error("Binning file not found");
  File file(filename, false);
  const hid_t f = file.id;
  unsigned int version = 0;
  if (H5LTread_dataset(f, "/binning_dataformat_version", H5T_NATIVE_UINT, &version) < 0)
    throw std::runtime_error("Error reading /binning_dataformat_version");
  if (version != 1) {
    std::stringstream ss;
    ss << "wrong format of binning file, found: " << version << ", needed: 1";
    throw std::runtime_error(ss.str());
  }
  unsigned int nbins = 0;
  if (H5LTread_dataset(f, "/numBins", H5T_NATIVE_UINT, &nbins) < 0)
    throw std::runtime_error("Error reading /numBins_dataformat_version");
  std::vector<std::vector<unsigned int>> bins;
  for (unsigned int i = 1; i <= nbins; ++i) {
    const std::string path = "/bin_" + std::to_string(i);
    std::vector<unsigned int> b(count(dims_of(f, path.c_str(), "Error reading bin info")));
    if (!b.empty() && H5LTread_dataset(f, path.c_str(), H5T_NATIVE_UINT, b.data()) < 0)
      throw std::runtime_error("Error reading bin data");
    bins.push_back(std::move(b));
  }
  binning_ = std::move(bins);
}

// python/l1shroud.py:139-150 layout
void CBCT_binning::save(const std::string &filename) const {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  File file(filename, true);
  const hid_t f = file.id;
  for (size_t i = 0; i < binning_.size(); ++i) {
    const std::string path = "/bin_" + std::to_string(i + 1);
    make_1d(f, path.c_str(), H5T_NATIVE_UINT, binning_[i].size(), binning_[i].data());
  }
  const unsigned int n = (unsigned int)binning_.size(), version = 1;
  make_1d(f, "/numBins", H5T_NATIVE_UINT, 1, &n);
  make_1d(f, "/binning_dataformat_version", H5T_NATIVE_UINT, 1, &version);
}

}  // namespace gtt

using namespace gtt;

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F &&fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

void need(bool ok) {
  if (!ok) throw std::runtime_error("illegal array pointer provided");
}
}  // namespace

extern "C" {

const char *gtt_h5_last_error(void) { return g_err.c_str(); }

int gtt_h5_acquisition_info(const char *path, int *nu, int *nv, int *nviews, int *airscan_dims) {
  return guarded([&] {
    need(path && nu && nv && nviews);
    CBCT_acquisition a;
    a.load(path);
    auto p = a.get_projections();
    *nu = (int)p->get_size(0);
    *nv = (int)p->get_size(1);
    *nviews = (int)p->get_size(2);
    if (airscan_dims) {
      auto I0 = a.get_I0();
      for (int i = 0; i < 3; ++i) airscan_dims[i] = I0 ? (int)I0->get_size(i) : 0;
    }
  });
}

int gtt_h5_acquisition_read(const char *path, float *sad, float *sdd, float *fov, float *angles,
                            float *offsets, float *projections, float *airscan) {
  return guarded([&] {
    need(path && sad && sdd && fov && angles && offsets && projections);
    CBCT_acquisition a;
    a.load(path);
    auto g = a.get_geometry();
    *sad = g->get_SAD();
    *sdd = g->get_SDD();
    fov[0] = g->get_FOV()[0];
    fov[1] = g->get_FOV()[1];
    auto &an = g->get_angles();
    std::memcpy(angles, an.data(), an.size() * sizeof(float));
    auto &of = g->get_offsets();
    for (size_t i = 0; i < of.size(); ++i) {
      offsets[2 * i] = of[i][0];
      offsets[2 * i + 1] = of[i][1];
    }
    auto p = a.get_projections();
    std::memcpy(projections, p->get_data_ptr(), p->get_number_of_bytes());
    if (airscan && a.get_I0())
      std::memcpy(airscan, a.get_I0()->get_data_ptr(), a.get_I0()->get_number_of_bytes());
  });
}

int gtt_h5_acquisition_write(const char *path, float sad, float sdd, const float *fov, int nu,
                             int nv, int nviews, const float *angles, const float *offsets,
                             const float *projections) {
  return guarded([&] {
    need(path && fov && angles && projections && nu > 0 && nv > 0 && nviews > 0);
    auto g = std::make_shared<CBCT_geometry>();
    g->set_SAD(sad);
    g->set_SDD(sdd);
    g->set_FOV(floatd2{{fov[0], fov[1]}});
    g->set_angles(std::vector<float>(angles, angles + nviews));
    std::vector<floatd2> o(nviews, floatd2{{0, 0}});
    if (offsets)
      for (int i = 0; i < nviews; ++i) o[i] = floatd2{{offsets[2 * i], offsets[2 * i + 1]}};
    g->set_offsets(o);
    auto p = std::make_shared<hoCuNDArray<float>>(std::vector<size_t>{(size_t)nu, (size_t)nv, (size_t)nviews});
    std::memcpy(p->get_data_ptr(), projections, p->get_number_of_bytes());
    CBCT_acquisition a(p, g);
    a.save(path);
  });
}

int gtt_h5_binning_info(const char *path, int *nbins, int *total) {
  return guarded([&] {
    need(path && nbins && total);
    CBCT_binning b;
    b.load(path);
    *nbins = (int)b.get_number_of_bins();
    *total = (int)b.get_number_of_projections();
  });
}

int gtt_h5_binning_read(const char *path, int *sizes, unsigned *index) {
  return guarded([&] {
    need(path && sizes && index);
    CBCT_binning b;
    b.load(path);
    size_t k = 0;
    for (unsigned int i = 0; i < b.get_number_of_bins(); ++i) {
      sizes[i] = (int)b.bin(i).size();
      for (auto v : b.bin(i)) index[k++] = v;
    }
  });
}

int gtt_h5_binning_write(const char *path, int nbins, const int *sizes, const unsigned *index) {
  return guarded([&] {
    need(path && nbins >= 0 && (nbins == 0 || (sizes && index)));
    std::vector<std::vector<unsigned int>> bins(nbins);
    size_t k = 0;
    for (int i = 0; i < nbins; ++i)
      for (int j = 0; j < sizes[i]; ++j) bins[i].push_back(index[k++]);
    CBCT_binning(bins).save(path);
  });
}

// proton/hdf5_utils.h:17-45 (3-D)
int gtt_h5_image_write(const char *path, const float *image, int nx, int ny, int nz,
                       const float *dims_mm, const float *origin, const char *arguments,
                       unsigned iterations) {
  return guarded([&] {
    need(path && image && dims_mm && origin && nx > 0 && ny > 0 && nz > 0);
    std::lock_guard<std::mutex> lk(g_h5_mu);
    File file(path, true);
    const hid_t f = file.id;
    const hsize_t d[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    if (H5LTmake_dataset(f, "/image", 3, d, H5T_NATIVE_FLOAT, image) < 0)
      throw std::runtime_error("Error writing /image");
    const int n[3] = {nx, ny, nz};
    float elem[3];
    for (int i = 0; i < 3; ++i) elem[3 - i - 1] = dims_mm[i] / n[i];
    if (H5LTset_attribute_string(f, "/image", "input", arguments ? arguments : "") < 0 ||
        H5LTset_attribute_float(f, "/image", "dimensions", dims_mm, 3) < 0 ||
        H5LTset_attribute_float(f, "/image", "element_size_um", elem, 3) < 0 ||
        H5LTset_attribute_float(f, "/image", "origin", origin, 3) < 0 ||
        H5LTset_attribute_uint(f, "/image", "iterations", &iterations, 1) < 0)
      throw std::runtime_error("Error writing /image attributes");
  });
}

}  // extern "C"
