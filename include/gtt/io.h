// ".real" array files: uint32 ndims, ndims x uint32 dims (first index
// fastest), then float32 data -- the format of write_nd_array and of
// python/readReal.py:41-76 in the reference.
#pragma once

#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gtt/ndarray.h"

namespace gtt {

inline void write_nd_array(const hoCuNDArray<float> *a, const std::string &path) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("write_nd_array: cannot open " + path);
  uint32_t nd = (uint32_t)a->get_number_of_dimensions();
  f.write(reinterpret_cast<const char *>(&nd), 4);
  for (size_t i = 0; i < nd; ++i) {
    uint32_t d = (uint32_t)a->get_size(i);
    f.write(reinterpret_cast<const char *>(&d), 4);
  }
  f.write(reinterpret_cast<const char *>(a->get_data_ptr()), a->get_number_of_bytes());
}

inline void write_nd_array(const cuNDArray<float> *a, const std::string &path) {
  hoCuNDArray<float> h(a->dims());
  a->to_host(&h);
  write_nd_array(&h, path);
}

// The sizes in a file's header (first index fastest); `f` is left at the data.
inline std::vector<size_t> read_nd_array_header(std::ifstream &f, const std::string &path) {
  uint32_t nd = 0;
  f.read(reinterpret_cast<char *>(&nd), 4);
  if (!f || nd == 0 || nd > 16) throw std::runtime_error("read_nd_array: bad header in " + path);
  std::vector<size_t> dims(nd);
  for (auto &d : dims) {
    uint32_t v = 0;
    f.read(reinterpret_cast<char *>(&v), 4);
    d = v;
  }
  if (!f) throw std::runtime_error("read_nd_array: bad header in " + path);
  return dims;
}

inline std::vector<size_t> read_nd_array_dims(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("read_nd_array: cannot open " + path);
  return read_nd_array_header(f, path);
}

inline hoCuNDArray<float> read_nd_array(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("read_nd_array: cannot open " + path);
  const std::vector<size_t> dims = read_nd_array_header(f, path);
  hoCuNDArray<float> a(dims);
  f.read(reinterpret_cast<char *>(a.get_data_ptr()), a.get_number_of_bytes());
  if (!f) throw std::runtime_error("read_nd_array: truncated data in " + path);
  return a;
}

}  // namespace gtt
