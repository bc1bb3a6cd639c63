// Device and host N-d arrays with the Gadgetron array conventions the
// reference's operators and solvers use (first index fastest, dimension
// vectors, views onto foreign memory, deep copy on copy-construction) and the
// element-wise / BLAS free functions of cuNDArray_math / cuNDArray_blas.
// Device arithmetic runs through libcbct (include/cbct.h) on the calling
// thread's current HIP stream: the default stream (like the reference's
// legacy default stream) unless a worker thread selected its own.
#pragma once

#include <cstddef>
#include <algorithm>
#include <functional>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "cbct.h"

namespace gtt {

// Per-device context used by the array math (one per thread and device, so
// that threads running on different streams share no reduction workspace).
cbct_ctx device_context();
// HIP stream (a hipStream_t) that the calling thread's array math, copies and
// operator launches go to; nullptr is the default stream.  Copies between
// host and device synchronise this stream only.
void *current_stream();
void set_current_stream(void *stream);
// Destroy the calling thread's contexts (worker threads, before they exit).
void release_thread_resources();
// A projector context for an operator or driver: taken from the calling
// thread's pool for the current device (settings reset, view table and
// filter to be set by the caller), created when the pool is empty.  Every
// context of a thread's pool shares the scratch of device_context(), so the
// operators of one reconstruction hold one padded volume and one padded
// sinogram copy between them; contexts and scratch are kept across
// reconstructions (no device synchronisation per operator).
cbct_ctx acquire_context();
// Returns a context to the pool of the thread that took it; a context given
// back on another thread is destroyed instead.
void release_context(cbct_ctx ctx);
// Storage type (CBCT_F32 or CBCT_BF16) of the projectors' working copies for
// projection operators the calling thread creates from now on (default FP32).
void set_default_storage(int dtype);
int default_storage();
// Whether reconstructions the calling thread starts tune the forward
// projector for their geometry first (cbct_fp_autotune; default off).
void set_default_autotune(bool on);
bool default_autotune();
// Whether OS-SART reconstructions the calling thread starts replay their
// iterations as captured HIP graphs (default off).
void set_default_graphs(bool on);
bool default_graphs();
// Graphs the calling thread has launched (StreamGraph::launch) so far.
long long graph_launches();
// HIP device current on the calling thread.
int current_device();
// Throw std::runtime_error with the library's message when rc != 0.
void check(int rc, cbct_ctx ctx);

void device_malloc(void **p, size_t bytes);
void device_free(void *p);
// Frees the device blocks the calling thread keeps for reuse (device_free
// caches blocks per thread, device and stream).
void trim_device_cache();
// Bytes of device memory the calling thread keeps between uses: its cached
// blocks and the scratch of its contexts on the current device.
size_t retained_device_bytes();
// Frees both (after the thread's last calls that used them have finished).
void trim_thread_memory();
// Executable HIP graph of the device work a function queues on the calling
// thread's current stream (which must not be the default stream: it cannot
// be captured).  The function runs once, during the capture, queueing
// nothing; launch() replays its work on that stream.  Device blocks freed
// into the thread's cache while capturing stay reserved for the graph until
// it is destroyed, so no allocation outside it can share them between
// replays; no block is returned to the device while capturing.
class StreamGraph {
 public:
  explicit StreamGraph(const std::function<void()> &work);
  ~StreamGraph();
  StreamGraph(const StreamGraph &) = delete;
  StreamGraph &operator=(const StreamGraph &) = delete;
  void launch();

  struct Block {
    void *p;
    size_t bytes;
    int device;
    void *stream;
  };

 private:
  void *graph_ = nullptr, *exec_ = nullptr, *stream_ = nullptr;
  std::vector<Block> held_;
};

// When `needed`, makes a new stream the calling thread's current stream for
// the lifetime of the object (after the device has finished its queued
// work); on destruction, waits for that stream, restores the previous one
// and destroys it.
class PrivateStream {
 public:
  explicit PrivateStream(bool needed);
  ~PrivateStream();
  PrivateStream(const PrivateStream &) = delete;
  PrivateStream &operator=(const PrivateStream &) = delete;

 private:
  void *stream_ = nullptr, *previous_ = nullptr;
};

void copy_d2d(void *dst, const void *src, size_t bytes);
void copy_h2d(void *dst, const void *src, size_t bytes);
void copy_d2h(void *dst, const void *src, size_t bytes);
void device_synchronize();

inline size_t prod(const std::vector<size_t> &d) {
  return std::accumulate(d.begin(), d.end(), size_t(1), std::multiplies<size_t>());
}

template <class T>
class hoCuNDArray;

// Device array.  Owning unless constructed as a view on foreign memory.
template <class T>
class cuNDArray {
 public:
  typedef T element_type;

  cuNDArray() = default;
  explicit cuNDArray(const std::vector<size_t> &dims) { create(dims); }
  explicit cuNDArray(const std::vector<size_t> *dims) { create(*dims); }
  explicit cuNDArray(std::shared_ptr<std::vector<size_t>> dims) { create(*dims); }
  cuNDArray(const std::vector<size_t> &dims, T *data) : dims_(dims), data_(data), owns_(false) {}
  cuNDArray(const cuNDArray &o) { *this = o; }
  cuNDArray(cuNDArray &&o) noexcept { swap(o); }
  explicit cuNDArray(const hoCuNDArray<T> &h);
  ~cuNDArray() { release(); }

  cuNDArray &operator=(const cuNDArray &o) {
    if (this == &o) return *this;
    if (!data_ || prod(dims_) != o.get_number_of_elements()) create(o.dims_);
    dims_ = o.dims_;
    if (o.get_number_of_elements()) copy_d2d(data_, o.data_, o.get_number_of_bytes());
    return *this;
  }
  cuNDArray &operator=(cuNDArray &&o) noexcept {
    swap(o);
    return *this;
  }

  void create(const std::vector<size_t> &dims) {
    release();
    dims_ = dims;
    const size_t n = prod(dims_);
    if (n) device_malloc(reinterpret_cast<void **>(&data_), n * sizeof(T));
    owns_ = true;
  }
  void create(const std::vector<size_t> *dims) { create(*dims); }

  std::shared_ptr<std::vector<size_t>> get_dimensions() const {
    return std::make_shared<std::vector<size_t>>(dims_);
  }
  const std::vector<size_t> &dims() const { return dims_; }
  size_t get_size(size_t i) const { return i < dims_.size() ? dims_[i] : 1; }
  size_t get_number_of_dimensions() const { return dims_.size(); }
  size_t get_number_of_elements() const { return dims_.empty() ? 0 : prod(dims_); }
  size_t get_number_of_bytes() const { return get_number_of_elements() * sizeof(T); }
  T *get_data_ptr() const { return data_; }
  bool dimensions_equal(const cuNDArray *o) const { return o && o->dims_ == dims_; }
  void to_host(hoCuNDArray<T> *out) const;

  // Element-wise operators (float arrays)
  cuNDArray &operator*=(T a);
  cuNDArray &operator+=(T a);
  cuNDArray &operator*=(const cuNDArray &a);
  cuNDArray &operator/=(const cuNDArray &a);
  cuNDArray &operator+=(const cuNDArray &a);
  cuNDArray &operator-=(const cuNDArray &a);

  void swap(cuNDArray &o) noexcept {
    std::swap(dims_, o.dims_);
    std::swap(data_, o.data_);
    std::swap(owns_, o.owns_);
  }

 private:
  void release() {
    if (owns_ && data_) device_free(data_);
    data_ = nullptr;
    owns_ = false;
  }
  std::vector<size_t> dims_;
  T *data_ = nullptr;
  bool owns_ = false;
};

// Host array (the reference's page-locked hoCuNDArray).  Owning arrays are
// zero-filled unless made with `no_init` (for arrays about to be overwritten
// whole); the (dims, data) form wraps caller memory without owning it.
struct no_init_t {};
constexpr no_init_t no_init{};

template <class T>
class hoCuNDArray {
 public:
  typedef T element_type;
  hoCuNDArray() = default;
  explicit hoCuNDArray(const std::vector<size_t> &dims)
      : dims_(dims), own_(new T[prod(dims)]()), data_(own_.get()) {}
  hoCuNDArray(const std::vector<size_t> &dims, no_init_t)
      : dims_(dims), own_(new T[prod(dims)]), data_(own_.get()) {}
  explicit hoCuNDArray(const std::vector<size_t> *dims) : hoCuNDArray(*dims) {}
  explicit hoCuNDArray(std::shared_ptr<std::vector<size_t>> dims) : hoCuNDArray(*dims) {}
  hoCuNDArray(const std::vector<size_t> &dims, T *data) : dims_(dims), data_(data) {}
  hoCuNDArray(const hoCuNDArray &o)
      : dims_(o.dims_), own_(new T[o.get_number_of_elements()]), data_(own_.get()) {
    std::copy(o.data_, o.data_ + o.get_number_of_elements(), data_);
  }
  hoCuNDArray(hoCuNDArray &&o) noexcept
      : dims_(std::move(o.dims_)), own_(std::move(o.own_)), data_(o.data_) {
    o.data_ = nullptr;
  }
  hoCuNDArray &operator=(const hoCuNDArray &o) {
    if (this == &o) return *this;
    std::unique_ptr<T[]> n(new T[o.get_number_of_elements()]);
    std::copy(o.data_, o.data_ + o.get_number_of_elements(), n.get());
    dims_ = o.dims_;
    own_ = std::move(n);
    data_ = own_.get();
    return *this;
  }
  hoCuNDArray &operator=(hoCuNDArray &&o) noexcept {
    dims_ = std::move(o.dims_);
    own_ = std::move(o.own_);
    data_ = o.data_;
    o.data_ = nullptr;
    return *this;
  }
  std::shared_ptr<std::vector<size_t>> get_dimensions() const {
    return std::make_shared<std::vector<size_t>>(dims_);
  }
  const std::vector<size_t> &dims() const { return dims_; }
  size_t get_size(size_t i) const { return i < dims_.size() ? dims_[i] : 1; }
  size_t get_number_of_dimensions() const { return dims_.size(); }
  size_t get_number_of_elements() const { return dims_.empty() ? 0 : prod(dims_); }
  size_t get_number_of_bytes() const { return get_number_of_elements() * sizeof(T); }
  T *get_data_ptr() const { return data_; }

 private:
  std::vector<size_t> dims_;
  std::unique_ptr<T[]> own_;
  T *data_ = nullptr;
};

template <class T>
cuNDArray<T>::cuNDArray(const hoCuNDArray<T> &h) {
  create(h.dims());
  if (get_number_of_elements()) copy_h2d(data_, h.get_data_ptr(), get_number_of_bytes());
}

template <class T>
void cuNDArray<T>::to_host(hoCuNDArray<T> *out) const {
  if (out->get_number_of_elements() != get_number_of_elements()) *out = hoCuNDArray<T>(dims_);
  if (get_number_of_elements()) copy_d2h(out->get_data_ptr(), data_, get_number_of_bytes());
}

// cuNDArray_math / cuNDArray_blas equivalents (float).  dot/nrm2/asum
// accumulate in FP64 on the device and return double.
void clear(cuNDArray<float> *x);
void fill(cuNDArray<float> *x, float v);
void clamp_min(cuNDArray<float> *x, float v);
void reciprocal_inplace(cuNDArray<float> *x);
void axpy(float a, const cuNDArray<float> *x, cuNDArray<float> *y);
double dot(const cuNDArray<float> *a, const cuNDArray<float> *b);
double nrm2(const cuNDArray<float> *a);
double asum(const cuNDArray<float> *a);

// Fused solver passes (libcbct): see cbct_sart_residual etc.
void sart_residual(cuNDArray<float> *r, const cuNDArray<float> *b, const cuNDArray<float> *w);
void sart_update(cuNDArray<float> *x, const cuNDArray<float> *img, const cuNDArray<float> *wx,
                 float step, bool nonneg);
void mom_gradient_step(cuNDArray<float> *z, const cuNDArray<float> *img,
                       const cuNDArray<float> *precon, float step);
void mom_momentum(cuNDArray<float> *x, const cuNDArray<float> *xold, cuNDArray<float> *z, float c,
                  bool nonneg);

// Host helpers
void clear(hoCuNDArray<float> *x);
double dot(const hoCuNDArray<float> *a, const hoCuNDArray<float> *b);

}  // namespace gtt
