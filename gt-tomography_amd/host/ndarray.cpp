// Device memory helpers and the cuNDArray math used by the operators and
// solvers, on top of libcbct.
#include "gtt/ndarray.h"

#include <hip/hip_runtime_api.h>

#include <map>
#include <set>
#include <vector>
#include <string>

namespace gtt {

namespace {
// Contexts of the calling thread, by device.  Not destroyed at thread exit
// (the HIP runtime may already be gone at process exit); worker threads call
// release_thread_resources().
thread_local std::map<int, cbct_ctx> t_ctx;
thread_local hipStream_t t_stream = nullptr;
thread_local int t_storage = CBCT_F32;
thread_local bool t_autotune = false;
thread_local bool t_graphs = false;
thread_local long long t_graph_launches = 0;

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

cbct_ctx device_context() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  auto it = t_ctx.find(dev);
  if (it != t_ctx.end()) return it->second;
  cbct_ctx c = nullptr;
  int rc = cbct_ctx_create(dev, &c);
  if (rc) throw std::runtime_error(cbct_last_error(nullptr));
  t_ctx[dev] = c;
  return c;
}

void *current_stream() { return t_stream; }
void set_default_storage(int dtype) { t_storage = dtype == CBCT_BF16 ? CBCT_BF16 : CBCT_F32; }
int default_storage() { return t_storage; }
void set_default_autotune(bool on) { t_autotune = on; }
bool default_autotune() { return t_autotune; }
void set_default_graphs(bool on) { t_graphs = on; }
bool default_graphs() { return t_graphs; }
long long graph_launches() { return t_graph_launches; }
void set_current_stream(void *stream) { t_stream = static_cast<hipStream_t>(stream); }
namespace {
// Pooled projector contexts of the calling thread, by device; and every
// context the thread handed out (to recognise returns from other threads).
thread_local std::map<int, std::vector<cbct_ctx>> t_pool;
thread_local std::map<cbct_ctx, int> t_pool_owned;   // context -> its device
}  // namespace

cbct_ctx acquire_context() {
  const int dev = current_device();
  auto &free_list = t_pool[dev];
  if (!free_list.empty()) {
    cbct_ctx c = free_list.back();
    free_list.pop_back();
    check(cbct_ctx_reset(c), c);
    return c;
  }
  cbct_ctx owner = device_context();
  cbct_ctx c = nullptr;
  check(cbct_ctx_create(dev, &c), nullptr);
  int rc = cbct_ctx_share_workspace(c, owner);
  if (rc) {
    std::string msg = cbct_last_error(c);
    cbct_ctx_destroy(c);
    throw std::runtime_error(msg);
  }
  t_pool_owned[c] = dev;
  return c;
}

void release_context(cbct_ctx c) {
  if (!c) return;
  auto it = t_pool_owned.find(c);
  if (it == t_pool_owned.end()) {   // taken on another thread
    cbct_ctx_destroy(c);
    return;
  }
  t_pool[it->second].push_back(c);
}

void release_thread_resources() {
  trim_device_cache();
  for (auto &kv : t_pool)
    for (cbct_ctx c : kv.second) cbct_ctx_destroy(c);
  t_pool.clear();
  t_pool_owned.clear();
  for (auto &kv : t_ctx) cbct_ctx_destroy(kv.second);
  t_ctx.clear();
}

void check(int rc, cbct_ctx ctx) {
  if (rc != CBCT_OK) throw std::runtime_error(cbct_last_error(ctx));
}

// Freed device blocks are kept for reuse by the same thread on the same
// device and stream: the operators and solvers create and drop arrays of the
// same few shapes every subset, and hipFree synchronises the device and costs
// milliseconds on large blocks.  Work queued on the freeing stream is ordered
// before any later use of the block on that stream, so reuse needs no wait.
namespace {
struct CachedBlock {
  void *p;
  size_t bytes;
  int device;
  hipStream_t stream;
};
thread_local std::vector<CachedBlock> t_free_blocks;
thread_local std::map<void *, size_t> t_block_sizes;   // live blocks from device_malloc
thread_local size_t t_cached_bytes = 0;
constexpr size_t kCacheLimit = size_t(32) << 30;
// While a StreamGraph captures: the blocks freed into the cache meanwhile
thread_local std::vector<void *> *t_capture_freed = nullptr;
}  // namespace

void trim_device_cache() {
  for (auto &b : t_free_blocks) (void)hipFree(b.p);
  t_free_blocks.clear();
  t_cached_bytes = 0;
}

size_t retained_device_bytes() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return t_cached_bytes;
  auto it = t_ctx.find(dev);
  return t_cached_bytes + (it == t_ctx.end() ? 0 : cbct_workspace_bytes(it->second));
}

void trim_thread_memory() {
  trim_device_cache();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  auto it = t_ctx.find(dev);
  if (it != t_ctx.end()) (void)cbct_ctx_trim(it->second);
}

void device_malloc(void **p, size_t bytes) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  for (size_t i = 0; i < t_free_blocks.size(); ++i) {
    const CachedBlock &b = t_free_blocks[i];
    if (b.bytes == bytes && b.device == dev && b.stream == t_stream) {
      *p = b.p;
      t_cached_bytes -= bytes;
      t_free_blocks.erase(t_free_blocks.begin() + i);
      t_block_sizes[*p] = bytes;
      return;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && retained_device_bytes() > 0 && !t_capture_freed) {
    (void)hipGetLastError();
    trim_thread_memory();
    e = hipMalloc(p, bytes);
  }
  hip_check(e, "hipMalloc");
  t_block_sizes[*p] = bytes;
}

void device_free(void *p) {
  auto it = t_block_sizes.find(p);
  if (it == t_block_sizes.end()) {   // not this thread's block
    (void)hipFree(p);
    return;
  }
  const size_t bytes = it->second;
  t_block_sizes.erase(it);
  int dev = 0;
  const bool dev_ok = hipGetDevice(&dev) == hipSuccess;
  if (t_capture_freed) {   // no hipFree while capturing: it synchronises
    t_capture_freed->push_back(p);
  } else if (t_cached_bytes + bytes > kCacheLimit || !dev_ok) {
    (void)hipFree(p);
    return;
  }
  t_free_blocks.push_back({p, bytes, dev, t_stream});
  t_cached_bytes += bytes;
}

StreamGraph::StreamGraph(const std::function<void()> &work) {
  hipStream_t s = t_stream;
  if (!s) throw std::runtime_error("StreamGraph: the default stream cannot be captured");
  std::vector<void *> freed;
  t_capture_freed = &freed;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    t_capture_freed = nullptr;
    hip_check(e, "hipStreamBeginCapture");
  }
  hipGraph_t g = nullptr;
  try {
    work();
  } catch (...) {
    t_capture_freed = nullptr;
    if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    throw;
  }
  t_capture_freed = nullptr;
  hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    hip_check(e, "hipGraphInstantiate");
  }
  graph_ = g;
  exec_ = x;
  stream_ = s;
  // reserve the blocks the captured work uses as temporaries
  const std::set<void *> mine(freed.begin(), freed.end());
  for (auto it = t_free_blocks.begin(); it != t_free_blocks.end();) {
    if (mine.count(it->p)) {
      held_.push_back({it->p, it->bytes, it->device, it->stream});
      t_cached_bytes -= it->bytes;
      it = t_free_blocks.erase(it);
    } else {
      ++it;
    }
  }
}

StreamGraph::~StreamGraph() {
  // the last replay may still run: finish it before the graph and its
  // blocks are let go
  if (stream_) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
  if (exec_) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec_));
  if (graph_) (void)hipGraphDestroy(static_cast<hipGraph_t>(graph_));
  // back into the cache: reused only on the same stream, after the replays
  for (const Block &b : held_) {
    t_free_blocks.push_back({b.p, b.bytes, b.device, static_cast<hipStream_t>(b.stream)});
    t_cached_bytes += b.bytes;
  }
}

void StreamGraph::launch() {
  hip_check(hipGraphLaunch(static_cast<hipGraphExec_t>(exec_), static_cast<hipStream_t>(stream_)),
            "hipGraphLaunch");
  ++t_graph_launches;
}

PrivateStream::PrivateStream(bool needed) {
  if (!needed) return;
  device_synchronize();
  hipStream_t s = nullptr;
  hip_check(hipStreamCreate(&s), "hipStreamCreate");
  stream_ = s;
  previous_ = t_stream;
  t_stream = s;
}

PrivateStream::~PrivateStream() {
  if (!stream_) return;
  (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
  t_stream = static_cast<hipStream_t>(previous_);
  (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
}
void copy_d2d(void *dst, const void *src, size_t bytes) {
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, t_stream), "hipMemcpyAsync");
}
// Host copies are ordered after the thread's queued work and complete before
// returning (the host buffer may be reused or read right away).
void copy_h2d(void *dst, const void *src, size_t bytes) {
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, t_stream), "hipMemcpyAsync H2D");
  hip_check(hipStreamSynchronize(t_stream), "hipStreamSynchronize");
}
void copy_d2h(void *dst, const void *src, size_t bytes) {
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, t_stream), "hipMemcpyAsync D2H");
  hip_check(hipStreamSynchronize(t_stream), "hipStreamSynchronize");
}
void device_synchronize() {
  if (t_stream)
    hip_check(hipStreamSynchronize(t_stream), "hipStreamSynchronize");
  else
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

#define CTX device_context()

void clear(cuNDArray<float> *x) { fill(x, 0.0f); }
void fill(cuNDArray<float> *x, float v) {
  cbct_ctx c = CTX;
  check(cbct_fill(c, x->get_number_of_elements(), x->get_data_ptr(), v, current_stream()), c);
}
void clamp_min(cuNDArray<float> *x, float v) {
  cbct_ctx c = CTX;
  check(cbct_clamp_min(c, x->get_number_of_elements(), x->get_data_ptr(), v, current_stream()), c);
}
void reciprocal_inplace(cuNDArray<float> *x) {
  cbct_ctx c = CTX;
  check(cbct_reciprocal(c, x->get_number_of_elements(), x->get_data_ptr(), current_stream()), c);
}
static void same_size(const cuNDArray<float> *a, const cuNDArray<float> *b, const char *who) {
  if (a->get_number_of_elements() != b->get_number_of_elements())
    throw std::runtime_error(std::string(who) + ": array sizes differ");
}
void axpy(float a, const cuNDArray<float> *x, cuNDArray<float> *y) {
  same_size(x, y, "axpy");
  cbct_ctx c = CTX;
  check(cbct_axpy(c, y->get_number_of_elements(), a, x->get_data_ptr(), y->get_data_ptr(), current_stream()), c);
}
double dot(const cuNDArray<float> *a, const cuNDArray<float> *b) {
  same_size(a, b, "dot");
  cbct_ctx c = CTX;
  double r = 0;
  check(cbct_reduce_f64(c, CBCT_DOT, a->get_number_of_elements(), a->get_data_ptr(), b->get_data_ptr(), &r, current_stream()), c);
  return r;
}
double nrm2(const cuNDArray<float> *a) {
  cbct_ctx c = CTX;
  double r = 0;
  check(cbct_reduce_f64(c, CBCT_NRM2, a->get_number_of_elements(), a->get_data_ptr(), nullptr, &r, current_stream()), c);
  return r;
}
double asum(const cuNDArray<float> *a) {
  cbct_ctx c = CTX;
  double r = 0;
  check(cbct_reduce_f64(c, CBCT_ASUM, a->get_number_of_elements(), a->get_data_ptr(), nullptr, &r, current_stream()), c);
  return r;
}

void sart_residual(cuNDArray<float> *r, const cuNDArray<float> *b, const cuNDArray<float> *w) {
  same_size(r, b, "sart_residual");
  same_size(r, w, "sart_residual");
  cbct_ctx c = CTX;
  check(cbct_sart_residual(c, r->get_number_of_elements(), r->get_data_ptr(), b->get_data_ptr(),
                           w->get_data_ptr(), current_stream()), c);
}
void sart_update(cuNDArray<float> *x, const cuNDArray<float> *img, const cuNDArray<float> *wx,
                 float step, bool nonneg) {
  same_size(x, img, "sart_update");
  same_size(x, wx, "sart_update");
  cbct_ctx c = CTX;
  check(cbct_sart_update(c, x->get_number_of_elements(), x->get_data_ptr(), img->get_data_ptr(),
                         wx->get_data_ptr(), step, nonneg ? 1 : 0, current_stream()), c);
}
void mom_gradient_step(cuNDArray<float> *z, const cuNDArray<float> *img,
                       const cuNDArray<float> *precon, float step) {
  same_size(z, img, "mom_gradient_step");
  same_size(z, precon, "mom_gradient_step");
  cbct_ctx c = CTX;
  check(cbct_mom_gradient_step(c, z->get_number_of_elements(), z->get_data_ptr(), img->get_data_ptr(),
                               precon->get_data_ptr(), step, current_stream()), c);
}
void mom_momentum(cuNDArray<float> *x, const cuNDArray<float> *xold, cuNDArray<float> *z, float cc,
                  bool nonneg) {
  same_size(x, xold, "mom_momentum");
  same_size(x, z, "mom_momentum");
  cbct_ctx c = CTX;
  check(cbct_mom_momentum(c, x->get_number_of_elements(), x->get_data_ptr(), xold->get_data_ptr(),
                          z->get_data_ptr(), cc, nonneg ? 1 : 0, current_stream()), c);
}

template <>
cuNDArray<float> &cuNDArray<float>::operator*=(float a) {
  cbct_ctx c = CTX;
  check(cbct_scale(c, get_number_of_elements(), data_, a, current_stream()), c);
  return *this;
}
template <>
cuNDArray<float> &cuNDArray<float>::operator+=(float a) {
  cbct_ctx c = CTX;
  check(cbct_add_scalar(c, get_number_of_elements(), data_, a, current_stream()), c);
  return *this;
}
template <>
cuNDArray<float> &cuNDArray<float>::operator*=(const cuNDArray<float> &a) {
  same_size(this, &a, "operator*=");
  cbct_ctx c = CTX;
  check(cbct_mul(c, get_number_of_elements(), a.data_, data_, current_stream()), c);
  return *this;
}
template <>
cuNDArray<float> &cuNDArray<float>::operator/=(const cuNDArray<float> &a) {
  same_size(this, &a, "operator/=");
  cbct_ctx c = CTX;
  check(cbct_div(c, get_number_of_elements(), a.data_, data_, current_stream()), c);
  return *this;
}
template <>
cuNDArray<float> &cuNDArray<float>::operator+=(const cuNDArray<float> &a) {
  same_size(this, &a, "operator+=");
  cbct_ctx c = CTX;
  check(cbct_add(c, get_number_of_elements(), a.data_, data_, current_stream()), c);
  return *this;
}
template <>
cuNDArray<float> &cuNDArray<float>::operator-=(const cuNDArray<float> &a) {
  same_size(this, &a, "operator-=");
  cbct_ctx c = CTX;
  check(cbct_sub(c, get_number_of_elements(), a.data_, data_, current_stream()), c);
  return *this;
}

void clear(hoCuNDArray<float> *x) {
  float *p = x->get_data_ptr();
  for (size_t i = 0, n = x->get_number_of_elements(); i < n; ++i) p[i] = 0.0f;
}

double dot(const hoCuNDArray<float> *a, const hoCuNDArray<float> *b) {
  if (a->get_number_of_elements() != b->get_number_of_elements())
    throw std::runtime_error("dot: array sizes differ");
  double acc = 0.0;
  const float *pa = a->get_data_ptr(), *pb = b->get_data_ptr();
  for (size_t i = 0, n = a->get_number_of_elements(); i < n; ++i) acc += (double)pa[i] * pb[i];
  return acc;
}

}  // namespace gtt
