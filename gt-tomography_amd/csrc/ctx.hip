// Context, view table, workspaces, element-wise solver passes and FP64
// reductions.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "cbct_internal.h"

namespace {
std::mutex g_create_err_mu;
std::string g_create_err;  // message of the last failed cbct_ctx_create
}  // namespace

int cbct_fail(cbct_ctx ctx, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) {
    ctx->err = buf;
  } else {
    std::lock_guard<std::mutex> lk(g_create_err_mu);
    g_create_err = buf;
  }
  return code;
}

void cbct_prof_mark(cbct_ctx ctx, int which, int stop, hipStream_t s) {
  if (!ctx->profile) return;
  (void)hipEventRecord(ctx->ev[which][stop ? 1 : 0], s);
  if (stop) ctx->ev_recorded[which] = true;
}

int cbct_set_device(cbct_ctx ctx) {
  CBCT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return CBCT_OK;
}

int cbct_reserve(cbct_ctx ctx, DeviceBuffer &b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return CBCT_OK;
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return cbct_fail(ctx, CBCT_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return cbct_fail(ctx, CBCT_ERR_NOMEM, "workspace allocation of %zu bytes failed: %s", want,
                     hipGetErrorString(e));
  }
  b.bytes = want;
  return CBCT_OK;
}

int cbct_check_geom(cbct_ctx ctx, const cbct_geom *g, const char *who) {
  if (!g) return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal geometry pointer provided", who);
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->nu <= 0 || g->nv <= 0)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: array dimensions must be positive", who);
  if (g->nx > 4096 || g->ny > 4096 || g->nz > 4096 || g->nu > 8192 || g->nv > 8192)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: array dimensions exceed the supported maximum",
                     who);
  if (!(g->lx > 0 && g->ly > 0 && g->lz > 0 && g->fov_u > 0 && g->fov_v > 0 && g->sad > 0 &&
        g->sdd > 0))
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: geometry extents and distances must be positive",
                     who);
  return CBCT_OK;
}

void Workspace::free_buffers() {
  DeviceBuffer *bufs[] = {&vol_pad, &proj_pad, &red, &filt_tmp, &ssim_ws, &bp_rects};
  for (auto *b : bufs) {
    if (b->ptr) (void)hipFree(b->ptr);
    b->ptr = nullptr;
    b->bytes = 0;
  }
}

Workspace::~Workspace() {
  (void)hipSetDevice(device);
  if (done) {
    if (recorded) (void)hipEventSynchronize(done);
    (void)hipEventDestroy(done);
  }
  free_buffers();
}

CallScope::CallScope(cbct_ctx c, hipStream_t st) : ctx(c), s(st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    cs = hipStreamCaptureStatusNone;
  }
  capturing = cs != hipStreamCaptureStatusNone;
  Workspace &w = *ctx->ws;
  if (!capturing && w.recorded && w.last != s) (void)hipStreamWaitEvent(s, w.done, 0);
}

CallScope::~CallScope() {
  if (capturing) return;
  Workspace &w = *ctx->ws;
  if (!w.done && hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess) w.done = nullptr;
  if (!ctx->done && hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming) != hipSuccess)
    ctx->done = nullptr;
  if (w.done && hipEventRecord(w.done, s) == hipSuccess) {
    w.recorded = true;
    w.last = s;
  }
  if (ctx->done && hipEventRecord(ctx->done, s) == hipSuccess) ctx->done_recorded = true;
}

namespace {
std::mutex g_upload_mu;
hipStream_t g_upload_stream[64] = {};
}  // namespace

int cbct_upload(cbct_ctx ctx, void *dst, const void *src, size_t bytes) {
  if (ctx->device < 0 || ctx->device >= 64)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "device %d out of range", ctx->device);
  std::lock_guard<std::mutex> lk(g_upload_mu);
  hipStream_t &s = g_upload_stream[ctx->device];
  if (!s) CBCT_HIP_CHECK(ctx, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  CBCT_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  CBCT_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return CBCT_OK;
}

int cbct_wait_tables(cbct_ctx ctx) {
  if (ctx->done_recorded) CBCT_HIP_CHECK(ctx, hipEventSynchronize(ctx->done));
  return CBCT_OK;
}

int cbct_check_views(cbct_ctx ctx, int view0, int nviews, const char *who) {
  if (ctx->nviews == 0)
    return cbct_fail(ctx, CBCT_ERR_STATE, "Error: %s: setup not performed (no views set)", who);
  if (view0 < 0 || nviews < 0 || view0 + nviews > ctx->nviews)
    return cbct_fail(ctx, CBCT_ERR_ARG,
                     "Error: %s: inconsistent sizes of input arrays/vectors (views [%d,%d) of %d)",
                     who, view0, view0 + nviews, ctx->nviews);
  return CBCT_OK;
}

extern "C" {

int cbct_ctx_create(int device, cbct_ctx *out) {
  if (!out) return cbct_fail(nullptr, CBCT_ERR_ARG, "cbct_ctx_create: null output pointer");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return cbct_fail(nullptr, CBCT_ERR_HIP, "cbct_ctx_create: no HIP device available (%s)",
                     hipGetErrorString(e));
  }
  if (device < 0 || device >= count)
    return cbct_fail(nullptr, CBCT_ERR_ARG, "cbct_ctx_create: device %d out of range [0,%d)",
                     device, count);
  cbct_ctx ctx = new cbct_ctx_s();
  ctx->device = device;
  ctx->ws = std::make_shared<Workspace>();
  ctx->ws->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&ctx->h_red), 64);
  if (e != hipSuccess) {
    delete ctx;
    return cbct_fail(nullptr, CBCT_ERR_HIP, "cbct_ctx_create: %s", hipGetErrorString(e));
  }
  *out = ctx;
  return CBCT_OK;
}

int cbct_ctx_trim(cbct_ctx ctx) {
  if (!ctx) return CBCT_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  Workspace &w = *ctx->ws;
  if (w.recorded) (void)hipEventSynchronize(w.done);
  w.free_buffers();
  return CBCT_OK;
}

int cbct_ctx_share_workspace(cbct_ctx ctx, cbct_ctx owner) {
  if (!ctx || !owner) return CBCT_ERR_ARG;
  if (ctx->device != owner->device)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_ctx_share_workspace: contexts on devices %d and %d",
                     ctx->device, owner->device);
  if (ctx->ws == owner->ws) return CBCT_OK;
  (void)hipSetDevice(ctx->device);
  ctx->ws = owner->ws;   // the previous workspace goes when its last holder does
  return CBCT_OK;
}

int cbct_ctx_reset(cbct_ctx ctx) {
  if (!ctx) return CBCT_ERR_ARG;
  const cbct_ctx_s d;
  ctx->filter_engine = d.filter_engine;
  ctx->fp_engine = d.fp_engine;
  ctx->bp_engine = d.bp_engine;
  ctx->bp_kz = d.bp_kz;
  ctx->fp_max_rpl = d.fp_max_rpl;
  ctx->fp_view_group = d.fp_view_group;
  ctx->fp_force_variant = d.fp_force_variant;
  ctx->fp_force_rpl = d.fp_force_rpl;
  ctx->fp_force_vg = d.fp_force_vg;
  ctx->work_dtype = d.work_dtype;
  ctx->profile = false;
  for (bool &r : ctx->ev_recorded) r = false;
  ctx->err.clear();
  return CBCT_OK;
}

int cbct_profile_enable(cbct_ctx ctx, int on) {
  if (!ctx) return CBCT_ERR_ARG;
  int rc = cbct_set_device(ctx);
  if (rc) return rc;
  if (on && !ctx->ev[0][0]) {
    for (auto &pair : ctx->ev)
      for (auto &e : pair) CBCT_HIP_CHECK(ctx, hipEventCreate(&e));
  }
  ctx->profile = on != 0;
  for (bool &r : ctx->ev_recorded) r = false;
  return CBCT_OK;
}

int cbct_profile_read(cbct_ctx ctx, double *fp_ms, double *bp_ms, double *filter_ms) {
  if (!ctx) return CBCT_ERR_ARG;
  double *outs[3] = {fp_ms, bp_ms, filter_ms};
  for (int i = 0; i < 3; ++i) {
    if (!outs[i]) continue;
    *outs[i] = -1.0;
    if (!ctx->ev_recorded[i]) continue;
    CBCT_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[i][1]));
    float ms = 0.0f;
    CBCT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[i][0], ctx->ev[i][1]));
    *outs[i] = ms;
  }
  return CBCT_OK;
}

int cbct_ctx_destroy(cbct_ctx ctx) {
  if (!ctx) return CBCT_OK;
  (void)hipSetDevice(ctx->device);
  (void)cbct_wait_tables(ctx);
  if (ctx->done) (void)hipEventDestroy(ctx->done);
  for (auto &pair : ctx->ev)
    for (auto &e : pair)
      if (e) (void)hipEventDestroy(e);
  void *ptrs[] = {ctx->d_views_d, ctx->d_views_f, ctx->d_angles, ctx->d_h,
                  ctx->d_thi,     ctx->d_tlo,     ctx->d_fft,    ctx->d_cosw};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (ctx->h_red) (void)hipHostFree(ctx->h_red);
  delete ctx;
  return CBCT_OK;
}

const char *cbct_last_error(cbct_ctx ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_create_err_mu);
  static thread_local std::string copy;
  copy = g_create_err;
  return copy.c_str();
}

size_t cbct_workspace_bytes(cbct_ctx ctx) {
  if (!ctx) return 0;
  const Workspace &w = *ctx->ws;
  return w.vol_pad.bytes + w.proj_pad.bytes + w.red.bytes + w.filt_tmp.bytes + w.ssim_ws.bytes +
         w.bp_rects.bytes;
}

int cbct_views_set(cbct_ctx ctx, const float *angles_deg, const float *offsets, int n) {
  if (!ctx) return CBCT_ERR_ARG;
  if (!angles_deg || n <= 0)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: cbct_views_set: illegal angle array provided");
  int rc = cbct_set_device(ctx);
  if (rc) return rc;
  std::vector<ViewD> vd(n);
  std::vector<ViewF> vf(n);
  ctx->angles_deg.assign(angles_deg, angles_deg + n);
  ctx->offsets.assign(2 * static_cast<size_t>(n), 0.0f);
  if (offsets) std::memcpy(ctx->offsets.data(), offsets, sizeof(float) * 2 * n);
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < n; ++k) {
    double th = static_cast<double>(angles_deg[k]) * (pi / 180.0);
    vd[k] = {std::cos(th), std::sin(th), ctx->offsets[2 * k], ctx->offsets[2 * k + 1]};
    vf[k] = {static_cast<float>(vd[k].c), static_cast<float>(vd[k].s), ctx->offsets[2 * k],
             ctx->offsets[2 * k + 1]};
  }
  if (n > ctx->nviews || !ctx->d_views_d) {
    if (ctx->d_views_d) (void)hipFree(ctx->d_views_d);
    if (ctx->d_views_f) (void)hipFree(ctx->d_views_f);
    if (ctx->d_angles) (void)hipFree(ctx->d_angles);
    ctx->d_views_d = nullptr;
    ctx->d_views_f = nullptr;
    ctx->d_angles = nullptr;
    ctx->nviews = 0;
    CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_views_d, sizeof(ViewD) * n));
    CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_views_f, sizeof(ViewF) * n));
    CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_angles, sizeof(float) * n));
  }
  // Synchronous upload: callers may set views between asynchronous launches;
  // work already enqueued that reads the old table must be complete first
  // (this context's calls only, on whatever stream they were made).
  rc = cbct_wait_tables(ctx);
  if (rc) return rc;
  if ((rc = cbct_upload(ctx, ctx->d_views_d, vd.data(), sizeof(ViewD) * n))) return rc;
  if ((rc = cbct_upload(ctx, ctx->d_views_f, vf.data(), sizeof(ViewF) * n))) return rc;
  if ((rc = cbct_upload(ctx, ctx->d_angles, angles_deg, sizeof(float) * n))) return rc;
  ctx->nviews = n;
  ctx->views_h = std::move(vf);
  ctx->bpp_fit_cache.clear();
  // the event that marks the end of the calls reading this table
  if (!ctx->done) CBCT_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming));
  return CBCT_OK;
}

int cbct_views_count(cbct_ctx ctx) { return ctx ? ctx->nviews : 0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Element-wise kernels (grid-stride, 64-bit indices)
// ---------------------------------------------------------------------------
namespace {

constexpr int EW_BLOCK = 256;

inline dim3 ew_grid(size_t n) {
  size_t blocks = (n + EW_BLOCK - 1) / EW_BLOCK;
  if (blocks > 65536u * 4u) blocks = 65536u * 4u;
  if (blocks == 0) blocks = 1;
  return dim3(static_cast<unsigned int>(blocks));
}

#define GRID_STRIDE(i, n)                                                              \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n);            \
       i += (size_t)gridDim.x * blockDim.x)

__global__ void k_fill(size_t n, float *x, float v) { GRID_STRIDE(i, n) x[i] = v; }
__global__ void k_scale(size_t n, float *x, float a) { GRID_STRIDE(i, n) x[i] *= a; }
__global__ void k_add_scalar(size_t n, float *x, float a) { GRID_STRIDE(i, n) x[i] += a; }
__global__ void k_axpy(size_t n, float a, const float *__restrict__ x, float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] = fmaf(a, x[i], y[i]);
}
__global__ void k_mul(size_t n, const float *__restrict__ a, float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] *= a[i];
}
__global__ void k_div(size_t n, const float *__restrict__ a, float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] /= a[i];
}
__global__ void k_sub(size_t n, const float *__restrict__ a, float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] -= a[i];
}
__global__ void k_add(size_t n, const float *__restrict__ a, float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] += a[i];
}
__global__ void k_clamp_min(size_t n, float *x, float lo) {
  GRID_STRIDE(i, n) x[i] = fmaxf(x[i], lo);
}
__global__ void k_reciprocal(size_t n, float *x) { GRID_STRIDE(i, n) x[i] = 1.0f / x[i]; }
__global__ void k_mul_periodic(size_t n, const float *__restrict__ a, size_t period,
                               float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] *= a[i % period];
}
__global__ void k_apply_mask(size_t n, float *x, const unsigned char *__restrict__ m) {
  GRID_STRIDE(i, n) x[i] = m[i] ? x[i] : 0.0f * x[i];
}
__global__ void k_to_bf16(size_t n, const float *__restrict__ x, unsigned short *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] = f32_to_bf16(x[i]);
}
__global__ void k_from_bf16(size_t n, const unsigned short *__restrict__ x, float *__restrict__ y) {
  GRID_STRIDE(i, n) y[i] = bf16_to_f32(x[i]);
}

__global__ void k_sart_residual(size_t n, float *__restrict__ r, const float *__restrict__ b,
                                const float *__restrict__ w) {
  GRID_STRIDE(i, n) r[i] = -(r[i] - b[i]) * w[i];
}
__global__ void k_sart_update(size_t n, float *__restrict__ x, const float *__restrict__ img,
                              const float *__restrict__ wx, float step, int nonneg) {
  GRID_STRIDE(i, n) {
    float v = fmaf(step, img[i] * wx[i], x[i]);
    x[i] = nonneg ? fmaxf(v, 0.0f) : v;
  }
}
__global__ void k_mom_gradient_step(size_t n, float *__restrict__ z, const float *__restrict__ img,
                                    const float *__restrict__ p, float step) {
  GRID_STRIDE(i, n) z[i] = fmaf(step, img[i] * p[i], z[i]);
}
__global__ void k_mom_momentum(size_t n, float *__restrict__ x, const float *__restrict__ xold,
                               float *__restrict__ z, float c, int nonneg) {
  GRID_STRIDE(i, n) {
    float xv = x[i];
    if (nonneg) {
      xv = fmaxf(xv, 0.0f);
      x[i] = xv;
    }
    z[i] = fmaf(-c, xold[i], xv * (1.0f + c));
  }
}
__global__ void k_update_g(size_t n, const float *__restrict__ g, float *__restrict__ x,
                           const float *__restrict__ s, const float *__restrict__ p, float tau,
                           float tl) {
  GRID_STRIDE(i, n) {
    float xv = fmaf(-tau, g[i], x[i]);
    xv = fmaf(tl, s[i] / p[i], xv);
    x[i] = xv / ((1.0f / p[i]) * tl + 1.0f);
  }
}

// Two-pass deterministic FP64 reduction.
constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_BLOCKS = 2048;

template <int OP>
__global__ void __launch_bounds__(RED_BLOCK) k_reduce_partial(size_t n, const float *__restrict__ a,
                                                              const float *__restrict__ b,
                                                              double *__restrict__ partial) {
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    double va = a[i];
    if (OP == CBCT_DOT) acc += va * (double)b[i];
    else if (OP == CBCT_NRM2) acc += va * va;
    else if (OP == CBCT_ASUM) acc += fabs(va);
    else acc += va;
  }
  __shared__ double sm[RED_BLOCK];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = RED_BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(RED_BLOCK) k_reduce_final(int nparts, const double *__restrict__ partial,
                                                            double *__restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += RED_BLOCK) acc += partial[i];
  __shared__ double sm[RED_BLOCK];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = RED_BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

}  // namespace

#define EW_PRE(ctx, n, ...)                                          \
  if (!(ctx)) return CBCT_ERR_ARG;                                   \
  if ((n) == 0) return CBCT_OK;                                      \
  if (!(__VA_ARGS__))                                                \
    return cbct_fail((ctx), CBCT_ERR_ARG, "%s: illegal array pointer provided", __func__); \
  {                                                                  \
    int _rc = cbct_set_device(ctx);                                  \
    if (_rc) return _rc;                                             \
  }

#define EW_LAUNCH(ctx, kern, n, stream, ...)                                     \
  hipLaunchKernelGGL(kern, ew_grid(n), dim3(EW_BLOCK), 0, as_stream(stream), __VA_ARGS__); \
  CBCT_LAUNCH_CHECK(ctx, #kern);                                                 \
  return CBCT_OK;

extern "C" {

int cbct_fill(cbct_ctx ctx, size_t n, float *x, float v, void *st) {
  EW_PRE(ctx, n, x);
  EW_LAUNCH(ctx, k_fill, n, st, n, x, v);
}
int cbct_scale(cbct_ctx ctx, size_t n, float *x, float a, void *st) {
  EW_PRE(ctx, n, x);
  EW_LAUNCH(ctx, k_scale, n, st, n, x, a);
}
int cbct_add_scalar(cbct_ctx ctx, size_t n, float *x, float a, void *st) {
  EW_PRE(ctx, n, x);
  EW_LAUNCH(ctx, k_add_scalar, n, st, n, x, a);
}
int cbct_axpy(cbct_ctx ctx, size_t n, float a, const float *x, float *y, void *st) {
  EW_PRE(ctx, n, x && y);
  EW_LAUNCH(ctx, k_axpy, n, st, n, a, x, y);
}
int cbct_mul(cbct_ctx ctx, size_t n, const float *a, float *y, void *st) {
  EW_PRE(ctx, n, a && y);
  EW_LAUNCH(ctx, k_mul, n, st, n, a, y);
}
int cbct_div(cbct_ctx ctx, size_t n, const float *a, float *y, void *st) {
  EW_PRE(ctx, n, a && y);
  EW_LAUNCH(ctx, k_div, n, st, n, a, y);
}
int cbct_sub(cbct_ctx ctx, size_t n, const float *a, float *y, void *st) {
  EW_PRE(ctx, n, a && y);
  EW_LAUNCH(ctx, k_sub, n, st, n, a, y);
}
int cbct_add(cbct_ctx ctx, size_t n, const float *a, float *y, void *st) {
  EW_PRE(ctx, n, a && y);
  EW_LAUNCH(ctx, k_add, n, st, n, a, y);
}
int cbct_clamp_min(cbct_ctx ctx, size_t n, float *x, float lo, void *st) {
  EW_PRE(ctx, n, x);
  EW_LAUNCH(ctx, k_clamp_min, n, st, n, x, lo);
}
int cbct_reciprocal(cbct_ctx ctx, size_t n, float *x, void *st) {
  EW_PRE(ctx, n, x);
  EW_LAUNCH(ctx, k_reciprocal, n, st, n, x);
}
int cbct_mul_periodic(cbct_ctx ctx, size_t n, const float *a, size_t period, float *y, void *st) {
  EW_PRE(ctx, n, a && y && period > 0);
  EW_LAUNCH(ctx, k_mul_periodic, n, st, n, a, period, y);
}
int cbct_apply_mask(cbct_ctx ctx, size_t n, float *x, const unsigned char *m, void *st) {
  EW_PRE(ctx, n, x && m);
  EW_LAUNCH(ctx, k_apply_mask, n, st, n, x, m);
}
int cbct_to_bf16(cbct_ctx ctx, size_t n, const float *x, unsigned short *y, void *st) {
  EW_PRE(ctx, n, x && y);
  EW_LAUNCH(ctx, k_to_bf16, n, st, n, x, y);
}
int cbct_from_bf16(cbct_ctx ctx, size_t n, const unsigned short *x, float *y, void *st) {
  EW_PRE(ctx, n, x && y);
  EW_LAUNCH(ctx, k_from_bf16, n, st, n, x, y);
}
int cbct_sart_residual(cbct_ctx ctx, size_t n, float *r, const float *b, const float *w, void *st) {
  EW_PRE(ctx, n, r && b && w);
  EW_LAUNCH(ctx, k_sart_residual, n, st, n, r, b, w);
}
int cbct_sart_update(cbct_ctx ctx, size_t n, float *x, const float *img, const float *wx,
                     float step, int nonneg, void *st) {
  EW_PRE(ctx, n, x && img && wx);
  EW_LAUNCH(ctx, k_sart_update, n, st, n, x, img, wx, step, nonneg);
}
int cbct_mom_gradient_step(cbct_ctx ctx, size_t n, float *z, const float *img, const float *p,
                           float step, void *st) {
  EW_PRE(ctx, n, z && img && p);
  EW_LAUNCH(ctx, k_mom_gradient_step, n, st, n, z, img, p, step);
}
int cbct_mom_momentum(cbct_ctx ctx, size_t n, float *x, const float *xold, float *z, float c,
                      int nonneg, void *st) {
  EW_PRE(ctx, n, x && xold && z);
  EW_LAUNCH(ctx, k_mom_momentum, n, st, n, x, xold, z, c, nonneg);
}
int cbct_update_g(cbct_ctx ctx, size_t n, const float *g, float *x, const float *s,
                  const float *p, float tau, float avg_lambda, void *st) {
  EW_PRE(ctx, n, g && x && s && p);
  if (!(avg_lambda != 0.0f))
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_update_g: avg_lambda must be non-zero");
  EW_LAUNCH(ctx, k_update_g, n, st, n, g, x, s, p, tau, tau / avg_lambda);
}

int cbct_reduce_f64(cbct_ctx ctx, int op, size_t n, const float *a, const float *b, double *out,
                    void *st) {
  if (!ctx) return CBCT_ERR_ARG;
  if (!out) return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_reduce_f64: null output");
  if (op < 0 || op > 3) return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_reduce_f64: unknown op %d", op);
  if (n == 0) {
    *out = 0.0;
    return CBCT_OK;
  }
  if (!a || (op == CBCT_DOT && !b))
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_reduce_f64: illegal array pointer provided");
  int rc = cbct_set_device(ctx);
  if (rc) return rc;
  hipStream_t s = as_stream(st);
  CallScope scope(ctx, s);
  rc = cbct_reserve(ctx, ctx->ws->red, sizeof(double) * (RED_MAX_BLOCKS + 8));
  if (rc) return rc;
  size_t blocks = (n + RED_BLOCK - 1) / RED_BLOCK;
  if (blocks > RED_MAX_BLOCKS) blocks = RED_MAX_BLOCKS;
  double *partial = static_cast<double *>(ctx->ws->red.ptr);
  double *res = partial + RED_MAX_BLOCKS;
  switch (op) {
    case CBCT_DOT:
      hipLaunchKernelGGL(k_reduce_partial<CBCT_DOT>, dim3(blocks), dim3(RED_BLOCK), 0, s, n, a, b, partial);
      break;
    case CBCT_NRM2:
      hipLaunchKernelGGL(k_reduce_partial<CBCT_NRM2>, dim3(blocks), dim3(RED_BLOCK), 0, s, n, a, b, partial);
      break;
    case CBCT_ASUM:
      hipLaunchKernelGGL(k_reduce_partial<CBCT_ASUM>, dim3(blocks), dim3(RED_BLOCK), 0, s, n, a, b, partial);
      break;
    default:
      hipLaunchKernelGGL(k_reduce_partial<CBCT_SUM>, dim3(blocks), dim3(RED_BLOCK), 0, s, n, a, b, partial);
  }
  CBCT_LAUNCH_CHECK(ctx, "k_reduce_partial");
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BLOCK), 0, s, static_cast<int>(blocks), partial, res);
  CBCT_LAUNCH_CHECK(ctx, "k_reduce_final");
  CBCT_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_red, res, sizeof(double), hipMemcpyDeviceToHost, s));
  CBCT_HIP_CHECK(ctx, hipStreamSynchronize(s));
  double v = ctx->h_red[0];
  *out = (op == CBCT_NRM2) ? std::sqrt(v) : v;
  return CBCT_OK;
}

}  // extern "C"
