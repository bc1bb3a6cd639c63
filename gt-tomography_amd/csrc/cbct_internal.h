// Internal declarations shared by the libcbct translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "cbct.h"

// Per-view projector constants.  Double precision copies feed the per-ray
// setup of the forward projector; float copies feed the inner loops.
struct ViewD {
  double c, s, off_u, off_v;
};
struct ViewF {
  float c, s, off_u, off_v;
};

struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Device workspaces (grow-only until trimmed) of one context, or of several
// that share them (cbct_ctx_share_workspace), with the event that marks the
// end of the last call that used them.  Freed with the last context holding
// them, after that call has finished.
struct Workspace {
  int device = 0;
  DeviceBuffer vol_pad;     // padded / repacked volume for the forward projector
  DeviceBuffer proj_pad;    // padded projections for the backprojector
  DeviceBuffer red;         // reduction partials
  DeviceBuffer filt_tmp;    // filter staging
  DeviceBuffer ssim_ws;     // MS-SSIM moment fields and taps
  DeviceBuffer bp_rects;    // backprojector: per (block, view) detector rectangles
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;   // stream of the last call (compared by value only)
  bool recorded = false;
  void free_buffers();
  ~Workspace();
};

struct cbct_ctx_s {
  int device = 0;
  std::string err;
  // view table
  int nviews = 0;
  std::vector<float> angles_deg;
  std::vector<float> offsets;  // interleaved (u, v)
  std::vector<ViewF> views_h;  // host copy of the float view table
  ViewD *d_views_d = nullptr;
  ViewF *d_views_f = nullptr;
  float *d_angles = nullptr;  // degrees, for the Parker weights
  std::shared_ptr<Workspace> ws;
  // end of the last call that read this context's tables (views, filter)
  hipEvent_t done = nullptr;
  bool done_recorded = false;
  // filter state
  int filter_nu = 0;
  int filter_engine = 2;
  int fp_engine = 1;             // 1: detector-column sharing, 0: one thread per ray
  int bp_engine = 2;             // 2: LDS-staged, 3: LDS-staged forced (BF16), 1: quad, 0: plain
  // backprojector: whether the LDS tile of the BF16 pair kernel holds the
  // blocks' detector rectangles, per (geometry, views, z-depth); cleared
  // with the view table
  std::vector<std::pair<std::vector<float>, bool>> bpp_fit_cache;
  int bp_kz = 0;                 // z-voxels per BP thread: 0 = chosen by size, else 8 or 16
  int fp_max_rpl = 8;            // column engine: most detector rows per lane
  int fp_view_group = 6;         // column engine: adjacent views launched together per column
  int fp_force_variant = 0;      // autotuner trials: variant and tile forced when > 0
  int fp_force_rpl = 0;
  int fp_force_vg = 0;
  int work_dtype = CBCT_F32;     // type of the projectors' working copies of FP32 arrays
  float *d_h = nullptr;          // fp32 spatial kernel, 2*nu entries (h[|m|] mirrored)
  std::vector<double> h64;       // the same kernel in FP64, h[0..nu-1]
  unsigned short *d_thi = nullptr; // fp16 hi Toeplitz row, 8 shifted copies (built on first use)
  unsigned short *d_tlo = nullptr; // fp16 lo parts, the same layout
  int toe_ld = 0;                // length of one copy (halves)
  float toe_inv_scale = 1.0f;    // inverse of the power-of-two scale applied to them
  float *d_fft = nullptr;        // FFT engine: twiddles W_N^m (N complex), then C[k] / N (N floats)
  int fft_log2n = 0;             // log2 N of those tables
  // cosine weights cache
  float *d_cosw = nullptr;
  cbct_geom cosw_geom{};
  bool cosw_valid = false;
  double *h_red = nullptr;       // pinned host slot for reduction results
  // kernel timing: event pairs for the FP, BP and filter kernels
  bool profile = false;
  hipEvent_t ev[3][2] = {};
  bool ev_recorded[3] = {false, false, false};
};

// A call that uses a context's tables or workspace: on entry the stream waits
// for the workspace's previous call when that was made on another stream; on
// exit the end of the call is recorded for the context and the workspace
// (not while the stream is being captured into a graph).
struct CallScope {
  cbct_ctx ctx;
  hipStream_t s;
  bool capturing = false;
  CallScope(cbct_ctx c, hipStream_t st);
  ~CallScope();
};
// Wait (host) for the end of the last call that read the context's tables.
int cbct_wait_tables(cbct_ctx ctx);
// Host-to-device copy of a context's table, complete on return.  It goes
// through a non-blocking stream of the library's own (one per device), not
// the legacy default stream, so it neither waits for nor breaks the work or
// graph captures of other threads' streams.
int cbct_upload(cbct_ctx ctx, void *dst, const void *src, size_t bytes);

enum { PROF_FP = 0, PROF_BP = 1, PROF_FILTER = 2 };
// Record the start/stop event of a profiled kernel on `s` (no-op unless enabled).
void cbct_prof_mark(cbct_ctx ctx, int which, int stop, hipStream_t s);

// Error plumbing -----------------------------------------------------------
int cbct_fail(cbct_ctx ctx, int code, const char *fmt, ...);

#define CBCT_HIP_CHECK(ctx, expr)                                                    \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return cbct_fail((ctx), CBCT_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

#define CBCT_LAUNCH_CHECK(ctx, what)                                                 \
  do {                                                                               \
    hipError_t _e = hipGetLastError();                                               \
    if (_e != hipSuccess)                                                            \
      return cbct_fail((ctx), CBCT_ERR_HIP, "%s launch: %s", (what), hipGetErrorString(_e)); \
  } while (0)

// Workspace growth (device current must be ctx->device).
int cbct_reserve(cbct_ctx ctx, DeviceBuffer &b, size_t bytes);
int cbct_set_device(cbct_ctx ctx);
int cbct_check_geom(cbct_ctx ctx, const cbct_geom *g, const char *who);
int cbct_check_views(cbct_ctx ctx, int view0, int nviews, const char *who);

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// bf16 helpers --------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __uint_as_float(static_cast<unsigned int>(h) << 16);
}
__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<unsigned short>(u >> 16);
}

// Padded layouts -----------------------------------------------------------
// The forward projector reads a zero-bordered copy of the volume: one voxel
// of zeros before index 0 and two after index n-1 on every axis, so that a
// clamped texel coordinate in [-1, n] never needs a bounds test.
static inline size_t padded_dim(int n) { return static_cast<size_t>(n) + 3; }

// XCD-aware block order.  The hardware deals linear block ids round-robin to
// the 8 XCDs (block b runs on XCD b mod 8).  This maps b to a logical id such
// that XCD x processes the contiguous logical range [start_x, start_x +
// count_x): neighbouring logical blocks -- which read neighbouring data --
// then share one XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
  const unsigned q = nblocks >> 3, rem = nblocks & 7u;
  const unsigned x = b & 7u, idx = b >> 3;
  return x * q + (x < rem ? x : rem) + idx;
}

// Launch helpers
static inline unsigned int ceil_div(size_t a, size_t b) {
  return static_cast<unsigned int>((a + b - 1) / b);
}
