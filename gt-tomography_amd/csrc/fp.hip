// Ray-driven cone-beam forward projector for gfx950.
//
// Math: xray/conebeam_projection.cu:412-528 (one ray per detector pixel, the
// source->pixel segment clipped to the volume box by the slab method,
// n = int(samples_per_pixel*nx) left-Riemann trilinear samples, result scaled
// by the step length in mm).
//
// Design
//  * The per-ray setup (rotation, slab clipping, step length, start point and
//    increment in texel space) runs in FP64, once per ray.  Only the sample
//    loop is FP32.  Each sample position is p0 + m*q with one FMA per axis, so
//    coordinate error does not grow along the ray.
//  * Interpolation is done by hand (no texture unit): the volume is first
//    copied into a zero-bordered layout (1 voxel before, 2 after on every
//    axis) so the inner loop needs no bounds tests -- coordinates are
//    clamped to [-1, n] and the border supplies the zeros of the reference's
//    border addressing mode.
//  * Storage may be FP32 or BF16 (bf16 halves the gather bytes; arithmetic is
//    FP32 either way).
//  * One thread per ray, 16x16 detector tiles per workgroup (a wave covers a
//    16x4 pixel tile), so the 64 samples a wave issues at step m are close
//    together in the volume and the gathers share cache lines.
#include "cbct_internal.h"

namespace {

struct FPParams {
  int nu, nv, n;
  int nx, ny, nz;
  unsigned px, pxy;           // padded strides (elements)
  double lx, ly, lz;
  double fu, fv, sad, sdd;
  float cx, cy, cz;           // clamp limits in padded texel space (= n_axis + 1)
};

template <typename T>
__device__ __forceinline__ float ld(const T *p, unsigned i);
template <>
__device__ __forceinline__ float ld<float>(const float *p, unsigned i) {
  return p[i];
}
template <>
__device__ __forceinline__ float ld<unsigned short>(const unsigned short *p, unsigned i) {
  return bf16_to_f32(p[i]);
}

// Per-ray setup in FP64.  Returns the padded-texel start (p) and per-sample
// increment (q) and the step length in mm.
__device__ __forceinline__ void ray_setup(const FPParams &P, const ViewD &vw, int i, int j,
                                          float p[3], float q[3], float *step) {
  const double c = vw.c, s = vw.s;
  const double S0 = s * P.sad, S1 = -c * P.sad;  // R(theta) (0,-SAD,0)
  const double pu = ((i + 0.5) / P.nu - 0.5) * P.fu + vw.off_u;
  const double pv = ((j + 0.5) / P.nv - 0.5) * P.fv + vw.off_v;
  const double add = P.sdd - P.sad;
  const double d0 = (c * pu - s * add) - S0;
  const double d1 = (s * pu + c * add) - S1;
  const double d2 = pv;
  const double hx = P.lx * 0.5, hy = P.ly * 0.5, hz = P.lz * 0.5;
  double tA = (hx - S0) / d0, tB = (-hx - S0) / d0;
  double lo = fmin(tB, tA), hi = fmax(tB, tA);
  tA = (hy - S1) / d1;
  tB = (-hy - S1) / d1;
  lo = fmax(lo, fmin(tB, tA));
  hi = fmin(hi, fmax(tB, tA));
  tA = hz / d2;
  tB = -hz / d2;
  lo = fmax(lo, fmin(tB, tA));
  hi = fmin(hi, fmax(tB, tA));
  const double a1 = fmax(lo, 0.0);
  const double a2 = fmin(hi, 1.0);
  const double span = a2 - a1;
  const double h = span / P.n;
  *step = (float)(sqrt(d0 * d0 + d1 * d1 + d2 * d2) * fabs(span) / P.n);
  // texel = normalised*N - 0.5, plus 1 for the leading zero border
  p[0] = (float)(((S0 + a1 * d0) / P.lx + 0.5) * P.nx + 0.5);
  p[1] = (float)(((S1 + a1 * d1) / P.ly + 0.5) * P.ny + 0.5);
  p[2] = (float)(((a1 * d2) / P.lz + 0.5) * P.nz + 0.5);
  q[0] = (float)(d0 / P.lx * h * P.nx);
  q[1] = (float)(d1 / P.ly * h * P.ny);
  q[2] = (float)(d2 / P.lz * h * P.nz);
}

template <typename T>
__global__ void __launch_bounds__(256) k_fp_ray(FPParams P, const ViewD *__restrict__ views,
                                                const T *__restrict__ vol, float *__restrict__ proj,
                                                int accumulate) {
  const int i = blockIdx.x * 16 + threadIdx.x;
  const int j = blockIdx.y * 16 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= P.nu || j >= P.nv) return;
  const ViewD vw = views[k];
  float p[3], q[3], step;
  ray_setup(P, vw, i, j, p, q, &step);
  float acc = 0.0f;
  const unsigned px = P.px, pxy = P.pxy;
  for (int m = 0; m < P.n; ++m) {
    const float fm = (float)m;
    float sx = fminf(fmaxf(fmaf(fm, q[0], p[0]), 0.0f), P.cx);
    float sy = fminf(fmaxf(fmaf(fm, q[1], p[1]), 0.0f), P.cy);
    float sz = fminf(fmaxf(fmaf(fm, q[2], p[2]), 0.0f), P.cz);
    const float flx = floorf(sx), fly = floorf(sy), flz = floorf(sz);
    const float ax = sx - flx, ay = sy - fly, az = sz - flz;
    const unsigned base = (unsigned)flx + px * (unsigned)fly + pxy * (unsigned)flz;
    const float v000 = ld(vol, base), v100 = ld(vol, base + 1);
    const float v010 = ld(vol, base + px), v110 = ld(vol, base + px + 1);
    const float v001 = ld(vol, base + pxy), v101 = ld(vol, base + pxy + 1);
    const float v011 = ld(vol, base + pxy + px), v111 = ld(vol, base + pxy + px + 1);
    const float c00 = fmaf(ax, v100 - v000, v000);
    const float c10 = fmaf(ax, v110 - v010, v010);
    const float c01 = fmaf(ax, v101 - v001, v001);
    const float c11 = fmaf(ax, v111 - v011, v011);
    const float c0 = fmaf(ay, c10 - c00, c00);
    const float c1 = fmaf(ay, c11 - c01, c01);
    acc += fmaf(az, c1 - c0, c0);
  }
  const size_t idx = (size_t)i + (size_t)P.nu * ((size_t)j + (size_t)P.nv * k);
  const float r = acc * step;
  proj[idx] = accumulate ? proj[idx] + r : r;
}

// Exact transpose of k_fp_ray: scatter y*step*w into the padded accumulator.
__global__ void __launch_bounds__(256) k_fp_ray_transpose(FPParams P, const ViewD *__restrict__ views,
                                                          const float *__restrict__ proj,
                                                          float *__restrict__ acc_pad) {
  const int i = blockIdx.x * 16 + threadIdx.x;
  const int j = blockIdx.y * 16 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= P.nu || j >= P.nv) return;
  const ViewD vw = views[k];
  float p[3], q[3], step;
  ray_setup(P, vw, i, j, p, q, &step);
  const float y = proj[(size_t)i + (size_t)P.nu * ((size_t)j + (size_t)P.nv * k)] * step;
  if (y == 0.0f) return;
  const unsigned px = P.px, pxy = P.pxy;
  for (int m = 0; m < P.n; ++m) {
    const float fm = (float)m;
    float sx = fminf(fmaxf(fmaf(fm, q[0], p[0]), 0.0f), P.cx);
    float sy = fminf(fmaxf(fmaf(fm, q[1], p[1]), 0.0f), P.cy);
    float sz = fminf(fmaxf(fmaf(fm, q[2], p[2]), 0.0f), P.cz);
    const float flx = floorf(sx), fly = floorf(sy), flz = floorf(sz);
    const float ax = sx - flx, ay = sy - fly, az = sz - flz;
    const unsigned base = (unsigned)flx + px * (unsigned)fly + pxy * (unsigned)flz;
    const float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    atomicAdd(acc_pad + base, bx * by * bz * y);
    atomicAdd(acc_pad + base + 1, ax * by * bz * y);
    atomicAdd(acc_pad + base + px, bx * ay * bz * y);
    atomicAdd(acc_pad + base + px + 1, ax * ay * bz * y);
    atomicAdd(acc_pad + base + pxy, bx * by * az * y);
    atomicAdd(acc_pad + base + pxy + 1, ax * by * az * y);
    atomicAdd(acc_pad + base + pxy + px, bx * ay * az * y);
    atomicAdd(acc_pad + base + pxy + px + 1, ax * ay * az * y);
  }
}

// Copy a volume into the zero-bordered layout.
template <typename T>
__global__ void __launch_bounds__(256) k_pad_volume(const T *__restrict__ vol, T *__restrict__ out,
                                                    int nx, int ny, int nz, unsigned px,
                                                    unsigned py, unsigned pz) {
  const size_t total = (size_t)px * py * pz;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const unsigned a = (unsigned)(t % px);
    const unsigned b = (unsigned)((t / px) % py);
    const unsigned c = (unsigned)(t / ((size_t)px * py));
    const int x = (int)a - 1, y = (int)b - 1, z = (int)c - 1;
    T v = T(0);
    if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
      v = vol[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)];
    out[t] = v;
  }
}

// vol = (accumulate ? vol : 0) + interior(acc_pad)
__global__ void __launch_bounds__(256) k_unpad_add(const float *__restrict__ acc_pad,
                                                   float *__restrict__ vol, int nx, int ny, int nz,
                                                   unsigned px, unsigned pxy, int accumulate) {
  const size_t total = (size_t)nx * ny * nz;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const unsigned x = (unsigned)(t % nx);
    const unsigned y = (unsigned)((t / nx) % ny);
    const unsigned z = (unsigned)(t / ((size_t)nx * ny));
    const float v = acc_pad[(x + 1) + px * (y + 1) + (size_t)pxy * (z + 1)];
    vol[t] = accumulate ? vol[t] + v : v;
  }
}

__global__ void k_zero(size_t n, float *x) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    x[t] = 0.0f;
}

inline dim3 stride_grid(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  if (b == 0) b = 1;
  return dim3((unsigned)b);
}

int make_params(cbct_ctx ctx, const cbct_geom *g, float spp, FPParams &P) {
  const float fn = spp * (float)g->nx;
  if (!(fn >= 1.0f) || fn > 1e6f)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: conebeam_forwards_projection: samples per ray %g out of range", (double)fn);
  P.nu = g->nu;
  P.nv = g->nv;
  P.n = (int)fn;
  P.nx = g->nx;
  P.ny = g->ny;
  P.nz = g->nz;
  const size_t px = padded_dim(g->nx), py = padded_dim(g->ny), pz = padded_dim(g->nz);
  if (px * py * pz >= (1ull << 32) - 1)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "Error: conebeam_forwards_projection: volume too large for 32-bit indexing");
  P.px = (unsigned)px;
  P.pxy = (unsigned)(px * py);
  P.lx = g->lx;
  P.ly = g->ly;
  P.lz = g->lz;
  P.fu = g->fov_u;
  P.fv = g->fov_v;
  P.sad = g->sad;
  P.sdd = g->sdd;
  P.cx = (float)(g->nx + 1);
  P.cy = (float)(g->ny + 1);
  P.cz = (float)(g->nz + 1);
  return CBCT_OK;
}

}  // namespace

extern "C" {

int cbct_fp(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *vol, int dtype,
            float *proj, float spp, int accumulate, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "conebeam_forwards_projection";
  if (!vol || !proj)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (dtype != CBCT_F32 && dtype != CBCT_BF16)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "Error: %s: unsupported volume dtype %d", who, dtype);
  if (nviews == 0) return CBCT_OK;
  if (nviews > 65535)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: at most 65535 views per call", who);
  FPParams P;
  rc = make_params(ctx, g, spp, P);
  if (rc) return rc;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  const size_t esz = dtype == CBCT_F32 ? 4 : 2;
  const size_t npad = (size_t)P.pxy * padded_dim(g->nz);
  rc = cbct_reserve(ctx, ctx->vol_pad, npad * esz);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  const unsigned px = P.px, py = (unsigned)padded_dim(g->ny), pz = (unsigned)padded_dim(g->nz);
  dim3 block(16, 16);
  dim3 grid(ceil_div(g->nu, 16), ceil_div(g->nv, 16), nviews);
  if (dtype == CBCT_F32) {
    float *vp = static_cast<float *>(ctx->vol_pad.ptr);
    hipLaunchKernelGGL(k_pad_volume<float>, stride_grid(npad), dim3(256), 0, s,
                       static_cast<const float *>(vol), vp, g->nx, g->ny, g->nz, px, py, pz);
    CBCT_LAUNCH_CHECK(ctx, "k_pad_volume");
    cbct_prof_mark(ctx, PROF_FP, 0, s);
    hipLaunchKernelGGL(k_fp_ray<float>, grid, block, 0, s, P, ctx->d_views_d + view0, vp, proj,
                       accumulate);
    cbct_prof_mark(ctx, PROF_FP, 1, s);
  } else {
    unsigned short *vp = static_cast<unsigned short *>(ctx->vol_pad.ptr);
    hipLaunchKernelGGL(k_pad_volume<unsigned short>, stride_grid(npad), dim3(256), 0, s,
                       static_cast<const unsigned short *>(vol), vp, g->nx, g->ny, g->nz, px, py, pz);
    CBCT_LAUNCH_CHECK(ctx, "k_pad_volume");
    cbct_prof_mark(ctx, PROF_FP, 0, s);
    hipLaunchKernelGGL(k_fp_ray<unsigned short>, grid, block, 0, s, P, ctx->d_views_d + view0, vp,
                       proj, accumulate);
    cbct_prof_mark(ctx, PROF_FP, 1, s);
  }
  CBCT_LAUNCH_CHECK(ctx, "k_fp_ray");
  return CBCT_OK;
}

int cbct_fp_transpose(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const float *proj,
                      float *vol, float spp, int accumulate, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "cbct_fp_transpose";
  if (!vol || !proj)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (nviews > 65535)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: at most 65535 views per call", who);
  FPParams P;
  rc = make_params(ctx, g, spp, P);
  if (rc) return rc;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  const size_t npad = (size_t)P.pxy * padded_dim(g->nz);
  rc = cbct_reserve(ctx, ctx->vol_pad, npad * sizeof(float));
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  float *acc = static_cast<float *>(ctx->vol_pad.ptr);
  hipLaunchKernelGGL(k_zero, stride_grid(npad), dim3(256), 0, s, npad, acc);
  CBCT_LAUNCH_CHECK(ctx, "k_zero");
  if (nviews > 0) {
    dim3 block(16, 16);
    dim3 grid(ceil_div(g->nu, 16), ceil_div(g->nv, 16), nviews);
    hipLaunchKernelGGL(k_fp_ray_transpose, grid, block, 0, s, P, ctx->d_views_d + view0, proj, acc);
    CBCT_LAUNCH_CHECK(ctx, "k_fp_ray_transpose");
  }
  const size_t nvox = (size_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_unpad_add, stride_grid(nvox), dim3(256), 0, s, acc, vol, g->nx, g->ny,
                     g->nz, P.px, P.pxy, accumulate);
  CBCT_LAUNCH_CHECK(ctx, "k_unpad_add");
  return CBCT_OK;
}

}  // extern "C"
