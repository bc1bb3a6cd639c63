// Ray-driven cone-beam forward projector for gfx950.
//
// Math: xray/conebeam_projection.cu:412-528 (one ray per detector pixel, the
// source->pixel segment clipped to the volume box by the slab method,
// n = int(samples_per_pixel*nx) left-Riemann trilinear samples, result scaled
// by the step length in mm).
//
// Common design
//  * The per-ray setup (rotation, slab clipping, step length, start point and
//    increment in texel space) runs in FP64, once per ray.  Only the sample
//    loop is FP32.  Each sample position is p0 + m*q with one FMA per axis, so
//    coordinate error does not grow along the ray.
//  * Interpolation is done by hand (no texture unit) from a zero-bordered copy
//    of the volume (1 voxel before, 2 after on every axis): coordinates are
//    clamped to [-1, n] and the border supplies the zeros of the reference's
//    border addressing mode, so the inner loops carry no bounds tests.
//  * Storage may be FP32 or BF16 (arithmetic is FP32 either way).
//
// Engine 1 (default): detector-column sharing.
//  All rays of one detector column (one view, one u) lie in the vertical plane
//  through the source.  Their x/y slab limits are identical, and the z slab
//  never raises a1 (its lower limit is <= 0), so unless a ray leaves through
//  the top or bottom face its (a1, a2) -- and therefore its x/y sample
//  positions -- equal the column's at every step m.  One wave owns a column:
//  per step it bilinearly interpolates the z-line at the shared (x, y) from a
//  z-fastest copy of the volume (coalesced loads, 64 consecutive z per
//  instruction) into LDS, and every ray then does one z-lerp from LDS.  The
//  arithmetic order (x-lerps, y-lerp, z-lerp) and sample coordinates are those
//  of engine 0, so both engines produce bitwise-identical projections.  Rays
//  clipped by the z slab run the per-ray loop.  Blocks are remapped so each
//  XCD walks a contiguous run of columns (neighbouring columns share z-lines
//  in that XCD's L2).
//
// Engine 0: one thread per ray, 16x16 detector tiles, x-fastest padded
// volume, 8 gathers per sample.  Also the basis of the exact transpose.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cbct_internal.h"

// Working copy of cbct_fp_noalloc: device memory the library owns for the
// life of the process (one copy per GPU, each at its own address), so that
// call allocates nothing.  Host code uses the per-device address from
// hipGetSymbolAddress, never this name.
constexpr size_t FP_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[FP_SCRATCH_BYTES];

namespace {

struct FPParams {
  int nu, nv, n;
  int nx, ny, nz;
  unsigned px, pxy;           // x-fastest padded strides (elements)
  unsigned pz, pzx;           // z-fastest padded strides: z + pz*(x + px*y)
  double lx, ly, lz;
  double fu, fv, sad, sdd;
  float cx, cy, cz;           // clamp limits in padded texel space (= n_axis + 1)
  unsigned vg;                // column engine: views per launch group (see col_block)
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

// Column part of the ray setup (shared by every ray with detector column i).
struct ColSetup {
  double S0, S1, d0, d1, lo, hi;  // lo/hi: slab limits from the x and y faces
};

__device__ __forceinline__ ColSetup col_setup(const FPParams &P, const ViewD &vw, int i) {
  ColSetup C;
  const double c = vw.c, s = vw.s;
  C.S0 = s * P.sad;  // R(theta) (0,-SAD,0)
  C.S1 = -c * P.sad;
  const double pu = ((i + 0.5) / P.nu - 0.5) * P.fu + vw.off_u;
  const double add = P.sdd - P.sad;
  C.d0 = (c * pu - s * add) - C.S0;
  C.d1 = (s * pu + c * add) - C.S1;
  const double hx = P.lx * 0.5, hy = P.ly * 0.5;
  double tA = (hx - C.S0) / C.d0, tB = (-hx - C.S0) / C.d0;
  C.lo = fmin(tB, tA);
  C.hi = fmax(tB, tA);
  tA = (hy - C.S1) / C.d1;
  tB = (-hy - C.S1) / C.d1;
  C.lo = fmax(C.lo, fmin(tB, tA));
  C.hi = fmin(C.hi, fmax(tB, tA));
  return C;
}

struct RaySetup {
  double a1, a2;
  float p[3], q[3], step;
};

// Per-ray completion of the setup (row j).
__device__ __forceinline__ RaySetup ray_finish(const FPParams &P, const ViewD &vw,
                                               const ColSetup &C, int j) {
  RaySetup R;
  const double pv = ((j + 0.5) / P.nv - 0.5) * P.fv + vw.off_v;
  const double d2 = pv;
  const double hz = P.lz * 0.5;
  const double tA = hz / d2, tB = -hz / d2;
  const double lo = fmax(C.lo, fmin(tB, tA));
  const double hi = fmin(C.hi, fmax(tB, tA));
  R.a1 = fmax(lo, 0.0);
  R.a2 = fmin(hi, 1.0);
  const double span = R.a2 - R.a1;
  const double h = span / P.n;
  R.step = (float)(sqrt(C.d0 * C.d0 + C.d1 * C.d1 + d2 * d2) * fabs(span) / P.n);
  // texel = normalised*N - 0.5, plus 1 for the leading zero border
  R.p[0] = (float)(((C.S0 + R.a1 * C.d0) / P.lx + 0.5) * P.nx + 0.5);
  R.p[1] = (float)(((C.S1 + R.a1 * C.d1) / P.ly + 0.5) * P.ny + 0.5);
  R.p[2] = (float)(((R.a1 * d2) / P.lz + 0.5) * P.nz + 0.5);
  R.q[0] = (float)(C.d0 / P.lx * h * P.nx);
  R.q[1] = (float)(C.d1 / P.ly * h * P.ny);
  R.q[2] = (float)(d2 / P.lz * h * P.nz);
  return R;
}

__device__ __forceinline__ float clampf(float v, float hi) { return fminf(fmaxf(v, 0.0f), hi); }

// ------------------------------------------------------------------ engine 0
template <typename T>
__global__ void __launch_bounds__(256) k_fp_ray(FPParams P, const ViewD *__restrict__ views,
                                                const T *__restrict__ vol, float *__restrict__ proj,
                                                int accumulate) {
  const int i = blockIdx.x * 16 + threadIdx.x;
  const int j = blockIdx.y * 16 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= P.nu || j >= P.nv) return;
  const ViewD vw = views[k];
  const RaySetup R = ray_finish(P, vw, col_setup(P, vw, i), j);
  float acc = 0.0f;
  const unsigned px = P.px, pxy = P.pxy;
  for (int m = 0; m < P.n; ++m) {
    const float fm = (float)m;
    const float sx = clampf(fmaf(fm, R.q[0], R.p[0]), P.cx);
    const float sy = clampf(fmaf(fm, R.q[1], R.p[1]), P.cy);
    const float sz = clampf(fmaf(fm, R.q[2], R.p[2]), P.cz);
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
  const float r = acc * R.step;
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
  const RaySetup R = ray_finish(P, vw, col_setup(P, vw, i), j);
  const float y = proj[(size_t)i + (size_t)P.nu * ((size_t)j + (size_t)P.nv * k)] * R.step;
  if (y == 0.0f) return;
  const unsigned px = P.px, pxy = P.pxy;
  for (int m = 0; m < P.n; ++m) {
    const float fm = (float)m;
    const float sx = clampf(fmaf(fm, R.q[0], R.p[0]), P.cx);
    const float sy = clampf(fmaf(fm, R.q[1], R.p[1]), P.cy);
    const float sz = clampf(fmaf(fm, R.q[2], R.p[2]), P.cz);
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

// ------------------------------------------------------------------ engine 1
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Raw buffer loads: 32-bit byte offsets, corner offsets in SGPRs, and zero
// for any offset past the end of the buffer.
template <typename T>
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  // the builtin returns the raw 32-bit word
  return __uint_as_float((unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ float bload<unsigned short>(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                                       unsigned soff) {
  return bf16_to_f32(__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0));
}

// Two consecutive elements (byte offset voff+soff, 4- or 8-byte aligned) as
// floats: one dwordx2 load for FP32, one dword load for BF16.
template <typename T>
__device__ __forceinline__ float2 bload2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float2 bload2<float>(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                                unsigned soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
  return make_float2(__uint_as_float((unsigned)v[0]), __uint_as_float((unsigned)v[1]));
}
template <>
__device__ __forceinline__ float2 bload2<unsigned short>(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                                         unsigned soff) {
  const unsigned w = (unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
  return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

// Two consecutive z as kept in registers by the cached column variant.
template <typename T>
struct PairWord;
template <>
struct PairWord<float> {
  using type = unsigned long long;   // both floats in one 64-bit word (kept in registers)
  __device__ static unsigned long long load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    return (unsigned long long)(unsigned)v[0] | ((unsigned long long)(unsigned)v[1] << 32);
  }
  __device__ static float2 unpack(unsigned long long w) {
    return make_float2(__uint_as_float((unsigned)w), __uint_as_float((unsigned)(w >> 32)));
  }
};
template <>
struct PairWord<unsigned short> {
  using type = unsigned;
  __device__ static unsigned load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
  }
  __device__ static float2 unpack(unsigned w) {
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
  }
};

// Order one wave's LDS writes before its later LDS reads of other lanes'
// data (and the reads before the next writes).  A wave's LDS operations
// complete in issue order, so only the compiler has to be held back.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Block order of the column engine.  xcd_remap gives every XCD a contiguous
// range of W.  Within it, the blocks of P.vg adjacent views of one (column,
// row group) come one after another, so they run at the same time on one XCD
// and their rays, close together in the volume, share that XCD's L2; views
// past the last whole group of P.vg follow view by view.  vg = 1: view by
// view (column fastest).
__device__ __forceinline__ void col_block(const FPParams &P, unsigned W, unsigned ngroups,
                                          unsigned nblocks, int &i, int &grp, int &k) {
  const unsigned per_view = (unsigned)P.nu * ngroups;
  const unsigned nviews = nblocks / per_view;
  const unsigned vg = P.vg;
  const unsigned whole = (nviews / vg) * vg * per_view;
  unsigned vin = 0, rest = W;
  if (W < whole) {
    vin = W % vg;
    rest = W / vg;
  }
  i = (int)(rest % (unsigned)P.nu);
  const unsigned r2 = rest / (unsigned)P.nu;
  grp = (int)(r2 % ngroups);
  k = W < whole ? (int)((r2 / ngroups) * vg + vin) : (int)(r2 / ngroups);
}

// One 64-lane workgroup per (detector column, row group, view).  RPL rays per
// lane: rows j = row0 + r*64 + lane.
//   SSTEP = false, ZP = 1: per-step setup on the vector unit, one z per lane
//     and load (FP32 storage: measured fastest there).
//   SSTEP = true, ZP = 2: integer step setup in scalar registers, no per-ray
//     z clamp (a column-level range check instead), two consecutive z per
//     lane and load (BF16 storage: one dword per lane instead of two 16-bit
//     loads, 1.5x faster there).
//   NCH > 0 (with SSTEP, ZP = 2): the z window of the whole ray (all its
//     steps, all shared rows) is fixed once; when it spans at most NCH
//     chunks of 128 z, the four corner z-lines of the current (x, y) cell
//     stay in registers while the ray stays in the cell.  FP32 copies: a
//     step into a neighbouring cell keeps the lines the two cells share (two
//     of four for a move along x or y, one for a diagonal move) and loads
//     only the others; the column's (x, y) is wave-uniform, so the case
//     split is a scalar branch, and the lines are kept as (00, 10, 01, 11)
//     of the cell and shifted by register moves.  BF16 copies: all four are
//     reloaded for the next step's cell right after this step's combine.
//     Every line load fetches all NCH chunks (no per-chunk branches); chunks
//     past the window are never read.
// All produce the same bits as the per-ray engine.
template <typename T, int RPL, bool SSTEP, int ZP, int NCH>
__device__ __forceinline__ void fp_col_body(const FPParams &P, const ViewD *__restrict__ views,
                                            const T *__restrict__ vt, float *__restrict__ proj,
                                            int accumulate, unsigned ngroups, unsigned nblocks,
                                            unsigned vt_bytes) {
  extern __shared__ float zl[];  // one interpolated z-line (whole 128-wide chunks)
  int i, grp, k;
  col_block(P, xcd_remap(blockIdx.x, nblocks), ngroups, nblocks, i, grp, k);
  const int lane = threadIdx.x;
  const ViewD vw = views[k];
  const ColSetup C = col_setup(P, vw, i);
  const double a1c = fmax(C.lo, 0.0);
  const double a2c = fmin(C.hi, 1.0);
  const bool col_ok = (a2c - a1c) >= 0.0;   // h >= 0 keeps tz monotone in the row

  const int row0 = grp * 64 * RPL;
  float pz[RPL], qz[RPL], acc[RPL], stp[RPL];
  unsigned shared_mask = 0, valid_mask = 0;
  int jmin = 0x7fffffff, jmax = -1;
  RaySetup Rc;  // the column's x/y coordinates (from any shared ray)
  Rc.p[0] = Rc.p[1] = Rc.q[0] = Rc.q[1] = 0.0f;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int j = row0 + r * 64 + lane;
    acc[r] = 0.0f;
    pz[r] = qz[r] = stp[r] = 0.0f;
    if (j < P.nv) {
      valid_mask |= 1u << r;
      const RaySetup R = ray_finish(P, vw, C, j);
      stp[r] = R.step;
      pz[r] = R.p[2];
      qz[r] = R.q[2];
      if (col_ok && R.a1 == a1c && R.a2 == a2c) {
        shared_mask |= 1u << r;
        jmin = min(jmin, j);
        jmax = max(jmax, j);
        Rc = R;
      }
    }
  }
  const int gjmin = wave_min_i(jmin);
  int gjmax = wave_max_i(jmax);

  RaySetup Rlo, Rhi;
  if (gjmax >= 0) {
    // x/y trace and the z range from the first/last shared rays (identical
    // FP64 code, so identical floats to the per-ray values)
    Rlo = ray_finish(P, vw, C, gjmin);
    Rhi = ray_finish(P, vw, C, gjmax);
    if (SSTEP) {
      // Every shared ray's z texel coordinate lies between the first and
      // last shared rays' (monotone in the row) and, being linear in the
      // step, between its values at the first and last step.  If all four
      // are inside [0, nz+1] the per-ray z clamp is the identity and is
      // dropped; otherwise the whole column takes the per-ray path.
      const float ml = (float)(P.n - 1);
      const float e0 = Rlo.p[2], e1 = fmaf(ml, Rlo.q[2], Rlo.p[2]);
      const float e2 = Rhi.p[2], e3 = fmaf(ml, Rhi.q[2], Rhi.p[2]);
      const bool z_ok = fminf(fminf(e0, e1), fminf(e2, e3)) >= 0.0f &&
                        fmaxf(fmaxf(e0, e1), fmaxf(e2, e3)) <= P.cz;
      if (!z_ok) {
        shared_mask = 0;
        gjmax = -1;
      }
    }
  }

  if (SSTEP && gjmax >= 0) {
    const float px0 = Rlo.p[0], qx = Rlo.q[0], py0 = Rlo.p[1], qy = Rlo.q[1];
    const float pzl = Rlo.p[2], qzl = Rlo.q[2], pzh = Rhi.p[2], qzh = Rhi.q[2];
#pragma unroll
    for (int r = 0; r < RPL; ++r)
      if (!(shared_mask & (1u << r))) {
        pz[r] = pzl;
        qz[r] = qzl;
      }
    const unsigned PZ = P.pz;
    const unsigned esz = sizeof(T);
    const unsigned so_x = PZ * esz, so_y = P.pzx * esz, so_xy = (P.pzx + PZ) * esz;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)vt, (short)0, (int)vt_bytes, 0x00020000);
    const unsigned lane_off = (unsigned)lane * esz;
    bool done = false;
    if constexpr (NCH > 0) {
      static_assert(ZP == 2, "cached z-lines use pair loads");
      // The window: per step the rays read z in [zlo(m), zhi(m)], both linear
      // in m through the same fmaf as below (so monotone), hence extreme at
      // the first or last step.
      const float ml = (float)(P.n - 1);
      const int zlo_min = __builtin_amdgcn_readfirstlane((int)fminf(pzl, fmaf(ml, qzl, pzl)));
      const int zhi_max = __builtin_amdgcn_readfirstlane((int)fmaxf(pzh, fmaf(ml, qzh, pzh))) + 1;
      const int zb0 = zlo_min & ~1;
      const int nchw = (zhi_max - zb0 + 128) >> 7;
      // (cell coordinates are packed as two 16-bit halves below)
      if (nchw <= NCH && P.cx < 65535.0f && P.cy < 65535.0f) {
        done = true;
        // the lines as loaded: FP32 pairs, or BF16 pairs packed in one word
        // (unpacked in the combine, which keeps the registers, and with them
        // the occupancy, of the uncached variant)
        using PW = PairWord<T>;
        using W = typename PW::type;
        W L00[NCH], L10[NCH], L01[NCH], L11[NCH];
        const unsigned zoff = (unsigned)zb0 * esz;
        int cx = -8, cy = -8;   // no cell loaded yet
        auto load_line = [&](W(&L)[NCH], unsigned B) __attribute__((always_inline)) {
#pragma unroll
          for (int c = 0; c < NCH; ++c) L[c] = PW::load(rsrc, 2 * lane_off + (unsigned)(128 * c) * esz, B);
        };
        auto copy_line = [&](W(&D)[NCH], const W(&S)[NCH]) __attribute__((always_inline)) {
#pragma unroll
          for (int c = 0; c < NCH; ++c) D[c] = S[c];
        };
        const float *zb = zl - zb0;
        // all four lines of cell (ix, iy)
        auto load_cell = [&](int ix, int iy) __attribute__((always_inline)) {
          const unsigned B = ((unsigned)iy * P.px + (unsigned)ix) * PZ * esz + zoff;
          load_line(L00, B); load_line(L10, B + so_x);
          load_line(L01, B + so_y); load_line(L11, B + so_xy);
        };
        // the four lines of cell (ix, iy), keeping those shared with the
        // current cell (FP32 copies, bound by their line bytes)
        auto update_cell = [&](int ix, int iy) __attribute__((always_inline)) {
            if (ix != cx || iy != cy) {
              const int dx = ix - cx, dy = iy - cy;
              const unsigned B = ((unsigned)iy * P.px + (unsigned)ix) * PZ * esz + zoff;
              if (dy == 0 && dx == 1) {
                copy_line(L00, L10); copy_line(L01, L11);
                load_line(L10, B + so_x); load_line(L11, B + so_xy);
              } else if (dy == 0 && dx == -1) {
                copy_line(L10, L00); copy_line(L11, L01);
                load_line(L00, B); load_line(L01, B + so_y);
              } else if (dx == 0 && dy == 1) {
                copy_line(L00, L01); copy_line(L10, L11);
                load_line(L01, B + so_y); load_line(L11, B + so_xy);
              } else if (dx == 0 && dy == -1) {
                copy_line(L01, L00); copy_line(L11, L10);
                load_line(L00, B); load_line(L10, B + so_x);
              } else if (dx == 1 && dy == 1) {
                copy_line(L00, L11);
                load_line(L10, B + so_x); load_line(L01, B + so_y); load_line(L11, B + so_xy);
              } else if (dx == 1 && dy == -1) {
                copy_line(L01, L10);
                load_line(L00, B); load_line(L10, B + so_x); load_line(L11, B + so_xy);
              } else if (dx == -1 && dy == 1) {
                copy_line(L10, L01);
                load_line(L00, B); load_line(L01, B + so_y); load_line(L11, B + so_xy);
              } else if (dx == -1 && dy == -1) {
                copy_line(L11, L00);
                load_line(L00, B); load_line(L10, B + so_x); load_line(L01, B + so_y);
              } else {
                load_line(L00, B); load_line(L10, B + so_x);
                load_line(L01, B + so_y); load_line(L11, B + so_xy);
              }
              cx = ix;
              cy = iy;
            }
        };
        // the combined z-line of the cell's lines into LDS
        auto combine = [&](const float ax, const float ay) __attribute__((always_inline)) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
              const float2 v00 = PairWord<T>::unpack(L00[c]), v10 = PairWord<T>::unpack(L10[c]);
              const float2 v01 = PairWord<T>::unpack(L01[c]), v11 = PairWord<T>::unpack(L11[c]);
              float2 o;
              {
                const float c0 = fmaf(ax, v10.x - v00.x, v00.x);
                const float c1 = fmaf(ax, v11.x - v01.x, v01.x);
                o.x = fmaf(ay, c1 - c0, c0);
              }
              {
                const float c0 = fmaf(ax, v10.y - v00.y, v00.y);
                const float c1 = fmaf(ax, v11.y - v01.y, v01.y);
                o.y = fmaf(ay, c1 - c0, c0);
              }
              *reinterpret_cast<float2 *>(&zl[2 * lane + 128 * c]) = o;
            }
            wave_lds_sync();
        };
        // every ray's z-lerp from the combined z-line
        auto rays = [&](const float fm) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
              const float tz = fmaf(fm, qz[r], pz[r]);  // in [0, nz+1]: checked above
              const int iz = (int)tz;
              const float az = __builtin_amdgcn_fractf(tz);
              const float a = zb[iz], b = zb[iz + 1];
              acc[r] += fmaf(az, b - a, a);
            }
            wave_lds_sync();
        };
        // one step with FP32 copies
        auto step = [&](const float fm, const float ax, const float ay, const int ix,
                        const int iy) __attribute__((always_inline)) {
          update_cell(ix, iy);
          combine(ax, ay);
          rays(fm);
        };
        // BF16 copies: the x/y part of a step, the same for the whole wave,
        // is evaluated for 64 steps at once, step mb + lane in each lane (the
        // same expressions as per step, so the same values), and each step
        // reads its cell and weights back into scalar registers.  The lines
        // of the next step's cell, when it changes, are all four reloaded
        // straight into their registers once this step's combine has used
        // them, so they arrive while this step's rays run (fewer
        // instructions than keeping the shared lines, whose register moves
        // wait on the new loads).  FP32 copies: per-step setup and shared
        // lines kept (measured faster there).
        if constexpr (sizeof(T) == 2) {
          {
            const float sx = __builtin_amdgcn_fmed3f(fmaf(0.0f, qx, px0), 0.0f, P.cx);
            const float sy = __builtin_amdgcn_fmed3f(fmaf(0.0f, qy, py0), 0.0f, P.cy);
            cx = __builtin_amdgcn_readfirstlane((int)sx);
            cy = __builtin_amdgcn_readfirstlane((int)sy);
            load_cell(cx, cy);
          }
          for (int mb = 0; mb < P.n; mb += 64) {
            const float fml = (float)(mb + lane);
            const float sxl = __builtin_amdgcn_fmed3f(fmaf(fml, qx, px0), 0.0f, P.cx);
            const float syl = __builtin_amdgcn_fmed3f(fmaf(fml, qy, py0), 0.0f, P.cy);
            const int axl = __float_as_int(__builtin_amdgcn_fractf(sxl));
            const int ayl = __float_as_int(__builtin_amdgcn_fractf(syl));
            // cell of step mb + lane + 1 (the next step's)
            const float fmn = (float)(mb + lane + 1);
            const float sxn = __builtin_amdgcn_fmed3f(fmaf(fmn, qx, px0), 0.0f, P.cx);
            const float syn = __builtin_amdgcn_fmed3f(fmaf(fmn, qy, py0), 0.0f, P.cy);
            const int cn = (int)((unsigned)sxn | ((unsigned)syn << 16));
            const int cnt = __builtin_amdgcn_readfirstlane(min(64, P.n - mb));
            for (int l = 0; l < cnt; ++l) {
              const float ax = __int_as_float(__builtin_amdgcn_readlane(axl, l));
              const float ay = __int_as_float(__builtin_amdgcn_readlane(ayl, l));
              combine(ax, ay);
              const unsigned c = (unsigned)__builtin_amdgcn_readlane(cn, l);
              const int nx = (int)(c & 0xffffu), ny = (int)(c >> 16);
              if (mb + l + 1 < P.n && (nx != cx || ny != cy)) {
                load_cell(nx, ny);
                cx = nx;
                cy = ny;
              }
              rays((float)(mb + l));
            }
          }
        }
        for (int m = 0; sizeof(T) != 2 && m < P.n; ++m) {
          const float fm = (float)m;
          const float sx = __builtin_amdgcn_fmed3f(fmaf(fm, qx, px0), 0.0f, P.cx);
          const float sy = __builtin_amdgcn_fmed3f(fmaf(fm, qy, py0), 0.0f, P.cy);
          const float ax = __builtin_amdgcn_fractf(sx), ay = __builtin_amdgcn_fractf(sy);
          const int ix = __builtin_amdgcn_readfirstlane((int)sx);
          const int iy = __builtin_amdgcn_readfirstlane((int)sy);
          step(fm, ax, ay, ix, iy);
        }
      }
    }
    for (int m = 0; !done && m < P.n; ++m) {
      // Uniform per-step values: the float part on the VALU (same
      // expressions as the per-ray loop), the integer part moved to scalar
      // registers so the addressing below is scalar.
      const float fm = (float)m;
      const float sx = __builtin_amdgcn_fmed3f(fmaf(fm, qx, px0), 0.0f, P.cx);
      const float sy = __builtin_amdgcn_fmed3f(fmaf(fm, qy, py0), 0.0f, P.cy);
      const float ax = __builtin_amdgcn_fractf(sx), ay = __builtin_amdgcn_fractf(sy);
      const int ix = __builtin_amdgcn_readfirstlane((int)sx);
      const int iy = __builtin_amdgcn_readfirstlane((int)sy);
      const int zlo = __builtin_amdgcn_readfirstlane((int)fmaf(fm, qzl, pzl));
      const int zhi = __builtin_amdgcn_readfirstlane((int)fmaf(fm, qzh, pzh)) + 1;
      // z-line chunks: corner bases in SGPRs, lane offsets in one VGPR.
      // Offsets past the column read the next column or the allocation's
      // slack; those LDS slots lie beyond zhi and are never read.
      int zb0;
      if constexpr (ZP == 1) {
        zb0 = zlo;
        const int nch = (zhi - zlo + 64) >> 6;
        const unsigned B = (((unsigned)iy * P.px + (unsigned)ix) * PZ + (unsigned)zlo) * esz;
        for (int c = 0; c < nch; ++c) {
          const unsigned voff = lane_off + (unsigned)(64 * c) * esz;
          const float v00 = bload<T>(rsrc, voff, B), v10 = bload<T>(rsrc, voff, B + so_x);
          const float v01 = bload<T>(rsrc, voff, B + so_y), v11 = bload<T>(rsrc, voff, B + so_xy);
          const float c0 = fmaf(ax, v10 - v00, v00);
          const float c1 = fmaf(ax, v11 - v01, v01);
          zl[lane + 64 * c] = fmaf(ay, c1 - c0, c0);
        }
      } else {
        // two consecutive z per lane and load (128 z per instruction); the
        // line starts at an even z so the pair loads are aligned
        zb0 = zlo & ~1;
        const int nch = (zhi - zb0 + 128) >> 7;
        const unsigned B = (((unsigned)iy * P.px + (unsigned)ix) * PZ + (unsigned)zb0) * esz;
        for (int c = 0; c < nch; ++c) {
          const unsigned voff = 2 * lane_off + (unsigned)(128 * c) * esz;
          const float2 v00 = bload2<T>(rsrc, voff, B), v10 = bload2<T>(rsrc, voff, B + so_x);
          const float2 v01 = bload2<T>(rsrc, voff, B + so_y), v11 = bload2<T>(rsrc, voff, B + so_xy);
          float2 o;
          {
            const float c0 = fmaf(ax, v10.x - v00.x, v00.x);
            const float c1 = fmaf(ax, v11.x - v01.x, v01.x);
            o.x = fmaf(ay, c1 - c0, c0);
          }
          {
            const float c0 = fmaf(ax, v10.y - v00.y, v00.y);
            const float c1 = fmaf(ax, v11.y - v01.y, v01.y);
            o.y = fmaf(ay, c1 - c0, c0);
          }
          *reinterpret_cast<float2 *>(&zl[2 * lane + 128 * c]) = o;
        }
      }
      wave_lds_sync();
      const float *zb = zl - zb0;
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const float tz = fmaf(fm, qz[r], pz[r]);  // in [0, nz+1]: checked above
        const int iz = (int)tz;
        const float az = __builtin_amdgcn_fractf(tz);
        const float a = zb[iz], b = zb[iz + 1];
        acc[r] += fmaf(az, b - a, a);
      }
      wave_lds_sync();
    }
  } else if (!SSTEP && gjmax >= 0) {
    const float px0 = Rlo.p[0], qx = Rlo.q[0], py0 = Rlo.p[1], qy = Rlo.q[1];
    const float pzl = Rlo.p[2], qzl = Rlo.q[2], pzh = Rhi.p[2], qzh = Rhi.q[2];
    // rays outside the shared set read harmless in-range LDS entries
#pragma unroll
    for (int r = 0; r < RPL; ++r)
      if (!(shared_mask & (1u << r))) {
        pz[r] = pzl;
        qz[r] = qzl;
      }
    // Coordinates here are finite and the clamps keep them in [0, n+1], so
    // med3 equals fminf(fmaxf()), truncation equals floor and fract(t)
    // equals t - floor(t): same values as the per-ray loop.
    const unsigned PZ = P.pz;
    const unsigned esz = sizeof(T);
    const unsigned so_x = PZ * esz, so_y = P.pzx * esz, so_xy = (P.pzx + PZ) * esz;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)vt, (short)0, (int)vt_bytes, 0x00020000);
    for (int m = 0; m < P.n; ++m) {
      const float fm = (float)m;
      const float sx = __builtin_amdgcn_fmed3f(fmaf(fm, qx, px0), 0.0f, P.cx);
      const float sy = __builtin_amdgcn_fmed3f(fmaf(fm, qy, py0), 0.0f, P.cy);
      const float ax = __builtin_amdgcn_fractf(sx), ay = __builtin_amdgcn_fractf(sy);
      const int zlo = (int)__builtin_amdgcn_fmed3f(fmaf(fm, qzl, pzl), 0.0f, P.cz);
      const int zhi = (int)__builtin_amdgcn_fmed3f(fmaf(fm, qzh, pzh), 0.0f, P.cz) + 1;
      const int nch = (zhi - zlo + 64) >> 6;
      const unsigned base = ((unsigned)(int)sy * P.px + (unsigned)(int)sx) * PZ + (unsigned)zlo;
      for (int c = 0; c < nch; ++c) {
        const unsigned zi = min(base + (unsigned)(lane + 64 * c), base - (unsigned)zlo + PZ - 1);
        const unsigned voff = zi * esz;
        const float v00 = bload<T>(rsrc, voff, 0), v10 = bload<T>(rsrc, voff, so_x);
        const float v01 = bload<T>(rsrc, voff, so_y), v11 = bload<T>(rsrc, voff, so_xy);
        const float c0 = fmaf(ax, v10 - v00, v00);
        const float c1 = fmaf(ax, v11 - v01, v01);
        zl[lane + 64 * c] = fmaf(ay, c1 - c0, c0);
      }
      wave_lds_sync();
      const float *zb = zl - zlo;
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const float tz = __builtin_amdgcn_fmed3f(fmaf(fm, qz[r], pz[r]), 0.0f, P.cz);
        const int iz = (int)tz;
        const float az = __builtin_amdgcn_fractf(tz);
        const float a = zb[iz], b = zb[iz + 1];
        acc[r] += fmaf(az, b - a, a);
      }
      wave_lds_sync();
    }
  }

  // Rays clipped by the z faces (or a column that misses the box): per-ray
  // loop on the z-fastest layout, same arithmetic order as engine 0.
  const unsigned todo = valid_mask & ~shared_mask;
  if (todo) {
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      if (!(todo & (1u << r))) continue;
      const int j = row0 + r * 64 + lane;
      const RaySetup R = ray_finish(P, vw, C, j);
      float a = 0.0f;
      for (int m = 0; m < P.n; ++m) {
        const float fm = (float)m;
        const float sx = clampf(fmaf(fm, R.q[0], R.p[0]), P.cx);
        const float sy = clampf(fmaf(fm, R.q[1], R.p[1]), P.cy);
        const float sz = clampf(fmaf(fm, R.q[2], R.p[2]), P.cz);
        const float flx = floorf(sx), fly = floorf(sy), flz = floorf(sz);
        const float ax = sx - flx, ay = sy - fly, az = sz - flz;
        const unsigned base = ((unsigned)fly * P.px + (unsigned)flx) * P.pz + (unsigned)flz;
        const float v000 = ld(vt, base), v100 = ld(vt, base + P.pz);
        const float v010 = ld(vt, base + P.pzx), v110 = ld(vt, base + P.pzx + P.pz);
        const float v001 = ld(vt, base + 1), v101 = ld(vt, base + P.pz + 1);
        const float v011 = ld(vt, base + P.pzx + 1), v111 = ld(vt, base + P.pzx + P.pz + 1);
        const float c00 = fmaf(ax, v100 - v000, v000);
        const float c10 = fmaf(ax, v110 - v010, v010);
        const float c01 = fmaf(ax, v101 - v001, v001);
        const float c11 = fmaf(ax, v111 - v011, v011);
        const float c0 = fmaf(ay, c10 - c00, c00);
        const float c1 = fmaf(ay, c11 - c01, c01);
        a += fmaf(az, c1 - c0, c0);
      }
      acc[r] = a;
    }
  }

#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    if (!(valid_mask & (1u << r))) continue;
    const int j = row0 + r * 64 + lane;
    const size_t idx = (size_t)i + (size_t)P.nu * ((size_t)j + (size_t)P.nv * k);
    const float v = acc[r] * stp[r];
    proj[idx] = accumulate ? proj[idx] + v : v;
  }
}

template <typename T, int RPL, bool SSTEP, int ZP>
__global__ void __launch_bounds__(64) k_fp_col(FPParams P, const ViewD *__restrict__ views,
                                               const T *__restrict__ vt, float *__restrict__ proj,
                                               int accumulate, unsigned ngroups, unsigned nblocks,
                                               unsigned vt_bytes) {
  fp_col_body<T, RPL, SSTEP, ZP, 0>(P, views, vt, proj, accumulate, ngroups, nblocks, vt_bytes);
}

// The cached variant, held to five waves per SIMD (five measured faster than
// four with more registers, and than six with BF16 copies).  With FP32
// copies and 6 rows per lane, five waves spill 3 VGPRs to scratch; four
// waves without the spill measured slower (162 against 143 ms at 512^3 x
// 720 views).
template <typename T, int RPL, int NCH>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
k_fp_colc(FPParams P, const ViewD *__restrict__ views, const T *__restrict__ vt,
          float *__restrict__ proj, int accumulate, unsigned ngroups, unsigned nblocks,
          unsigned vt_bytes) {
  fp_col_body<T, RPL, true, 2, NCH>(P, views, vt, proj, accumulate, ngroups, nblocks, vt_bytes);
}

// ------------------------------------------------------------ layouts
// Element conversion of the padded copies: identity, or FP32 -> BF16 when the
// context keeps its working copies in BF16.
template <typename TO, typename TI>
__device__ __forceinline__ TO to_storage(TI v) {
  return v;
}
template <>
__device__ __forceinline__ unsigned short to_storage<unsigned short, float>(float v) {
  return f32_to_bf16(v);
}

// Copy a volume into the x-fastest zero-bordered layout.
template <typename TI, typename T>
__global__ void __launch_bounds__(256) k_pad_volume(const TI *__restrict__ vol, T *__restrict__ out,
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
      v = to_storage<T>(vol[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]);
    out[t] = v;
  }
}

// Transposing copy into the z-fastest zero-bordered layout
// out[z' + pz*(x' + px*y')] = vol[x'-1, y'-1, z'-1] (0 outside).
// 32x32 (x', z') tiles through LDS so both the reads (along x) and the writes
// (along z) are coalesced.  grid = (ceil(pz/32), ceil(px/32), py), block 32x8.
template <typename TI, typename T>
__global__ void __launch_bounds__(256) k_pad_transpose(const TI *__restrict__ vol, T *__restrict__ out,
                                                       int nx, int ny, int nz, unsigned px,
                                                       unsigned pz) {
  __shared__ T tile[32][33];
  const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, yp = blockIdx.z;
  const int y = yp - 1;
  const int tx = threadIdx.x, ty = threadIdx.y;
  // read: rows of x for 32 z
  for (int r = ty; r < 32; r += 8) {
    const int zp = z0 + r, xp = x0 + tx;
    const int x = xp - 1, z = zp - 1;
    T v = T(0);
    if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
      v = to_storage<T>(vol[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]);
    tile[r][tx] = v;
  }
  __syncthreads();
  // write: columns of z for 32 x
  for (int r = ty; r < 32; r += 8) {
    const int xp = x0 + r, zp = z0 + tx;
    if (xp < (int)px && zp < (int)pz)
      out[(size_t)zp + (size_t)pz * ((size_t)xp + (size_t)px * yp)] = tile[tx][r];
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

// z-fastest padded length: nz + 3 rounded up to 16 elements
inline size_t zpadded(int nz) { return (padded_dim(nz) + 15) & ~size_t(15); }

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
  const size_t pzz = zpadded(g->nz);
  if (px * py * pzz >= (1ull << 32) - 1)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "Error: conebeam_forwards_projection: volume too large for 32-bit indexing");
  (void)pz;
  P.px = (unsigned)px;
  P.pxy = (unsigned)(px * py);
  P.pz = (unsigned)pzz;
  P.pzx = (unsigned)(pzz * px);
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
  P.vg = (unsigned)ctx->fp_view_group;
  return CBCT_OK;
}

template <typename T, int RPL, bool SSTEP, int ZP, int NCH>
void launch_col(const FPParams &P, hipStream_t s, const ViewD *views, const T *vt, float *proj,
                int acc, int nviews, unsigned vt_bytes) {
  const unsigned ngroups = (unsigned)((P.nv + 64 * RPL - 1) / (64 * RPL));
  // one z-line; the cached variants write all NCH chunks of their window
  const size_t zline = ((P.pz + 127) / 128) * 128;
  const size_t lds = (zline > (size_t)128 * NCH ? zline : (size_t)128 * NCH) * sizeof(float);
  // a launch covers at most 2^32 - 1 work-items: larger calls go out as
  // several launches over consecutive views (views are independent)
  const size_t per_view = (size_t)P.nu * ngroups;
  const int max_views = (int)std::max<size_t>(1, std::min<size_t>(nviews, ((1ull << 32) - 1) / 64 / per_view));
  for (int v0 = 0; v0 < nviews; v0 += max_views) {
    const int nvc = std::min(max_views, nviews - v0);
    const unsigned nblocks = (unsigned)(per_view * (size_t)nvc);
    const ViewD *vw = views + v0;
    float *pj = proj + (size_t)v0 * P.nu * P.nv;
    if constexpr (NCH > 0)
      hipLaunchKernelGGL((k_fp_colc<T, RPL, NCH>), dim3(nblocks), dim3(64), lds, s, P, vw, vt, pj, acc,
                         ngroups, nblocks, vt_bytes);
    else
      hipLaunchKernelGGL((k_fp_col<T, RPL, SSTEP, ZP>), dim3(nblocks), dim3(64), lds, s, P, vw, vt, pj,
                         acc, ngroups, nblocks, vt_bytes);
  }
}

// variant 1: vector step setup, one z per load; variant 2: scalar step
// setup, two z per load; variants 3 and 4: variant 2 with the corner z-lines
// cached in registers over a whole-ray window of at most 2 (3: 3) chunks of
// 128 z (k_fp_colc).
template <typename T, int RPL>
void launch_col_w(const FPParams &P, int variant, hipStream_t s, const ViewD *views, const T *vt,
                  float *proj, int acc, int nviews, unsigned vt_bytes) {
  switch (variant) {
    case 2: launch_col<T, RPL, true, 2, 0>(P, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 3: launch_col<T, RPL, true, 2, 2>(P, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 4: launch_col<T, RPL, true, 2, 3>(P, s, views, vt, proj, acc, nviews, vt_bytes); break;
    default: launch_col<T, RPL, false, 1, 0>(P, s, views, vt, proj, acc, nviews, vt_bytes); break;
  }
}

// Rows per lane that dispatch_col picks for max_rpl.
inline int rows_per_lane(int nv, int max_rpl) {
  const int need = (nv + 63) / 64;
  const int groups = (need + max_rpl - 1) / max_rpl;
  return (need + groups - 1) / groups;
}

// Launch with `rpl` rays per lane (1..12; rows_per_lane splits the rows
// into equal groups of whole 64-row blocks).  More rows per wave share one
// z-line load; fewer keep registers, and so occupancy, in check.
template <typename T>
void dispatch_col(const FPParams &P, int variant, int rpl, hipStream_t s, const ViewD *views,
                  const T *vt, float *proj, int acc, int nviews, unsigned vt_bytes) {
  switch (rpl) {
    case 1: launch_col_w<T, 1>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 2: launch_col_w<T, 2>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 3: launch_col_w<T, 3>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 4: launch_col_w<T, 4>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 5: launch_col_w<T, 5>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 6: launch_col_w<T, 6>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 7: launch_col_w<T, 7>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 8: launch_col_w<T, 8>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 9: launch_col_w<T, 9>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 10: launch_col_w<T, 10>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    case 11: launch_col_w<T, 11>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
    default: launch_col_w<T, 12>(P, variant, s, views, vt, proj, acc, nviews, vt_bytes); break;
  }
}

// Engine 1's choice.  The cached variant pays when every row group's
// whole-ray z window fits its two chunks.  A bound: a ray through detector
// height v (offset included) is at z = v t / SDD where t, its depth along the
// central ray, lies in [SAD - R, SAD + R] inside the box (R: half the box's
// xy diagonal); the window spans the extremes of v t over the group's rows
// and that depth range, plus the texel, the alignment and rounding slack.
// The largest tile (rows per lane, at least 2) whose groups fit is taken;
// if none does, the per-step variant measured fastest for the storage type
// with the configured tile.
struct FPChoice {
  int variant;
  int rpl;
  int vg = 0;   // views per launch group; 0 = the context's setting
};

double window_bound(const cbct_geom *g, int nv, int rpl, float omin, float omax) {
  const double R = 0.5 * std::sqrt((double)g->lx * g->lx + (double)g->ly * g->ly);
  const double t0 = (double)g->sad - R, t1 = (double)g->sad + R;
  const double pitch = (double)g->fov_v / nv, vz = (double)g->lz / g->nz;
  double worst = 0.0;
  for (int row0 = 0; row0 < nv; row0 += 64 * rpl) {
    const int row1 = std::min(nv, row0 + 64 * rpl) - 1;
    const double v0 = ((row0 + 0.5) / nv - 0.5) * g->fov_v + omin - 0.5 * pitch;
    const double v1 = ((row1 + 0.5) / nv - 0.5) * g->fov_v + omax + 0.5 * pitch;
    const double e[4] = {v0 * t0, v0 * t1, v1 * t0, v1 * t1};
    const double lo = std::min(std::min(e[0], e[1]), std::min(e[2], e[3]));
    const double hi = std::max(std::max(e[0], e[1]), std::max(e[2], e[3]));
    worst = std::max(worst, (hi - lo) / g->sdd / vz);
  }
  return worst + 4.0;
}

FPChoice default_variant(cbct_ctx ctx, const cbct_geom *g, const FPParams &P, int view0, int nviews,
                         int per_step) {
  float omin = 0.0f, omax = 0.0f;
  for (int k = 0; k < nviews; ++k) {
    const float ov = ctx->offsets[2 * (size_t)(view0 + k) + 1];
    omin = k ? fminf(omin, ov) : ov;
    omax = k ? fmaxf(omax, ov) : ov;
  }
  for (int r = ctx->fp_max_rpl; r >= 2; --r) {
    const int rpl = rows_per_lane(P.nv, r);
    if (rpl >= 2 && window_bound(g, P.nv, rpl, omin, omax) <= 256.0) return {3, rpl};
  }
  return {per_step, rows_per_lane(P.nv, ctx->fp_max_rpl)};
}

// Autotuned choices, process-wide (cbct_fp_autotune): the geometry, the
// samples per ray and the working-copy element size pick (variant, tile).
struct TuneKey {
  int nx, ny, nz, nu, nv, esz;
  float lx, ly, lz, fu, fv, sad, sdd, spp;
  bool operator<(const TuneKey &o) const {
    return std::tie(nx, ny, nz, nu, nv, esz, lx, ly, lz, fu, fv, sad, sdd, spp) <
           std::tie(o.nx, o.ny, o.nz, o.nu, o.nv, o.esz, o.lx, o.ly, o.lz, o.fu, o.fv, o.sad, o.sdd,
                    o.spp);
  }
};
struct Tuned {
  FPChoice choice;
  double ms;
};
std::mutex g_tune_mu;
std::map<TuneKey, Tuned> g_tuned;

TuneKey tune_key(const cbct_geom *g, float spp, size_t esz) {
  return {g->nx, g->ny, g->nz, g->nu, g->nv, (int)esz, g->lx, g->ly, g->lz, g->fov_u, g->fov_v,
          g->sad, g->sdd, spp};
}

// The variant cbct_fp runs and its tile: variant 0 = per ray (engine 0, or a
// volume or z-line beyond the column engine's 32-bit byte offsets / 64 KB of
// LDS); engine 1: default_variant; engine 2: the other per-step variant;
// engines 3 and 4: the cached variants (configured tile).
FPChoice select_variant(cbct_ctx ctx, const cbct_geom *g, const FPParams &P, int view0, int nviews,
                        size_t esz, float spp) {
  const size_t vt_bytes = (size_t)P.pzx * padded_dim(g->ny) * esz;
  const size_t zline_bytes = ((P.pz + 127) / 128) * 128 * sizeof(float);
  const int rpl = rows_per_lane(P.nv, ctx->fp_max_rpl);
  if (ctx->fp_engine == 0 || vt_bytes + (size_t)P.pzx * esz * 2 >= (1ull << 31) || zline_bytes > 65536)
    return {0, rpl};
  if (ctx->fp_force_variant > 0) return {ctx->fp_force_variant, ctx->fp_force_rpl, ctx->fp_force_vg};
  const int per_step = esz == 2 ? 2 : 1;
  if (ctx->fp_engine == 2) return {3 - per_step, rpl};
  if (ctx->fp_engine >= 3) return {ctx->fp_engine, rpl};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(g, spp, esz));
    if (it != g_tuned.end()) return it->second.choice;
  }
  return default_variant(ctx, g, P, view0, nviews, per_step);
}

// Bytes of the padded working copy cbct_fp builds for this call (the
// variant's layout, element size esz).
size_t fp_working_bytes(cbct_ctx ctx, const cbct_geom *g, const FPParams &P, int view0, int nviews,
                        size_t esz, float spp) {
  const FPChoice choice = select_variant(ctx, g, P, view0, nviews, esz, spp);
  if (choice.variant == 0) return (size_t)P.pxy * padded_dim(g->nz) * esz;
  return ((size_t)P.pzx * padded_dim(g->ny) + 1024) * esz;
}

// TI: the caller's volume type; T: the padded working copy's type.  The
// working copy goes to the context's workspace, or to `ext` when given (the
// caller has checked it holds fp_working_bytes).
template <typename TI, typename T>
int run_fp(cbct_ctx ctx, const cbct_geom *g, const FPParams &P, int view0, int nviews,
           const TI *vol, float *proj, int accumulate, hipStream_t s, float spp,
           void *ext = nullptr) {
  const size_t esz = sizeof(T);
  const ViewD *views = ctx->d_views_d + view0;
  const size_t vt_bytes = (size_t)P.pzx * padded_dim(g->ny) * esz;
  const FPChoice choice = select_variant(ctx, g, P, view0, nviews, esz, spp);
  if (choice.variant == 0) {
    const size_t npad = (size_t)P.pxy * padded_dim(g->nz);
    if (!ext) {
      int rc = cbct_reserve(ctx, ctx->ws->vol_pad, npad * esz);
      if (rc) return rc;
    }
    T *vp = static_cast<T *>(ext ? ext : ctx->ws->vol_pad.ptr);
    hipLaunchKernelGGL((k_pad_volume<TI, T>), stride_grid(npad), dim3(256), 0, s, vol, vp, g->nx, g->ny,
                       g->nz, P.px, (unsigned)padded_dim(g->ny), (unsigned)padded_dim(g->nz));
    CBCT_LAUNCH_CHECK(ctx, "k_pad_volume");
    dim3 block(16, 16);
    dim3 grid(ceil_div(g->nu, 16), ceil_div(g->nv, 16), nviews);
    cbct_prof_mark(ctx, PROF_FP, 0, s);
    hipLaunchKernelGGL(k_fp_ray<T>, grid, block, 0, s, P, views, vp, proj, accumulate);
    cbct_prof_mark(ctx, PROF_FP, 1, s);
    CBCT_LAUNCH_CHECK(ctx, "k_fp_ray");
    return CBCT_OK;
  }
  const size_t py = padded_dim(g->ny);
  const size_t npad = (size_t)P.pzx * py;
  // z-line chunk loads may run up to 127 elements past a column's end (those
  // values are never used); the slack keeps such reads inside this allocation
  if (!ext) {
    int rc = cbct_reserve(ctx, ctx->ws->vol_pad, (npad + 1024) * esz);
    if (rc) return rc;
  }
  T *vt = static_cast<T *>(ext ? ext : ctx->ws->vol_pad.ptr);
  dim3 tgrid(ceil_div(P.pz, 32), ceil_div(P.px, 32), (unsigned)py);
  hipLaunchKernelGGL((k_pad_transpose<TI, T>), tgrid, dim3(32, 8), 0, s, vol, vt, g->nx, g->ny, g->nz,
                     P.px, P.pz);
  CBCT_LAUNCH_CHECK(ctx, "k_pad_transpose");
  FPParams Q = P;
  if (choice.vg > 0) Q.vg = (unsigned)choice.vg;
  cbct_prof_mark(ctx, PROF_FP, 0, s);
  dispatch_col<T>(Q, choice.variant, choice.rpl, s, views, vt, proj, accumulate, nviews,
                  (unsigned)vt_bytes);
  cbct_prof_mark(ctx, PROF_FP, 1, s);
  CBCT_LAUNCH_CHECK(ctx, "k_fp_col");
  return CBCT_OK;
}

}  // namespace

extern "C" {

int cbct_fp_engine(cbct_ctx ctx, int engine) {
  if (!ctx) return CBCT_ERR_ARG;
  if (engine < 0 || engine > 4)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "cbct_fp_engine: engine %d not available", engine);
  ctx->fp_engine = engine;
  return CBCT_OK;
}

int cbct_fp_autotune(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, int dtype, float spp,
                     int retune, void *stream, double *best_ms) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "cbct_fp_autotune";
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (nviews <= 0) return cbct_fail(ctx, CBCT_ERR_ARG, "%s: no views", who);
  if (dtype != CBCT_F32 && dtype != CBCT_BF16)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "%s: unsupported volume dtype %d", who, dtype);
  FPParams P;
  rc = make_params(ctx, g, spp, P);
  if (rc) return rc;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  const size_t esz = (dtype == CBCT_BF16 || ctx->work_dtype == CBCT_BF16) ? 2 : 4;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(g, spp, esz));
    if (it != g_tuned.end()) {
      if (!retune) {
        if (best_ms) *best_ms = it->second.ms;
        return CBCT_OK;
      }
      g_tuned.erase(it);   // time the built-in choice itself
    }
  }
  const int saved_engine = ctx->fp_engine;
  ctx->fp_engine = 1;
  const FPChoice heuristic = select_variant(ctx, g, P, view0, nviews, esz, spp);
  ctx->fp_engine = saved_engine;
  if (heuristic.variant == 0) return CBCT_OK;   // the column engines do not apply
  // candidates: the heuristic's; the cached variants with 2- and 3-chunk
  // windows at several tiles; the per-step variant of the storage type
  std::vector<FPChoice> cand{heuristic};
  auto add = [&](int v, int r) {
    for (auto &c : cand)
      if (c.variant == v && c.rpl == r) return;
    cand.push_back({v, r});
  };
  for (int m : {4, 6, 8, 12}) add(3, rows_per_lane(P.nv, m));
  for (int m : {6, 8}) add(4, rows_per_lane(P.nv, m));
  add(esz == 2 ? 2 : 1, rows_per_lane(P.nv, 8));
  // scratch volume and projections (contents do not change the timing)
  const size_t nvox = (size_t)g->nx * g->ny * g->nz;
  const size_t nproj = (size_t)g->nu * g->nv * (size_t)nviews;
  void *vol = nullptr, *proj = nullptr;
  CBCT_HIP_CHECK(ctx, hipMalloc(&vol, nvox * (dtype == CBCT_BF16 ? 2 : 4)));
  if (hipMalloc(&proj, nproj * sizeof(float)) != hipSuccess) {
    (void)hipFree(vol);
    (void)hipGetLastError();
    return cbct_fail(ctx, CBCT_ERR_NOMEM, "%s: scratch allocation failed", who);
  }
  hipStream_t s = as_stream(stream);
  CallScope scope(ctx, s);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const size_t vol_words = nvox * (dtype == CBCT_BF16 ? 2 : 4) / 4;   // whole 4-byte words
  hipLaunchKernelGGL(k_zero, stride_grid(vol_words), dim3(256), 0, s, vol_words, static_cast<float *>(vol));
  FPChoice best = heuristic;
  double best_t = 1e300;
  auto trial = [&](const FPChoice &c) {
    ctx->fp_force_variant = c.variant;
    ctx->fp_force_rpl = c.rpl;
    ctx->fp_force_vg = c.vg;
    double t = 1e300;
    for (int rep = 0; rep < 3 && rc == CBCT_OK; ++rep) {   // first run warms up
      (void)hipEventRecord(e0, s);
      rc = cbct_fp(ctx, g, view0, nviews, vol, dtype, static_cast<float *>(proj), spp, 0, stream);
      (void)hipEventRecord(e1, s);
      if (rc == CBCT_OK && hipEventSynchronize(e1) == hipSuccess && rep > 0) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) t = std::min(t, (double)ms);
      }
    }
    if (t < best_t) {
      best_t = t;
      best = c;
    }
  };
  // the variants and tiles at the context's launch order, then the launch
  // order (views per group) for the fastest of them
  for (size_t c = 0; c < cand.size() && rc == CBCT_OK; ++c) trial(cand[c]);
  const FPChoice base = best;
  for (int vg : {1, 3, 6, 8})
    if (rc == CBCT_OK && vg != ctx->fp_view_group) trial({base.variant, base.rpl, vg});
  ctx->fp_force_variant = 0;
  ctx->fp_force_rpl = 0;
  ctx->fp_force_vg = 0;
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(vol);
  (void)hipFree(proj);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[tune_key(g, spp, esz)] = {best, best_t};
  }
  if (best_ms) *best_ms = best_t;
  return CBCT_OK;
}

int cbct_fp_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
  return CBCT_OK;
}

int cbct_fp_tile(cbct_ctx ctx, int max_rows_per_lane) {
  if (!ctx) return CBCT_ERR_ARG;
  if (max_rows_per_lane < 1 || max_rows_per_lane > 12)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_fp_tile: rows per lane %d outside 1..12", max_rows_per_lane);
  ctx->fp_max_rpl = max_rows_per_lane;
  return CBCT_OK;
}

int cbct_fp_view_group(cbct_ctx ctx, int views) {
  if (!ctx) return CBCT_ERR_ARG;
  if (views < 1 || views > 64)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_fp_view_group: %d views outside 1..64", views);
  ctx->fp_view_group = views;
  return CBCT_OK;
}

const char *cbct_fp_kernel(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, int dtype,
                           float spp) {
  if (!ctx || !g || cbct_check_geom(ctx, g, "cbct_fp_kernel") ||
      cbct_check_views(ctx, view0, nviews, "cbct_fp_kernel"))
    return "";
  FPParams P;
  if (make_params(ctx, g, spp, P)) return "";
  const size_t esz = (dtype == CBCT_BF16 || ctx->work_dtype == CBCT_BF16) ? 2 : 4;
  const int v = select_variant(ctx, g, P, view0, nviews, esz, spp).variant;
  return v == 0 ? "k_fp_ray" : v >= 3 ? "k_fp_colc" : "k_fp_col";
}

}  // extern "C"

namespace {

// Argument checks and parameters shared by cbct_fp and cbct_fp_noalloc.
int fp_prepare(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *vol, int dtype,
               float *proj, float spp, FPParams &P) {
  const char *who = "conebeam_forwards_projection";
  if (!vol || !proj)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (dtype != CBCT_F32 && dtype != CBCT_BF16)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "Error: %s: unsupported volume dtype %d", who, dtype);
  if (nviews == 0) return CBCT_OK;   // nothing to do (the caller returns)
  if (nviews > 65535)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: at most 65535 views per call", who);
  if ((size_t)g->nu * (size_t)((g->nv + 63) / 64) * (size_t)nviews >= (1ull << 31))
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: too many rays per call", who);
  rc = make_params(ctx, g, spp, P);
  if (rc) return rc;
  return cbct_set_device(ctx);
}

template <typename TI, typename T>
int run_fp_as(cbct_ctx ctx, const cbct_geom *g, const FPParams &P, int view0, int nviews,
              const void *vol, float *proj, int accumulate, hipStream_t s, float spp, void *ext) {
  return run_fp<TI, T>(ctx, g, P, view0, nviews, static_cast<const TI *>(vol), proj, accumulate, s,
                       spp, ext);
}

int run_fp_dtype(cbct_ctx ctx, const cbct_geom *g, const FPParams &P, int view0, int nviews,
                 const void *vol, int dtype, float *proj, int accumulate, hipStream_t s, float spp,
                 void *ext) {
  if (dtype == CBCT_BF16)
    return run_fp_as<unsigned short, unsigned short>(ctx, g, P, view0, nviews, vol, proj, accumulate,
                                                     s, spp, ext);
  if (ctx->work_dtype == CBCT_BF16)
    return run_fp_as<float, unsigned short>(ctx, g, P, view0, nviews, vol, proj, accumulate, s, spp,
                                            ext);
  return run_fp_as<float, float>(ctx, g, P, view0, nviews, vol, proj, accumulate, s, spp, ext);
}

// Per-device address of `scratch`, looked up on a device's first
// cbct_fp_noalloc call.
std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};

int fp_scratch(cbct_ctx ctx, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (ctx->device < 0 || ctx->device >= 64)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "cbct_fp_noalloc: device %d out of range", ctx->device);
  if (!g_scratch_ptr[ctx->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return cbct_fail(ctx, CBCT_ERR_HIP, "cbct_fp_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = p;
  }
  *out = g_scratch_ptr[ctx->device];
  return CBCT_OK;
}

}  // namespace

extern "C" {

int cbct_fp(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *vol, int dtype,
            float *proj, float spp, int accumulate, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  FPParams P;
  int rc = fp_prepare(ctx, g, view0, nviews, vol, dtype, proj, spp, P);
  if (rc || nviews == 0) return rc;
  hipStream_t s = as_stream(stream);
  CallScope scope(ctx, s);
  return run_fp_dtype(ctx, g, P, view0, nviews, vol, dtype, proj, accumulate, s, spp, nullptr);
}

int cbct_fp_noalloc(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *vol,
                    int dtype, float *proj, float spp, int accumulate, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  FPParams P;
  int rc = fp_prepare(ctx, g, view0, nviews, vol, dtype, proj, spp, P);
  if (rc || nviews == 0) return rc;
  const size_t esz = (dtype == CBCT_BF16 || ctx->work_dtype == CBCT_BF16) ? 2 : 4;
  const size_t need = fp_working_bytes(ctx, g, P, view0, nviews, esz, spp);
  if (need > FP_SCRATCH_BYTES)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED,
                     "cbct_fp_noalloc: the volume's working copy needs %zu bytes, more than the %zu "
                     "bytes of scratch (use cbct_fp)", need, FP_SCRATCH_BYTES);
  void *ptr = nullptr;
  rc = fp_scratch(ctx, &ptr);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  rc = run_fp_dtype(ctx, g, P, view0, nviews, vol, dtype, proj, accumulate, s, spp, ptr);
  // this call read the view table: a later cbct_views_set waits for it
  if (rc == CBCT_OK && ctx->done && hipEventRecord(ctx->done, s) == hipSuccess) ctx->done_recorded = true;
  return rc;
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
  hipStream_t s = as_stream(stream);
  CallScope scope(ctx, s);
  const size_t npad = (size_t)P.pxy * padded_dim(g->nz);
  rc = cbct_reserve(ctx, ctx->ws->vol_pad, npad * sizeof(float));
  if (rc) return rc;
  float *acc = static_cast<float *>(ctx->ws->vol_pad.ptr);
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
