// Voxel-driven cone-beam backprojector (optionally FDK weighted) for gfx950.
//
// Math: xray/conebeam_projection.cu:922-1037.  For every voxel centre and
// view: rotate by -theta, project from the source onto the detector plane,
// subtract the view offset, take a bilinear sample (zero outside the
// detector) and accumulate, with the FDK weight 1/U^2 in FBP mode.
//
// Design
//  * Rotation about z leaves z unchanged, so for a fixed (x, y) the rotated
//    coordinate, the magnification SDD/(SAD+q_y), the detector u-coordinate,
//    its interpolation weight and the FDK weight are identical for every
//    voxel of the column.  Each thread owns KZ voxels of one column and
//    computes those once per view; per voxel-view only the v-coordinate (one
//    FMA), its weight, four gathers and three lerps remain.
//  * The FDK weight uses the identity r*sin(beta-phi) = x sin(beta) - y
//    cos(beta) = -q_y (no hypot/atan2/sin per voxel-view).
//  * Projections are read from a zero-bordered copy (1 pixel before, 2 after
//    in u and v) so no bounds tests are needed; the copy may be BF16.
//  * Per-view constants (cos, sin, offsets) are wave-uniform and arrive
//    through scalar loads.
#include "cbct_internal.h"

namespace {

constexpr int BP_TX = 16, BP_TY = 16;

struct BPParams {
  int nx, ny, nz, nviews;
  float lx, ly, lz;
  float sad, sdd, inv_sad;
  float su, sv;        // nu/fov_u, nv/fov_v
  float cu, cv;        // padded texel offset: n/2 - 0.5 + 1
  float clu, clv;      // clamp limits (= n + 1)
  unsigned pnu;        // padded row length
  unsigned pnv;        // padded rows
  unsigned pview;      // padded elements per view
  unsigned pnu2;       // pair copy: row length (pnu rounded up to 4)
  unsigned pview2;     // pair copy: dwords per view
  unsigned bpp_rowb;   // LDS row of the pair kernel in bytes (a parameter, so
                       // the compiler keeps iv * row + col one 24-bit mad)
};

template <typename T>
__device__ __forceinline__ float ldp(const T *p, unsigned i);
template <>
__device__ __forceinline__ float ldp<float>(const float *p, unsigned i) {
  return p[i];
}
template <>
__device__ __forceinline__ float ldp<unsigned short>(const unsigned short *p, unsigned i) {
  return bf16_to_f32(p[i]);
}

// Raw buffer load of one padded-projection texel (32-bit byte offsets within
// one view; the descriptor covers exactly that view, so out-of-range offsets
// read zero rather than another view).
template <typename T>
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bld<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  // the builtin returns the raw 32-bit word
  return __uint_as_float((unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ float bld<unsigned short>(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                                     unsigned soff) {
  return bf16_to_f32(__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0));
}

// The four bilinear taps (u, v), (u+1, v), (u, v+1), (u+1, v+1) of a quad
// texel (see k_pad_quad) from one load: 16 bytes for FP32, 8 for BF16.
template <typename T>
__device__ __forceinline__ float4 qld(__amdgpu_buffer_rsrc_t r, unsigned voff);
template <>
__device__ __forceinline__ float4 qld<float>(__amdgpu_buffer_rsrc_t r, unsigned voff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0);
  return make_float4(__uint_as_float((unsigned)v[0]), __uint_as_float((unsigned)v[1]),
                     __uint_as_float((unsigned)v[2]), __uint_as_float((unsigned)v[3]));
}
template <>
__device__ __forceinline__ float4 qld<unsigned short>(__amdgpu_buffer_rsrc_t r, unsigned voff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, 0);
  const unsigned w0 = (unsigned)v[0], w1 = (unsigned)v[1];
  return make_float4(__uint_as_float(w0 << 16), __uint_as_float(w0 & 0xffff0000u),
                     __uint_as_float(w1 << 16), __uint_as_float(w1 & 0xffff0000u));
}

// QUAD: projections in the quad layout (one load per voxel-view), else the
// plain zero-bordered layout (four taps, two loads after merging).
template <typename T, bool FBP, int KZ, bool QUAD>
__global__ void __launch_bounds__(BP_TX *BP_TY) k_bp(BPParams P, const ViewF *__restrict__ views,
                                                     const T *__restrict__ ppad,
                                                     float *__restrict__ vol, int accumulate,
                                                     unsigned gx, unsigned gy) {
  // only the explicit fmaf calls fuse: the kernels' bits must not depend on
  // how the compiler contracts what it simplifies (fmaf(1, a*b, c) = a*b + c)
#pragma clang fp contract(off)
  // linear tile id, x fastest, z slowest.  Consecutive tiles land on
  // different XCDs, so all XCDs work on one band of the volume -- and one
  // band of detector rows per view -- at a time (measured faster than giving
  // each XCD its own z-slab).
  const unsigned L = blockIdx.x;
  // a wave's 64 columns are an 8 x 8 quarter of the 16 x 16 tile: their
  // projections span fewer detector cache lines at every view angle than a
  // 16 x 4 strip (measured 10% faster for FP32 working copies, 2% for BF16)
  const int tid = threadIdx.x + BP_TX * threadIdx.y, wv = tid >> 6, ln = tid & 63;
  const int lx = (wv & 1) * 8 + (ln & 7);
  const int ly = (wv >> 1) * 8 + (ln >> 3);
  const int x = (int)(L % gx) * BP_TX + lx;
  const int y = (int)((L / gx) % gy) * BP_TY + ly;
  const int z0 = (int)(L / (gx * gy)) * KZ;
  if (x >= P.nx || y >= P.ny) return;
  // voxel centre in mm (xray/conebeam_projection.cu:945-959)
  const float px = ((x + 0.5f) / P.nx - 0.5f) * P.lx;
  const float py = ((y + 0.5f) / P.ny - 0.5f) * P.ly;
  const float dz = P.lz / P.nz;
  const float pz0 = ((z0 + 0.5f) / P.nz - 0.5f) * P.lz;
  float acc[KZ];
#pragma unroll
  for (int t = 0; t < KZ; ++t) acc[t] = 0.0f;
  const unsigned esz = QUAD ? 4 * sizeof(T) : sizeof(T);  // bytes per (quad) texel
  const unsigned row_bytes = P.pnu * esz;  // < 2^24: 24-bit multiplies suffice
  const int view_bytes = (int)(P.pview * esz);

  ViewF vw = views[0];
  for (int k = 0; k < P.nviews; ++k) {
    const ViewF nxt = views[k + 1 < P.nviews ? k + 1 : k];  // prefetch the next view
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(ppad + (size_t)k * P.pview * (QUAD ? 4 : 1)), (short)0, view_bytes, 0x00020000);
    const float qx = fmaf(vw.c, px, vw.s * py);
    const float qy = fmaf(vw.c, py, -vw.s * px);
    const float mag = P.sdd * __builtin_amdgcn_rcpf(qy + P.sad);
    // coordinates are finite here; med3 = clamp, truncation = floor
    const float tu = __builtin_amdgcn_fmed3f(fmaf(fmaf(qx, mag, -vw.off_u), P.su, P.cu), 0.0f, P.clu);
    const float au = __builtin_amdgcn_fractf(tu);
    const unsigned ub = (unsigned)(int)tu * esz;
    const float tv0 = fmaf(fmaf(pz0, mag, -vw.off_v), P.sv, P.cv);
    const float dtv = dz * mag * P.sv;
    float w = 1.0f;
    if (FBP) {
      const float U = (P.sad - qy) * P.inv_sad;
      w = __builtin_amdgcn_rcpf(U * U);
    }
#pragma unroll
    for (int t = 0; t < KZ; ++t) {
      const float tv = __builtin_amdgcn_fmed3f(fmaf((float)t, dtv, tv0), 0.0f, P.clv);
      const float av = __builtin_amdgcn_fractf(tv);
      const unsigned o = __umul24((unsigned)(int)tv, row_bytes) + ub;
      float a, b, c, d;
      if constexpr (QUAD) {
        const float4 q = qld<T>(rsrc, o);
        a = q.x; b = q.y; c = q.z; d = q.w;
      } else {
        a = bld<T>(rsrc, o, 0); b = bld<T>(rsrc, o + esz, 0);
        c = bld<T>(rsrc, o, row_bytes); d = bld<T>(rsrc, o + esz, row_bytes);
      }
      const float r0 = fmaf(au, b - a, a);
      const float r1 = fmaf(au, d - c, c);
      acc[t] = fmaf(w, fmaf(av, r1 - r0, r0), acc[t]);
    }
    vw = nxt;
  }
#pragma unroll
  for (int t = 0; t < KZ; ++t) {
    const int z = z0 + t;
    if (z < P.nz) {
      const size_t idx = (size_t)x + (size_t)P.nx * ((size_t)y + (size_t)P.ny * z);
      vol[idx] = accumulate ? vol[idx] + acc[t] : acc[t];
    }
  }
}

// LDS-staged backprojector.  Per view, each block's detector rectangle
// (k_bpp_rects) is copied into LDS by 16-byte direct global-to-LDS loads, one
// view ahead, and every voxel takes its taps from there.  Two layouts, both
// 4-byte elements in rows of pnu2 (16-byte aligned): PAIRS (BF16 working
// copies, k_pad_staged): element (a, b) holds the taps (a, b) and (a+1, b), so
// the four taps of a voxel-view are two elements one LDS row apart, one
// ds_read2_b32; plain FP32 (k_pad_staged): the four taps are two adjacent
// pairs of elements one row apart, two ds_read2_b32.  LDS rows are BPP_RW =
// 68 elements (adjacent rows 4 banks apart); a buffer holds BPP_CAP
// elements, a multiple of one 64-lane copy (256), so whole copies never write
// past it.  Views whose rectangle does not fit read the copy directly.  Same
// arithmetic, same bits as k_bp.
constexpr int BPP_RW = 68;
constexpr int BPP_CAP = 13 * 256;   // 48 rows of 68 elements, rounded up to whole copies
constexpr int BPP_CHUNKS = BPP_CAP / 256;

// The rectangle of block L (x fastest, z slowest, as k_bpp) at view k, as
// entry L * nviews + k: (u0 | v0 << 16, H | inner << 8), 0 when it does not
// fit.
// Its bounds: the projections of the block's four corner columns at its
// first and last z bound every voxel of the block (u is a linear-fractional
// function of (x, y), v is linear in the magnification, so both are extreme
// at the corners), with one texel of margin for rounding (columns up to the
// right tap's); u0 is rounded down to 4 elements so the copies are 16-byte
// aligned.
template <int KZ>
__global__ void __launch_bounds__(256) k_bpp_rects(BPParams P, const ViewF *__restrict__ views,
                                                   uint2 *__restrict__ rects, unsigned gx, unsigned gy,
                                                   unsigned nblocks) {
#pragma clang fp contract(off)
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)nblocks * P.nviews) return;
  const unsigned L = (unsigned)(t / P.nviews), k = (unsigned)(t % P.nviews);
  const int xb = (int)(L % gx) * BP_TX, yb = (int)((L / gx) % gy) * BP_TY;
  const int z0 = (int)(L / (gx * gy)) * KZ;
  const float dz = P.lz / P.nz;
  const float pz0 = ((z0 + 0.5f) / P.nz - 0.5f) * P.lz;
  const ViewF vw = views[k];
  int ulo = 1 << 30, uhi = -(1 << 30), vlo = 1 << 30, vhi = -(1 << 30);
  float rlo = 3.0e38f, rhi = -3.0e38f;   // unclamped row coordinates
  for (int c = 0; c < 4; ++c) {
    const int cxi = xb + ((c & 1) ? BP_TX - 1 : 0), cyi = yb + ((c & 2) ? BP_TY - 1 : 0);
    const float cpx = ((cxi + 0.5f) / P.nx - 0.5f) * P.lx;
    const float cpy = ((cyi + 0.5f) / P.ny - 0.5f) * P.ly;
    const float cqx = fmaf(vw.c, cpx, vw.s * cpy);
    const float cqy = fmaf(vw.c, cpy, -vw.s * cpx);
    const float cmag = P.sdd * __builtin_amdgcn_rcpf(cqy + P.sad);
    const float ctu = __builtin_amdgcn_fmed3f(fmaf(fmaf(cqx, cmag, -vw.off_u), P.su, P.cu), 0.0f, P.clu);
    const float ctv0 = fmaf(fmaf(pz0, cmag, -vw.off_v), P.sv, P.cv);
    const float ctv1 = fmaf((float)(KZ - 1), dz * cmag * P.sv, ctv0);
    rlo = fminf(rlo, fminf(ctv0, ctv1));
    rhi = fmaxf(rhi, fmaxf(ctv0, ctv1));
    const float ca = __builtin_amdgcn_fmed3f(ctv0, 0.0f, P.clv);
    const float cb = __builtin_amdgcn_fmed3f(ctv1, 0.0f, P.clv);
    ulo = min(ulo, (int)ctu);
    uhi = max(uhi, (int)ctu);
    vlo = min(vlo, (int)fminf(ca, cb));
    vhi = max(vhi, (int)fmaxf(ca, cb));
  }
  const int u0 = max(ulo - 1, 0) & ~3, v0 = max(vlo - 1, 0);
  const int W = min(uhi + 2, (int)P.pnu2 - 1) - u0 + 1;
  const int H = min(vhi + 2, (int)P.pnv - 1) - v0 + 1;
  const bool staged = W <= BPP_RW && H * BPP_RW <= BPP_CAP;
  // bit 8: every voxel's row coordinate lies inside [0, nv + 1] (half a
  // texel of margin), so its clamp is the identity and is skipped
  const unsigned inner = rlo >= 0.5f && rhi <= P.clv - 0.5f ? 0x100u : 0u;
  rects[t] = make_uint2((unsigned)u0 | ((unsigned)v0 << 16), staged ? (unsigned)H | inner : 0u);
}

template <bool PAIRS, bool FBP, int KZ>
__global__ void __launch_bounds__(BP_TX *BP_TY) k_bpp(BPParams P, const ViewF *__restrict__ views,
                                                      const unsigned *__restrict__ copy,
                                                      const uint2 *__restrict__ rects,
                                                      float *__restrict__ vol, int accumulate,
                                                      unsigned gx, unsigned gy) {
#pragma clang fp contract(off)
  __shared__ unsigned buf[2 * BPP_CAP];
  // linear tile id, x fastest, z slowest (consecutive tiles land on
  // different XCDs, so all XCDs work on one band of detector rows per view)
  const unsigned L = blockIdx.x;
  const int tid = threadIdx.x + BP_TX * threadIdx.y, ln = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar branches
  // a wave's 64 columns are an 8 x 8 quarter of the 16 x 16 tile
  const int x = (int)(L % gx) * BP_TX + (wv & 1) * 8 + (ln & 7);
  const int y = (int)((L / gx) % gy) * BP_TY + (wv >> 1) * 8 + (ln >> 3);
  const int z0 = (int)(L / (gx * gy)) * KZ;
  const bool active = x < P.nx && y < P.ny;   // the others stage and synchronise only
  const float px = ((x + 0.5f) / P.nx - 0.5f) * P.lx;
  const float py = ((y + 0.5f) / P.ny - 0.5f) * P.ly;
  const float dz = P.lz / P.nz;
  const float pz0 = ((z0 + 0.5f) / P.nz - 0.5f) * P.lz;
  const unsigned view_bytes = P.pview2 * 4u;
  const unsigned row_bytes = P.pnu2 * 4u;
  const uint2 *rb = rects + (size_t)L * P.nviews;
  // byte offset, within a rectangle, of the 4 elements each lane copies for
  // this wave's copies (copy i = wv + 4 j covers LDS elements [256 i, 256 i +
  // 256)); rows past the rectangle land in the buffer's unused tail
  unsigned loff[(BPP_CHUNKS + 3) / 4];
#pragma unroll
  for (int j = 0; j < (BPP_CHUNKS + 3) / 4; ++j) {
    const int p = (wv + 4 * j) * 256 + ln * 4;
    const int r = p / BPP_RW;
    loff[j] = (unsigned)(r * (int)P.pnu2 + (p - r * BPP_RW)) * 4u;
  }
  auto issue = [&](int k, uint2 R, int base) {
    // the rectangle's origin goes into the descriptor, so the hardware range
    // check (offsets past the view read zero) covers the whole copy
    const unsigned org = ((R.x >> 16) * P.pnu2 + (R.x & 0xffffu)) * 4u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)((const char *)(copy + (size_t)k * P.pview2) + org), (short)0, (int)(view_bytes - org),
        0x00020000);
    const int n = ((int)(R.y & 0xffu) * BPP_RW + 255) >> 8;
#pragma unroll
    for (int j = 0; j < (BPP_CHUNKS + 3) / 4; ++j)
      if (wv + 4 * j < n)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void *)&buf[base + (wv + 4 * j) * 256], 16, loff[j], 0,
            0, 0);
  };
  float acc[KZ];
#pragma unroll
  for (int t = 0; t < KZ; ++t) acc[t] = 0.0f;
  uint2 cur = rb[0];
  uint2 nxt = rb[P.nviews > 1 ? 1 : 0];
  if (cur.y) issue(0, cur, 0);
  ViewF vw = views[0];
  for (int k = 0; k < P.nviews; ++k) {
    // scalar loads one iteration ahead of their use
    const uint2 nn = rb[k + 2 < P.nviews ? k + 2 : P.nviews - 1];
    const ViewF vnext = views[k + 1 < P.nviews ? k + 1 : k];
    const float qx = fmaf(vw.c, px, vw.s * py);
    const float qy = fmaf(vw.c, py, -vw.s * px);
    const float mag = P.sdd * __builtin_amdgcn_rcpf(qy + P.sad);
    const float tu = __builtin_amdgcn_fmed3f(fmaf(fmaf(qx, mag, -vw.off_u), P.su, P.cu), 0.0f, P.clu);
    const float au = __builtin_amdgcn_fractf(tu);
    const int iu = (int)tu;
    const float tv0 = fmaf(fmaf(pz0, mag, -vw.off_v), P.sv, P.cv);
    const float dtv = dz * mag * P.sv;
    float w = 1.0f;
    if (FBP) {
      const float U = (P.sad - qy) * P.inv_sad;
      w = __builtin_amdgcn_rcpf(U * U);
    }
    // this wave's copies of view k have landed; after the barrier everyone's
    // have, and nobody reads view k-1's buffer any more
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (k + 1 < P.nviews && nxt.y) issue(k + 1, nxt, ((k + 1) & 1) * BPP_CAP);
    auto tap = [&](int t, float a, float b, float c, float d, float av) {
      const float r0 = fmaf(au, b - a, a);
      const float r1 = fmaf(au, d - c, c);
      acc[t] = fmaf(w, fmaf(av, r1 - r0, r0), acc[t]);
    };
    // the taps (iv, iu), (iv, iu+1), (iv+1, iu), (iv+1, iu+1) from elements
    // at byte offsets o (row iv) and o + row (row iv+1)
    auto taps = [&](int t, float av, auto ld, unsigned o, unsigned row) {
      if constexpr (PAIRS) {
        const unsigned w0 = ld(o), w1 = ld(o + row);
        tap(t, __uint_as_float(w0 << 16), __uint_as_float(w0 & 0xffff0000u), __uint_as_float(w1 << 16),
            __uint_as_float(w1 & 0xffff0000u), av);
      } else {
        tap(t, __uint_as_float(ld(o)), __uint_as_float(ld(o + 4)), __uint_as_float(ld(o + row)),
            __uint_as_float(ld(o + row + 4)), av);
      }
    };
    if (cur.y) {
      // LDS byte address of element (iv, iu): iv * row + iu * 4 + s, with s
      // uniform (readfirstlane keeps the compiler from splitting the sum)
      const unsigned s = __builtin_amdgcn_readfirstlane(
          (unsigned)(((k & 1) * BPP_CAP - (int)(cur.x & 0xffffu) - (int)(cur.x >> 16) * BPP_RW) * 4));
      const unsigned colb = (unsigned)iu * 4u + s;
      const char *lds = (const char *)buf;
      auto ld = [&](unsigned o) { return *(const unsigned *)(lds + o); };
      if (cur.y & 0x100u) {
        // the row clamp is the identity for the whole block (same values)
#pragma unroll
        for (int t = 0; t < KZ; ++t) {
          const float tv = fmaf((float)t, dtv, tv0);
          taps(t, __builtin_amdgcn_fractf(tv), ld, __umul24((unsigned)(int)tv, P.bpp_rowb) + colb,
               BPP_RW * 4u);
        }
      } else {
#pragma unroll
        for (int t = 0; t < KZ; ++t) {
          const float tv = __builtin_amdgcn_fmed3f(fmaf((float)t, dtv, tv0), 0.0f, P.clv);
          taps(t, __builtin_amdgcn_fractf(tv), ld, __umul24((unsigned)(int)tv, P.bpp_rowb) + colb,
               BPP_RW * 4u);
        }
      }
    } else {
      const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(copy + (size_t)k * P.pview2), (short)0, (int)view_bytes, 0x00020000);
      auto ld = [&](unsigned o) { return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)o, 0, 0); };
      const unsigned ub = (unsigned)iu * 4u;
#pragma unroll
      for (int t = 0; t < KZ; ++t) {
        const float tv = __builtin_amdgcn_fmed3f(fmaf((float)t, dtv, tv0), 0.0f, P.clv);
        taps(t, __builtin_amdgcn_fractf(tv), ld, __umul24((unsigned)(int)tv, row_bytes) + ub, row_bytes);
      }
    }
    cur = nxt;
    nxt = nn;
    vw = vnext;
  }
  if (!active) return;
#pragma unroll
  for (int t = 0; t < KZ; ++t) {
    const int z = z0 + t;
    if (z < P.nz) {
      const size_t idx = (size_t)x + (size_t)P.nx * ((size_t)y + (size_t)P.ny * z);
      vol[idx] = accumulate ? vol[idx] + acc[t] : acc[t];
    }
  }
}

// Zero-bordered copy of the projections: (nu+3) x (nv+3) per view.
template <typename TI, typename TO>
__global__ void __launch_bounds__(256) k_pad_proj(const TI *__restrict__ in, TO *__restrict__ out,
                                                  int nu, int nv, unsigned pnu, unsigned pnv,
                                                  size_t total);
template <>
__global__ void __launch_bounds__(256) k_pad_proj<float, float>(const float *__restrict__ in,
                                                                float *__restrict__ out, int nu,
                                                                int nv, unsigned pnu, unsigned pnv,
                                                                size_t total) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const unsigned a = (unsigned)(t % pnu);
    const unsigned b = (unsigned)((t / pnu) % pnv);
    const size_t k = t / ((size_t)pnu * pnv);
    const int u = (int)a - 1, v = (int)b - 1;
    float val = 0.0f;
    if (u >= 0 && u < nu && v >= 0 && v < nv) val = in[(size_t)u + (size_t)nu * ((size_t)v + (size_t)nv * k)];
    out[t] = val;
  }
}
template <>
__global__ void __launch_bounds__(256) k_pad_proj<unsigned short, unsigned short>(
    const unsigned short *__restrict__ in, unsigned short *__restrict__ out, int nu, int nv,
    unsigned pnu, unsigned pnv, size_t total) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const unsigned a = (unsigned)(t % pnu);
    const unsigned b = (unsigned)((t / pnu) % pnv);
    const size_t k = t / ((size_t)pnu * pnv);
    const int u = (int)a - 1, v = (int)b - 1;
    unsigned short val = 0;
    if (u >= 0 && u < nu && v >= 0 && v < nv) val = in[(size_t)u + (size_t)nu * ((size_t)v + (size_t)nv * k)];
    out[t] = val;
  }
}

// Quad layout: texel (a, b) of the zero-bordered projection P (P(a, b) =
// proj[a-1, b-1], 0 outside) holds P(a,b), P(a+1,b), P(a,b+1), P(a+1,b+1),
// for a < nu+3, b < nv+3.  One thread per quad texel.
template <typename TO, typename TI>
__device__ __forceinline__ TO to_storage(TI v) {
  return v;
}
template <>
__device__ __forceinline__ unsigned short to_storage<unsigned short, float>(float v) {
  return f32_to_bf16(v);
}

// The four taps of quad texel (a, b) as one store: 16 bytes (FP32), 8 (BF16).
template <typename T>
__device__ __forceinline__ void store_quad(T *o, const T (&q)[4]);
template <>
__device__ __forceinline__ void store_quad<float>(float *o, const float (&q)[4]) {
  *reinterpret_cast<float4 *>(o) = make_float4(q[0], q[1], q[2], q[3]);
}
template <>
__device__ __forceinline__ void store_quad<unsigned short>(unsigned short *o,
                                                           const unsigned short (&q)[4]) {
  *reinterpret_cast<uint2 *>(o) = make_uint2((unsigned)q[0] | ((unsigned)q[1] << 16),
                                             (unsigned)q[2] | ((unsigned)q[3] << 16));
}

// Padded copies: each wave covers 64 texel columns a (blockIdx.x) of one row
// b, strided over rows by the block's four waves and gridDim.y, and over
// views by gridDim.z, so no thread divides and each does many texels.
// TI: the caller's projection type; T: the copy's (FP32 -> BF16 converts).
template <typename TI, typename T>
__global__ void __launch_bounds__(256) k_pad_quad(const TI *__restrict__ in, T *__restrict__ out,
                                                  int nu, int nv, unsigned pnu, unsigned pnv,
                                                  int nviews) {
  const int a = (int)(blockIdx.x * 64 + (threadIdx.x & 63));
  if (a >= (int)pnu) return;
  const int b0 = (int)(blockIdx.y * 4 + (threadIdx.x >> 6)), bs = (int)(gridDim.y * 4);
  for (int k = (int)blockIdx.z; k < nviews; k += (int)gridDim.z) {
    const TI *pk = in + (size_t)nu * nv * k;
    T *ok = out + 4 * (size_t)pnu * pnv * k;
    for (int b = b0; b < (int)pnv; b += bs) {
      T q[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int u = a - 1 + (e & 1), v = b - 1 + (e >> 1);
        q[e] = (u >= 0 && u < nu && v >= 0 && v < nv) ? to_storage<T>(pk[(size_t)u + (size_t)nu * v])
                                                       : T(0);
      }
      store_quad<T>(ok + 4 * ((size_t)pnu * b + a), q);
    }
  }
}

// Launch grid of the padded copies (see k_pad_quad).
inline dim3 pad_grid(unsigned pnu, int nviews) {
  return dim3((pnu + 63) / 64, 16, (unsigned)std::min(nviews, 1024));
}

// The staged backprojector's copies, rows of pnu2 (a multiple of 4) 4-byte
// elements; element (a, b) of a view is, with u = a-1, v = b-1 and zero
// outside the projection: PAIRS, the BF16 taps (u, v) (low half) and
// (u+1, v); else the FP32 value at (u, v).  Each thread writes 4 elements
// (one 16-byte store) from the projection values at u = a0-1 .. a0+3: one
// 16-byte (FP32) or 8-byte (BF16) load plus the value before when the group
// lies inside the row and vec (rows aligned for it) is set, else value by
// value.  Grid as k_pad_quad over pnu2 / 4 columns.
template <typename TI, bool PAIRS>
__global__ void __launch_bounds__(256) k_pad_staged(const TI *__restrict__ in, unsigned *__restrict__ out,
                                                    int nu, int nv, unsigned pnu2, unsigned pnv,
                                                    int nviews, int vec) {
  const int a0 = 4 * (int)(blockIdx.x * 64 + (threadIdx.x & 63));
  if (a0 >= (int)pnu2) return;
  const int b0 = (int)(blockIdx.y * 4 + (threadIdx.x >> 6)), bs = (int)(gridDim.y * 4);
  const bool inner = vec && a0 >= 4 && a0 + 3 < nu;
  for (int k = (int)blockIdx.z; k < nviews; k += (int)gridDim.z) {
    const TI *pk = in + (size_t)nu * nv * k;
    unsigned *ok = out + (size_t)pnu2 * pnv * k;
    for (int b = b0; b < (int)pnv; b += bs) {
      const int v = b - 1;
      uint4 o = make_uint4(0u, 0u, 0u, 0u);
      if (v >= 0 && v < nv) {
        const TI *row = pk + (size_t)nu * v;
        TI e[5];
        if (inner) {
          e[0] = row[a0 - 1];
          if constexpr (sizeof(TI) == 4) {
            const float4 q = *reinterpret_cast<const float4 *>(row + a0);
            e[1] = q.x; e[2] = q.y; e[3] = q.z; e[4] = q.w;
          } else {
            const uint2 q = *reinterpret_cast<const uint2 *>(row + a0);
            e[1] = (TI)(q.x & 0xffffu); e[2] = (TI)(q.x >> 16);
            e[3] = (TI)(q.y & 0xffffu); e[4] = (TI)(q.y >> 16);
          }
        } else {
#pragma unroll
          for (int i = 0; i < 5; ++i) {
            const int u = a0 - 1 + i;
            e[i] = (u >= 0 && u < nu) ? row[u] : TI(0);
          }
        }
        if constexpr (PAIRS) {
          unsigned h[5];
#pragma unroll
          for (int i = 0; i < 5; ++i) h[i] = to_storage<unsigned short>(e[i]);
          o = make_uint4(h[0] | (h[1] << 16), h[1] | (h[2] << 16), h[2] | (h[3] << 16), h[3] | (h[4] << 16));
        } else {
          o = make_uint4(__float_as_uint(e[0]), __float_as_uint(e[1]), __float_as_uint(e[2]),
                         __float_as_uint(e[3]));
        }
      }
      *reinterpret_cast<uint4 *>(ok + (size_t)pnu2 * b + a0) = o;
    }
  }
}

// Space carver (xray/conebeam_projection.cu:822-921): nearest-texel lookups,
// z axis flipped, count of views with 0 != p < limit; mask = count < 3.
__global__ void __launch_bounds__(256) k_spacecarver(BPParams P, int nu, int nv,
                                                     const ViewF *__restrict__ views,
                                                     const float *__restrict__ proj,
                                                     unsigned char *__restrict__ mask, float limit) {
  const size_t nvox = (size_t)P.nx * P.ny * P.nz;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nvox;
       t += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(t % P.nx);
    const int y = (int)((t / P.nx) % P.ny);
    const int z = (int)(t / ((size_t)P.nx * P.ny));
    const float px = ((x + 0.5f) / P.nx - 0.5f) * P.lx;
    const float py = ((y + 0.5f) / P.ny - 0.5f) * P.ly;
    const float pz = -((z + 0.5f) / P.nz - 0.5f) * P.lz;
    int count = 0;
    for (int k = 0; k < P.nviews; ++k) {
      const ViewF vw = views[k];
      const float qx = fmaf(vw.c, px, vw.s * py);
      const float qy = fmaf(vw.c, py, -vw.s * px);
      const float mag = P.sdd / (qy + P.sad);
      // normalised detector coordinates, nearest texel = floor(t*N)
      const float nu_f = (qx * mag - vw.off_u) * (P.su / nu) + 0.5f;
      const float nv_f = (pz * mag - vw.off_v) * (P.sv / nv) + 0.5f;
      const float fu = floorf(nu_f * nu), fv = floorf(nv_f * nv);
      float val = 0.0f;
      if (fu >= 0.0f && fu < (float)nu && fv >= 0.0f && fv < (float)nv)
        val = proj[(size_t)fu + (size_t)nu * ((size_t)fv + (size_t)nv * k)];
      if (val < limit && val != 0.0f) ++count;
    }
    mask[t] = count < 3 ? 1 : 0;
  }
}

inline dim3 stride_grid(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  if (b == 0) b = 1;
  return dim3((unsigned)b);
}

BPParams make_params(const cbct_geom *g, int nviews) {
  BPParams P;
  P.nx = g->nx;
  P.ny = g->ny;
  P.nz = g->nz;
  P.nviews = nviews;
  P.lx = g->lx;
  P.ly = g->ly;
  P.lz = g->lz;
  P.sad = g->sad;
  P.sdd = g->sdd;
  P.inv_sad = 1.0f / g->sad;
  P.su = (float)g->nu / g->fov_u;
  P.sv = (float)g->nv / g->fov_v;
  P.cu = 0.5f * g->nu - 0.5f + 1.0f;
  P.cv = 0.5f * g->nv - 0.5f + 1.0f;
  P.clu = (float)(g->nu + 1);
  P.clv = (float)(g->nv + 1);
  P.pnu = (unsigned)padded_dim(g->nu);
  P.pnv = (unsigned)padded_dim(g->nv);
  P.pview = (unsigned)(padded_dim(g->nu) * padded_dim(g->nv));
  P.pnu2 = (P.pnu + 3u) & ~3u;
  P.pview2 = P.pnu2 * P.pnv;
  P.bpp_rowb = BPP_RW * 4;
  return P;
}

// z-voxels per thread: the per-(x, y, view) work (rotation, magnification,
// u coordinate and weight, FDK weight) is shared by KZ voxel-views (measured:
// 16 beats 8 by 4-5% and 32 by a third at 512^3).  8 when 16 would make
// fewer tiles than the chip holds blocks at once (1024 at four per CU): a
// 128^3 volume makes 512 tiles of 16 z.
constexpr int KZ = 16;
inline int bp_kz(const cbct_ctx ctx, const cbct_geom *g) {
  if (ctx->bp_kz) return ctx->bp_kz;
  const size_t tiles16 = (size_t)ceil_div(g->nx, BP_TX) * ceil_div(g->ny, BP_TY) * ceil_div(g->nz, 16);
  return tiles16 < 1024 ? 8 : KZ;
}

// Whether the LDS-staged kernel should run: at least half of the blocks'
// detector rectangles fit its LDS tile (staged block-views take about 0.7 of
// the quad kernel's time; with a quarter staged the two were measured equal
// at BF16), estimated on up to 8 evenly spaced views x a grid of up to 16 x
// 16 block columns x every z-band, with one texel of slack on each side.
// Cached per geometry, view range and z-depth until the view table changes.
bool bpp_fits(cbct_ctx ctx, const cbct_geom *g, const BPParams &P, int view0, int nviews, int kz) {
  std::vector<float> key = {(float)g->nx, (float)g->ny, (float)g->nz, g->lx, g->ly, g->lz,
                            (float)g->nu, (float)g->nv, g->fov_u, g->fov_v, g->sad, g->sdd,
                            (float)view0, (float)nviews, (float)kz};
  for (const auto &e : ctx->bpp_fit_cache)
    if (e.first == key) return e.second;
  const int gx = (int)ceil_div(g->nx, BP_TX), gy = (int)ceil_div(g->ny, BP_TY);
  const int gzb = (int)ceil_div(g->nz, kz);
  const float dz = P.lz / P.nz;
  const int vstep = std::max(1, nviews / 8), sx = std::max(1, gx / 16), sy = std::max(1, gy / 16);
  size_t total = 0, staged = 0;
  for (int k = view0; k < view0 + nviews; k += vstep) {
    const ViewF vw = ctx->views_h[k];
    for (int by = 0; by < gy; by += sy)
      for (int bx = 0; bx < gx; bx += sx) {
        float ulo = 1e30f, uhi = -1e30f, mlo = 1e30f, mhi = -1e30f;
        for (int c = 0; c < 4; ++c) {
          const int xi = bx * BP_TX + ((c & 1) ? BP_TX - 1 : 0);
          const int yi = by * BP_TY + ((c & 2) ? BP_TY - 1 : 0);
          const float px = ((xi + 0.5f) / P.nx - 0.5f) * P.lx;
          const float py = ((yi + 0.5f) / P.ny - 0.5f) * P.ly;
          const float qx = vw.c * px + vw.s * py, qy = vw.c * py - vw.s * px;
          const float mag = P.sdd / (qy + P.sad);
          const float tu = std::min(std::max((qx * mag - vw.off_u) * P.su + P.cu, 0.0f), P.clu);
          ulo = std::min(ulo, tu);
          uhi = std::max(uhi, tu);
          mlo = std::min(mlo, mag);
          mhi = std::max(mhi, mag);
        }
        total += gzb;
        const int u0 = std::max((int)std::floor(ulo) - 2, 0) & ~3;
        const int W = std::min((int)std::floor(uhi) + 3, (int)P.pnu2 - 1) - u0 + 1;
        if (W > BPP_RW) continue;
        for (int zb = 0; zb < gzb; ++zb) {
          // v is linear in pz * mag: extreme at the band's end slices and the
          // extreme magnifications
          const float pz0 = ((zb * kz + 0.5f) / P.nz - 0.5f) * P.lz, pz1 = pz0 + (kz - 1) * dz;
          const float e[4] = {pz0 * mlo, pz0 * mhi, pz1 * mlo, pz1 * mhi};
          const float plo = std::min(std::min(e[0], e[1]), std::min(e[2], e[3]));
          const float phi = std::max(std::max(e[0], e[1]), std::max(e[2], e[3]));
          const float vlo = std::min(std::max((plo - vw.off_v) * P.sv + P.cv, 0.0f), P.clv);
          const float vhi = std::min(std::max((phi - vw.off_v) * P.sv + P.cv, 0.0f), P.clv);
          const int v0 = std::max((int)std::floor(vlo) - 2, 0);
          const int H = std::min((int)std::floor(vhi) + 3, (int)P.pnv - 1) - v0 + 1;
          staged += H * BPP_RW <= BPP_CAP;
        }
      }
  }
  const bool fits = staged * 2 >= total;
  if (ctx->bpp_fit_cache.size() >= 256) ctx->bpp_fit_cache.clear();
  ctx->bpp_fit_cache.emplace_back(std::move(key), fits);
  return fits;
}

// Whether the call runs the LDS-staged kernel (engine 2 where it pays,
// engine 3 always; rectangle origins are packed in 16 bits).
bool use_staged(cbct_ctx ctx, const cbct_geom *g, const BPParams &P, int view0, int nviews, int kz) {
  return ctx->bp_engine >= 2 && P.pnu2 < 65536 && P.pnv < 65536 &&
         (ctx->bp_engine == 3 || bpp_fits(ctx, g, P, view0, nviews, kz));
}

template <bool PAIRS, int Z>
void launch_bpp(const BPParams &P, dim3 grid, hipStream_t s, const ViewF *views, const unsigned *cp,
                const uint2 *rects, float *vol, int fbp, int acc) {
  const unsigned nblocks = grid.x * grid.y * grid.z;
  if (fbp)
    hipLaunchKernelGGL((k_bpp<PAIRS, true, Z>), dim3(nblocks), dim3(BP_TX, BP_TY), 0, s, P, views, cp,
                       rects, vol, acc, grid.x, grid.y);
  else
    hipLaunchKernelGGL((k_bpp<PAIRS, false, Z>), dim3(nblocks), dim3(BP_TX, BP_TY), 0, s, P, views, cp,
                       rects, vol, acc, grid.x, grid.y);
}

template <typename T, bool QUAD, int Z>
void launch_bp_kz(const BPParams &P, dim3 grid, hipStream_t s, const ViewF *views, const T *pp,
                  float *vol, int fbp, int acc) {
  dim3 block(BP_TX, BP_TY);
  const unsigned nblocks = grid.x * grid.y * grid.z;
  if (fbp)
    hipLaunchKernelGGL((k_bp<T, true, Z, QUAD>), dim3(nblocks), block, 0, s, P, views, pp, vol, acc,
                       grid.x, grid.y);
  else
    hipLaunchKernelGGL((k_bp<T, false, Z, QUAD>), dim3(nblocks), block, 0, s, P, views, pp, vol,
                       acc, grid.x, grid.y);
}
template <typename T, bool QUAD>
void launch_bp(const BPParams &P, dim3 grid, int kz, hipStream_t s, const ViewF *views, const T *pp,
               float *vol, int fbp, int acc) {
  if (kz == 8) launch_bp_kz<T, QUAD, 8>(P, grid, s, views, pp, vol, fbp, acc);
  else launch_bp_kz<T, QUAD, KZ>(P, grid, s, views, pp, vol, fbp, acc);
}

// Padded copy in the selected layout (and storage type), then the kernel.
// TI: the caller's projection type; T: the copy's.
template <typename TI, typename T>
int run_bp(cbct_ctx ctx, const cbct_geom *g, const BPParams &P, dim3 grid, int kz, int view0,
           int nviews, const TI *proj, float *vol, int fbp, int accumulate, hipStream_t s) {
  // the LDS-staged kernel on a pair copy for BF16 working copies, a plain
  // one for FP32, after the pass that computes its rectangles
  if (use_staged(ctx, g, P, view0, nviews, kz)) {
    int rc2 = cbct_reserve(ctx, ctx->ws->proj_pad, (size_t)P.pview2 * nviews * 4);
    if (rc2) return rc2;
    const size_t nblocks = (size_t)grid.x * grid.y * grid.z;
    rc2 = cbct_reserve(ctx, ctx->ws->bp_rects, nblocks * nviews * sizeof(uint2));
    if (rc2) return rc2;
    unsigned *cp = static_cast<unsigned *>(ctx->ws->proj_pad.ptr);
    uint2 *rects = static_cast<uint2 *>(ctx->ws->bp_rects.ptr);
    const int vec = g->nu % 4 == 0 && reinterpret_cast<uintptr_t>(proj) % (4 * sizeof(TI)) == 0;
    if (sizeof(T) == 2)
      hipLaunchKernelGGL((k_pad_staged<TI, true>), pad_grid(P.pnu2 / 4, nviews), dim3(256), 0, s, proj, cp,
                         g->nu, g->nv, P.pnu2, P.pnv, nviews, vec);
    else
      hipLaunchKernelGGL((k_pad_staged<TI, false>), pad_grid(P.pnu2 / 4, nviews), dim3(256), 0, s, proj,
                         cp, g->nu, g->nv, P.pnu2, P.pnv, nviews, vec);
    CBCT_LAUNCH_CHECK(ctx, "k_pad_staged");
    const ViewF *vf = ctx->d_views_f + view0;
    const size_t nr = nblocks * nviews;
    if (kz == 8)
      hipLaunchKernelGGL((k_bpp_rects<8>), dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, P, vf, rects,
                         grid.x, grid.y, (unsigned)nblocks);
    else
      hipLaunchKernelGGL((k_bpp_rects<KZ>), dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, P, vf, rects,
                         grid.x, grid.y, (unsigned)nblocks);
    CBCT_LAUNCH_CHECK(ctx, "k_bpp_rects");
    cbct_prof_mark(ctx, PROF_BP, 0, s);
    if (kz == 8) launch_bpp<sizeof(T) == 2, 8>(P, grid, s, vf, cp, rects, vol, fbp, accumulate);
    else launch_bpp<sizeof(T) == 2, KZ>(P, grid, s, vf, cp, rects, vol, fbp, accumulate);
    cbct_prof_mark(ctx, PROF_BP, 1, s);
    CBCT_LAUNCH_CHECK(ctx, "k_bpp");
    return CBCT_OK;
  }
  // engine 0's plain layout converts nothing: BF16 working copies of FP32
  // projections use quads
  const bool quad = ctx->bp_engine >= 1 || sizeof(TI) != sizeof(T);
  const unsigned pnu = (unsigned)padded_dim(g->nu), pnv = (unsigned)padded_dim(g->nv);
  const size_t total = (size_t)P.pview * nviews;
  int rc = cbct_reserve(ctx, ctx->ws->proj_pad, total * sizeof(T) * (quad ? 4 : 1));
  if (rc) return rc;
  T *pp = static_cast<T *>(ctx->ws->proj_pad.ptr);
  if (quad) {
    hipLaunchKernelGGL((k_pad_quad<TI, T>), pad_grid(pnu, nviews), dim3(256), 0, s, proj, pp, g->nu,
                       g->nv, pnu, pnv, nviews);
  } else {
    hipLaunchKernelGGL((k_pad_proj<T, T>), stride_grid(total), dim3(256), 0, s,
                       reinterpret_cast<const T *>(proj), pp, g->nu, g->nv, pnu, pnv, total);
  }
  CBCT_LAUNCH_CHECK(ctx, "k_pad_proj");
  cbct_prof_mark(ctx, PROF_BP, 0, s);
  if (quad)
    launch_bp<T, true>(P, grid, kz, s, ctx->d_views_f + view0, pp, vol, fbp, accumulate);
  else
    launch_bp<T, false>(P, grid, kz, s, ctx->d_views_f + view0, pp, vol, fbp, accumulate);
  cbct_prof_mark(ctx, PROF_BP, 1, s);
  CBCT_LAUNCH_CHECK(ctx, "k_bp");
  return CBCT_OK;
}

}  // namespace

extern "C" {

int cbct_bp(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *proj, int dtype,
            float *vol, int fbp, int accumulate, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "conebeam_backwards_projection";
  if (!vol || !proj)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (dtype != CBCT_F32 && dtype != CBCT_BF16)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "Error: %s: unsupported projection dtype %d", who, dtype);
  const size_t pview = padded_dim(g->nu) * padded_dim(g->nv);
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  CallScope scope(ctx, s);
  BPParams P = make_params(g, nviews);
  const int kz = bp_kz(ctx, g);
  dim3 grid(ceil_div(g->nx, BP_TX), ceil_div(g->ny, BP_TY), ceil_div(g->nz, kz));
  if (nviews == 0) {
    // No views: the reference clears the image unless accumulating
    // (xray/conebeam_projection.cu:1343-1350).  A fill kernel, not a memset,
    // so the call stays capturable in a HIP graph.
    if (!accumulate) return cbct_fill(ctx, (size_t)g->nx * g->ny * g->nz, vol, 0.0f, stream);
    return CBCT_OK;
  }
  // quad texels: one view of the padded copy must stay below 2^31 bytes
  if (pview * 16 >= (1ull << 31))
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "Error: %s: detector too large (%d x %d)", who,
                     g->nu, g->nv);
  if (dtype == CBCT_BF16)
    return run_bp<unsigned short, unsigned short>(ctx, g, P, grid, kz, view0, nviews,
                                                  static_cast<const unsigned short *>(proj), vol,
                                                  fbp, accumulate, s);
  if (ctx->work_dtype == CBCT_BF16)
    return run_bp<float, unsigned short>(ctx, g, P, grid, kz, view0, nviews,
                                         static_cast<const float *>(proj), vol, fbp, accumulate, s);
  return run_bp<float, float>(ctx, g, P, grid, kz, view0, nviews, static_cast<const float *>(proj),
                              vol, fbp, accumulate, s);
}

const char *cbct_bp_kernel(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews) {
  if (!ctx || !g || cbct_check_geom(ctx, g, "cbct_bp_kernel") ||
      cbct_check_views(ctx, view0, nviews, "cbct_bp_kernel") || nviews == 0)
    return "";
  const BPParams P = make_params(g, nviews);
  return use_staged(ctx, g, P, view0, nviews, bp_kz(ctx, g)) ? "k_bpp" : "k_bp";
}

int cbct_storage(cbct_ctx ctx, int dtype) {
  if (!ctx) return CBCT_ERR_ARG;
  if (dtype != CBCT_F32 && dtype != CBCT_BF16)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "cbct_storage: unsupported storage type %d", dtype);
  ctx->work_dtype = dtype;
  return CBCT_OK;
}

int cbct_bp_engine(cbct_ctx ctx, int engine) {
  if (!ctx) return CBCT_ERR_ARG;
  if (engine < 0 || engine > 3)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "cbct_bp_engine: engine %d not available", engine);
  ctx->bp_engine = engine;
  return CBCT_OK;
}

int cbct_bp_kz(cbct_ctx ctx, int kz) {
  if (!ctx) return CBCT_ERR_ARG;
  if (kz != 0 && kz != 8 && kz != KZ)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "cbct_bp_kz: %d z-voxels per thread not available (0, 8, %d)",
                     kz, KZ);
  ctx->bp_kz = kz;
  return CBCT_OK;
}

int cbct_spacecarver(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const float *proj,
                     unsigned char *mask, float limit, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "conebeam_spacecarver";
  if (!proj || !mask)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  BPParams P = make_params(g, nviews);
  const size_t nvox = (size_t)g->nx * g->ny * g->nz;
  CallScope scope(ctx, as_stream(stream));
  hipLaunchKernelGGL(k_spacecarver, stride_grid(nvox), dim3(256), 0, as_stream(stream), P, g->nu,
                     g->nv, ctx->d_views_f + view0, proj, mask, limit);
  CBCT_LAUNCH_CHECK(ctx, "k_spacecarver");
  return CBCT_OK;
}

}  // extern "C"
