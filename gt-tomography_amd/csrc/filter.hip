// FDK pre-filter chain: cosine weights, Parker short-scan weights, the ramp
// filter as a Toeplitz contraction, and half-fan offset weights.
//
// Reference chain (xray/conebeam_projection.cu:1483-1521): P *= C; [Parker];
// pad each row to 2Nu (data at offset Nu/2); R2C; multiply bins 0..Nu by the
// real response F; C2R; divide by 2Nu; crop [Nu/2, Nu/2+Nu); [offset weights].
// Because the data occupy Nu of the 2Nu padded samples, the circular
// convolution never wraps, and the whole FFT chain is exactly the Toeplitz
// product  y[i] = sum_j h[|i-j|] x[j]  with
//   h[m] = (F0 + 2 sum_{k=1}^{Nu-1} Fk cos(pi k m / Nu) + F_Nu (-1)^m) / (2Nu).
// That is a dense (rows x Nu) * (Nu x Nu) product, run as a tiled
// contraction (FP32 VALU engine, or FP16 hi/lo split MFMA engine with FP32
// accumulation), or, like the reference, by FFT (the default engine: two rows
// per complex transform, N >= 2Nu - 1, so the convolution never wraps).
#include <atomic>
#include <cmath>
#include <vector>

#include "cbct_internal.h"

namespace {

// ------------------------------------------------------------------ weights
// Parker-type short-scan weights, conebeam_projection.cu:149-194.
__device__ __forceinline__ float parker_S(float b) {
  if (b <= -0.5f) return 0.0f;
  if (b < 0.5f) return 0.5f * (1.0f + sinf(3.14159265358979f * b));
  return 1.0f;
}
__device__ __forceinline__ float parker_b(float alpha, float delta) {
  return 0.1f * (2.0f * (delta - alpha) + 0.001f);
}
__device__ float parker_weight(int u, int nu, float beta, float delta, float tan_delta) {
  const float eps = 0.001f;
  const float pi = 3.14159265358979f;
  const float alpha = -atanf(((float)u / (float)nu - 0.5f) * 2.0f * tan_delta);
  return 0.5f * (parker_S(beta / parker_b(alpha, delta) - 0.5f) +
                 parker_S((beta + 2.0f * (alpha - delta) - eps) / parker_b(alpha, delta) + 0.5f) -
                 parker_S((beta - pi + 2.0f * alpha) / parker_b(-alpha, delta) - 0.5f) -
                 parker_S((beta - pi - 2.0f * delta - eps) / parker_b(-alpha, delta) + 0.5f));
}

// out = in * C[u,v] * (parker ? w_parker : 1)
__global__ void __launch_bounds__(256) k_cos_parker(const float *__restrict__ in,
                                                    float *__restrict__ out,
                                                    const float *__restrict__ cosw, int nu, int nv,
                                                    int nviews, const float *__restrict__ angles,
                                                    int parker, float delta, float tan_delta) {
  const size_t plane = (size_t)nu * nv;
  const size_t total = plane * nviews;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t k = t / plane;
    const size_t uv = t - k * plane;
    float v = in[t] * cosw[uv];
    if (parker) {
      const int u = (int)(uv % nu);
      const float beta = angles[k] * (3.14159265358979f / 180.0f);
      v *= parker_weight(u, nu, beta, delta, tan_delta);
    }
    out[t] = v;
  }
}

// Half-fan weights: conebeam_projection.cu:233-262 (full) and :303-331 (sqrt).
__global__ void __launch_bounds__(256) k_offset_weights(float *__restrict__ p, int nu, int nv,
                                                        int nviews, const ViewF *__restrict__ views,
                                                        float fu, int sqrt_variant) {
  const size_t plane = (size_t)nu * nv;
  const size_t total = plane * nviews;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t k = t / plane;
    const int u = (int)((t - k * plane) % nu);
    const float ou = views[k].off_u;
    const float sign = ou / fabsf(ou);
    const float omega = fu / 2.0f - fabsf(ou);
    const float pi = 3.14159265358979f;
    if (!sqrt_variant) {
      const float tt = (fu * ((float)u / (float)nu - 0.5f) + ou) * sign;
      float v = p[t];
      if (fabsf(tt) <= omega) {
        const float sw = sinf(pi * (tt + omega) / (4.0f * omega));
        v *= sw * sw;
      }
      if (tt < -omega) v *= 0.0f;
      p[t] = v;
    } else {
      const float tt = (fu * (((float)u + 0.5f) / (float)nu - 0.5f) + ou) * sign;
      if (fabsf(tt) <= fabsf(omega)) p[t] *= sinf(pi * (tt + omega) / (4.0f * omega));
    }
  }
}

// ------------------------------------------------------- FP32 Toeplitz engine
// Y[r][i] = sum_j hm[i - j + nu - 1] * X[r][j], hm = mirrored kernel (2nu-1).
// Tile: 64 rows x 64 outputs per 256-thread block, 4x4 per thread, K-chunks
// of 64 staged in LDS together with the 127-entry kernel window.
constexpr int TT = 64;

__global__ void __launch_bounds__(256) k_toeplitz_f32(const float *__restrict__ X,
                                                      float *__restrict__ Y,
                                                      const float *__restrict__ hm, int nu,
                                                      size_t nrows) {
  __shared__ float xs[TT][TT + 1];
  __shared__ float hs[2 * TT];
  const int tid = threadIdx.x;
  const int ti = tid & 15;   // output column group
  const int tr = tid >> 4;   // row group
  const int i0 = blockIdx.x * TT;
  const size_t r0 = (size_t)blockIdx.y * TT;
  float acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;

  for (int j0 = 0; j0 < nu; j0 += TT) {
    __syncthreads();
    // stage X[r0..r0+63][j0..j0+63]
    for (int e = tid; e < TT * TT; e += 256) {
      const int rr = e / TT, jj = e % TT;
      const size_t r = r0 + rr;
      const int j = j0 + jj;
      xs[rr][jj] = (r < nrows && j < nu) ? X[r * nu + j] : 0.0f;
    }
    // kernel window: w = (i - i0) - (j - j0) + 63, index hm[i0 - j0 + w - 63 + nu - 1]
    if (tid < 2 * TT - 1) {
      const int m = i0 - j0 + tid - (TT - 1) + nu - 1;
      hs[tid] = (m >= 0 && m < 2 * nu - 1) ? hm[m] : 0.0f;
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < TT; ++jj) {
      float xv[4], hv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) xv[a] = xs[tr * 4 + a][jj];
#pragma unroll
      for (int b = 0; b < 4; ++b) hv[b] = hs[ti * 4 + b - jj + TT - 1];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fmaf(xv[a], hv[b], acc[a][b]);
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const size_t r = r0 + tr * 4 + a;
    if (r >= nrows) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + ti * 4 + b;
      if (i < nu) Y[r * nu + i] = acc[a][b];
    }
  }
}

// ------------------------------------------------- FP16 split MFMA engine
// Y = X T with T[k][n] = h[|n-k|] (symmetric).  Both operands are split into
// fp16 hi + lo parts after a power-of-two scaling (per row for X, one scale
// for T), and three products hi*hi + hi*lo + lo*hi accumulate in FP32; the
// dropped lo*lo term is below 2^-22 of |x||t| per product.
//   X parts: [mp][kp] halves (mp = rows rounded up to MB, kp = nu rounded up
//            to KB, zero padded), row scale inverse rinv[mp].
//   T parts: 8 shifted copies of T's generating row (ensure_toeplitz_f16),
//            np = nu rounded up to NB; T is symmetric, so rows n over k are
//            also its transpose.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Block tile MB x NB, K-steps of KB (one LDS row of a part = KB halves =
// 128 bytes = eight 16-byte chunks).
constexpr int MB = 256, NB = 256, KB = 64;
// the Toeplitz operand advances in K-chunks of one 32-wide MFMA tile, two per
// LDS K-step (its fragment reuse needs chunk = tile width)
constexpr int KC = 32;
static_assert(KB == 2 * KC, "two Toeplitz chunks per K-step");

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// One wave per row: weights (cosine plane, Parker), per-row power-of-two
// scale putting max|x| in [2^13, 2^14), then the fp16 hi/lo split.  Rows of
// up to SPLIT_V x 256 elements (vec: nu a multiple of 4, the input 16-byte
// aligned) stay in registers between the max and the split (16-byte loads,
// 8-byte stores); others are read twice, one element per lane and step.
constexpr int SPLIT_V = 8;

// x[u] with its weights, multiplied in the order of the FP32 chain
__device__ __forceinline__ float weighted(const float *x, const float *cw, int parker, int u, int nu,
                                          float beta, float delta, float tan_delta) {
  float v = x[u];
  if (cw) v *= cw[u];
  if (parker) v *= parker_weight(u, nu, beta, delta, tan_delta);
  return v;
}

__global__ void __launch_bounds__(256) k_split_rows(const float *__restrict__ in,
                                                    const float *__restrict__ cosw, int nu, int nv,
                                                    size_t nrows, size_t mp, int kp,
                                                    const float *__restrict__ angles, int parker,
                                                    float delta, float tan_delta, int vec,
                                                    _Float16 *__restrict__ xh,
                                                    _Float16 *__restrict__ xl,
                                                    float *__restrict__ rinv) {
  const size_t r = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= mp) return;
  _Float16 *oh = xh + r * kp, *ol = xl + r * kp;
  if (r >= nrows) {
    for (int u = lane; u < kp; u += 64) oh[u] = ol[u] = (_Float16)0.0f;
    if (lane == 0) rinv[r] = 0.0f;
    return;
  }
  const float *x = in + r * nu;
  const float *cw = cosw ? cosw + (r % nv) * nu : nullptr;
  const float beta = parker ? angles[r / nv] * (3.14159265358979f / 180.0f) : 0.0f;
  auto scale_of = [](float mx, float &sc, float &inv) {
    sc = inv = 1.0f;
    if (mx > 0.0f && mx <= 3.4e38f) {
      int ex;
      (void)frexpf(mx, &ex);               // mx = f * 2^ex, f in [0.5, 1)
      const int s = min(14 - ex, 126);
      sc = ldexpf(1.0f, s);
      inv = ldexpf(1.0f, -s);
    }
  };
  auto split = [](float v, _Float16 &h, _Float16 &l) {
    h = (_Float16)v;
    l = (_Float16)(v - (float)h);
  };
  if (vec) {
    float4 v[SPLIT_V];
    float mx = 0.0f;
#pragma unroll
    for (int i = 0; i < SPLIT_V; ++i) {
      const int u = 4 * lane + 256 * i;
      v[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (u < nu) {
        v[i] = *reinterpret_cast<const float4 *>(x + u);
        if (cw) {
          const float4 c = *reinterpret_cast<const float4 *>(cw + u);
          v[i].x *= c.x; v[i].y *= c.y; v[i].z *= c.z; v[i].w *= c.w;
        }
        if (parker) {
          v[i].x *= parker_weight(u, nu, beta, delta, tan_delta);
          v[i].y *= parker_weight(u + 1, nu, beta, delta, tan_delta);
          v[i].z *= parker_weight(u + 2, nu, beta, delta, tan_delta);
          v[i].w *= parker_weight(u + 3, nu, beta, delta, tan_delta);
        }
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[i].x), fabsf(v[i].y)), fmaxf(fabsf(v[i].z), fabsf(v[i].w))));
      }
    }
    mx = wave_max(mx);
    float sc, inv;
    scale_of(mx, sc, inv);
#pragma unroll
    for (int i = 0; i < SPLIT_V; ++i) {
      const int u = 4 * lane + 256 * i;
      if (u < kp) {   // kp - nu < 64 and both are multiples of 4: whole groups
        _Float16 h[4], l[4];
        split(v[i].x * sc, h[0], l[0]);
        split(v[i].y * sc, h[1], l[1]);
        split(v[i].z * sc, h[2], l[2]);
        split(v[i].w * sc, h[3], l[3]);
        typedef _Float16 half4 __attribute__((ext_vector_type(4)));
        *reinterpret_cast<half4 *>(oh + u) = half4{h[0], h[1], h[2], h[3]};
        *reinterpret_cast<half4 *>(ol + u) = half4{l[0], l[1], l[2], l[3]};
      }
    }
    if (lane == 0) rinv[r] = inv;
    return;
  }
  float mx = 0.0f;
  for (int u = lane; u < nu; u += 64)
    mx = fmaxf(mx, fabsf(weighted(x, cw, parker, u, nu, beta, delta, tan_delta)));
  mx = wave_max(mx);
  float sc, inv;
  scale_of(mx, sc, inv);
  for (int u = lane; u < kp; u += 64) {
    float v = 0.0f;
    if (u < nu) v = weighted(x, cw, parker, u, nu, beta, delta, tan_delta) * sc;
    _Float16 h, l;
    split(v, h, l);
    oh[u] = h;
    ol[u] = l;
  }
  if (lane == 0) rinv[r] = inv;
}

// 512 threads: eight waves in a 2 x 4 arrangement, each owning a 128 x 64
// output block as 4 x 2 tiles of the 32x32x16 fp16 MFMA.  Per K-step the
// block's 256 x 64 row slice (hi and lo parts) goes straight from global
// memory into LDS (16-byte direct loads, no registers), double-buffered: the
// next step's copies are issued before this step's products, one barrier
// per step.  The Toeplitz operand never goes through LDS: each fragment is 8
// consecutive entries of its generating row (ensure_toeplitz_f16), loaded
// into registers, and one K-chunk (32) along k equals one tile along n, so a
// wave loads one new tile per chunk (two chunks per K-step) and shifts the
// other.  LDS rows are 128 bytes; chunk c of row r is stored at chunk
// c ^ (r & 7), so the 8 rows a group of 8 lanes reads at one chunk fall in 8
// different bank groups, and a direct load (64 lanes -> 1 KB contiguous)
// fills 8 whole rows.  One block (two
// waves per SIMD) per CU.  The products accumulate in the order of K, hi*hi,
// hi*lo, lo*hi per 16 of K.
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) k_toeplitz_mfma(const _Float16 *__restrict__ xh,
                                                       const _Float16 *__restrict__ xl,
                                                       const float *__restrict__ rinv,
                                                       const _Float16 *__restrict__ th,
                                                       const _Float16 *__restrict__ tl, int tlen,
                                                       float *__restrict__ Y, size_t nrows, int nu,
                                                       int kp, unsigned ntn, unsigned nblocks,
                                                       float inv_st) {
  // [stage][part][row][KB halves]: rows of the block's X slice
  __shared__ __attribute__((aligned(1024))) _Float16 sa[2][2][MB * KB];
  const unsigned L = xcd_remap(blockIdx.x, nblocks);
  const unsigned tn = L % ntn, tm = L / ntn;
  const size_t m0 = (size_t)tm * MB;
  const int n0 = (int)tn * NB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 2, wn = wave & 3;
  // direct loads: a 1 KB copy fills 8 rows of 128 bytes; lane -> row lane / 8
  // of them, stored chunk lane & 7, holding global chunk (lane & 7) ^ (row & 7)
  // (8-row groups start at multiples of 8, so row & 7 = lane / 8)
  const int crow = lane >> 3, cchunk = (lane & 7) ^ crow;
  const size_t rowb = (size_t)kp * sizeof(_Float16);
  const int coff = crow * (int)rowb + cchunk * 16;   // bytes, + 8 rows per copy
  auto rsrc = [](const _Float16 *p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
  };
  const __amdgpu_buffer_rsrc_t rah = rsrc(xh + m0 * kp, MB * rowb), ral = rsrc(xl + m0 * kp, MB * rowb);
  const __amdgpu_buffer_rsrc_t rbh = rsrc(th, 8 * (size_t)tlen * sizeof(_Float16)),
                               rbl = rsrc(tl, 8 * (size_t)tlen * sizeof(_Float16));
  typedef __attribute__((address_space(3))) void *lds_ptr;
  // wave w copies 8-row groups w, w + 8, w + 16 and w + 24 of both parts
  // (the group's row offset is uniform: scalar offset; the lane's, one VGPR)
  auto stage = [&](int k0, int st) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = wave + 8 * j;
      const int so = k0 * (int)sizeof(_Float16) + q * 8 * (int)rowb;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rah, (lds_ptr)&sa[st][0][q * 8 * KB], 16, coff, so, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(ral, (lds_ptr)&sa[st][1][q * 8 * KB], 16, coff, so, 0, 0);
    }
  };
  floatx16 acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.0f;

  // fragment reads: lane (fr, fk) reads row fr of a 32-row tile, global
  // chunk g = 2 kk + fk (8 halves each), stored at g ^ (fr & 7)
  const int fr = lane & 31, fk = lane >> 5, sw = fr & 7;
  // T's fragments come from the phase copies of its generating row (see
  // ensure_toeplitz_f16): lane (fr, fk) of tile t needs T[n][k .. k+7] =
  // G[S0 - n + k ..] with k = k0 + 16 kk + 8 fk, 16-byte aligned in copy
  // q = (S0 - n) & 7 (k is a multiple of 8).  Byte offsets at k0 = 0:
  const int S0 = (int)ntn * NB - 1;
  unsigned boff[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int i0 = S0 - (n0 + wn * 64 + t * 32 + fr);
    const int q = i0 & 7;
    boff[t] = (unsigned)(q * tlen + (i0 - q) + 8 * fk) * (unsigned)sizeof(_Float16);
  }
  // fragment (tile t, k-half kk) of part r (hi or lo) at K-chunk k0
  auto load_b = [&](const __amdgpu_buffer_rsrc_t &r, int k0, int t, int kk) __attribute__((always_inline)) {
    return __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(
                                         r, (int)boff[t] + 32 * kk, k0 * (int)sizeof(_Float16), 0));
  };
  half8 bh[2][2], bl[2][2];
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      bh[kk][t] = load_b(rbh, 0, t, kk);
      bl[kk][t] = load_b(rbl, 0, t, kk);
    }
  stage(0, 0);
  for (int k0 = 0, st = 0; k0 < kp; k0 += KB, st ^= 1) {
    // this step's copies have landed (every wave's, after the barrier), and
    // every wave is done with the other buffer's products
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int kc = k0 + h * KC;   // this half's 32 of K
      // the half's T fragments are in registers.  T is Toeplitz and KC is
      // the tile width, so the next chunk's tile 1 is this chunk's tile 0
      // (T[n + 32][k + 32] = T[n][k]); only its tile 0 is loaded, now, so it
      // arrives during this half's products
      half8 bhn[2], bln[2];
      if (kc + KC < kp) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          bhn[kk] = load_b(rbh, kc + KC, 0, kk);
          bln[kk] = load_b(rbl, kc + KC, 0, kk);
        }
      }
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        // one 16-wide slice's fragments, then its products (registers)
        const int ch = ((2 * (2 * h + kk) + fk) ^ sw) * 8;   // halves
        half8 ah[4], al[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int ar = (wm * 128 + t * 32 + fr) * KB + ch;
          ah[t] = *reinterpret_cast<const half8 *>(&sa[st][0][ar]);
          al[t] = *reinterpret_cast<const half8 *>(&sa[st][1][ar]);
        }
        // the next step's copies go out after this step's last fragment
        // reads (the compiler makes LDS reads that follow a direct copy wait
        // for it, whatever the buffer)
        if (h == 1 && kk == 1 && k0 + KB < kp) stage(k0 + KB, st ^ 1);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bh[kk][b], acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bl[kk][b], acc[a][b], 0, 0, 0);
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bh[kk][b], acc[a][b], 0, 0, 0);
          }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        bh[kk][1] = bh[kk][0];
        bl[kk][1] = bl[kk][0];
        bh[kk][0] = bhn[kk];
        bl[kk][0] = bln[kk];
      }
    }
  }
  // D layout: lane holds column n = lane % 32, rows 8(i/4) + 4(lane/32) + i%4
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = n0 + wn * 64 + b * 32 + fr;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const size_t m = m0 + wm * 128 + a * 32 + 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
        if (m < nrows && n < nu) Y[m * nu + n] = acc[a][b][i] * (rinv[m] * inv_st);
      }
    }
}

inline dim3 stride_grid(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 262144) b = 262144;
  if (b == 0) b = 1;
  return dim3((unsigned)b);
}

inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Host: dense fp16 hi/lo Toeplitz matrices from the FP64 kernel, built once
// per filter on first use of the MFMA engine.
int ensure_toeplitz_f16(cbct_ctx ctx) {
  if (ctx->d_thi) return CBCT_OK;
  const int nu = ctx->filter_nu;
  const size_t kp = round_up(nu, KB), np = round_up(nu, NB);
  double hmax = 0.0;
  for (int m = 0; m < nu; ++m) hmax = std::fmax(hmax, std::fabs(ctx->h64[m]));
  int ex = 0;
  if (hmax > 0.0) (void)std::frexp(hmax, &ex);
  const double sc = std::ldexp(1.0, 14 - ex);
  // T[n][k] = h[|n - k|] for n, k < nu (scaled, hi + lo fp16).  Stored as
  // its generating row read backwards, G[s] = g[S0 - s] with g[t] = h[|t|]
  // (0 for |t| >= nu), S0 = np - 1, so that T[n][k + j] = G[S0 - n + k + j]
  // for j = 0..7: one MFMA fragment is 8 consecutive entries.  Eight copies,
  // copy q holding G shifted by q (G_q[s] = G[s + q]), make every fragment a
  // 16-byte aligned load.  Rows n >= nu and columns k >= nu of this T need
  // not be zero: those outputs are discarded and those row entries are.
  const long S0 = (long)np - 1;
  const size_t tlen = round_up((size_t)S0 + kp, 8);
  std::vector<_Float16> hi(8 * tlen), lo(8 * tlen);
  for (size_t q = 0; q < 8; ++q)
    for (size_t i = 0; i < tlen; ++i) {
      const long t = S0 - (long)(i + q);
      const long at = t < 0 ? -t : t;
      const double v = at < nu ? ctx->h64[at] * sc : 0.0;
      const _Float16 h = (_Float16)v;
      hi[q * tlen + i] = h;
      lo[q * tlen + i] = (_Float16)(v - (double)h);
    }
  CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_thi, sizeof(_Float16) * hi.size()));
  CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_tlo, sizeof(_Float16) * lo.size()));
  int rc = cbct_upload(ctx, ctx->d_thi, hi.data(), sizeof(_Float16) * hi.size());
  if (rc) return rc;
  rc = cbct_upload(ctx, ctx->d_tlo, lo.data(), sizeof(_Float16) * lo.size());
  if (rc) return rc;
  ctx->toe_ld = (int)tlen;
  ctx->toe_inv_scale = (float)std::ldexp(1.0, ex - 14);
  return CBCT_OK;
}

// Weighted (cosine plane, Parker) rows -> MFMA Toeplitz product into out.
// The split stages the input in the workspace, so in may alias out.
int run_toeplitz_mfma(cbct_ctx ctx, int nu, int nv, size_t nrows, const float *in, float *out,
                      const float *cosw, const float *angles, int parker, float delta,
                      float tan_delta, hipStream_t s) {
  if (nrows == 0) return CBCT_OK;
  int rc = ensure_toeplitz_f16(ctx);
  if (rc) return rc;
  const size_t kp = round_up(nu, KB);
  const size_t mp = round_up(nrows, MB);
  const unsigned ntn = (unsigned)(round_up(nu, NB) / NB);
  const size_t nblocks = mp / MB * ntn;
  if (nblocks >= (1ull << 31) || mp / 4 >= (1ull << 31))
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_ramp_filter: too many rows");
  const size_t part = mp * kp * sizeof(_Float16);
  rc = cbct_reserve(ctx, ctx->ws->filt_tmp, 2 * part + mp * sizeof(float));
  if (rc) return rc;
  char *ws = static_cast<char *>(ctx->ws->filt_tmp.ptr);
  _Float16 *xh = reinterpret_cast<_Float16 *>(ws);
  _Float16 *xl = reinterpret_cast<_Float16 *>(ws + part);
  float *rinv = reinterpret_cast<float *>(ws + 2 * part);
  const int vec = nu % 4 == 0 && nu <= SPLIT_V * 256 && reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                  reinterpret_cast<uintptr_t>(cosw) % 16 == 0;
  hipLaunchKernelGGL(k_split_rows, dim3((unsigned)(mp / 4)), dim3(256), 0, s, in, cosw, nu, nv,
                     nrows, mp, (int)kp, angles, parker, delta, tan_delta, vec, xh, xl, rinv);
  CBCT_LAUNCH_CHECK(ctx, "k_split_rows");
  cbct_prof_mark(ctx, PROF_FILTER, 0, s);
  hipLaunchKernelGGL(k_toeplitz_mfma, dim3((unsigned)nblocks), dim3(512), 0, s, xh, xl, rinv,
                     reinterpret_cast<const _Float16 *>(ctx->d_thi),
                     reinterpret_cast<const _Float16 *>(ctx->d_tlo), ctx->toe_ld, out, nrows, nu,
                     (int)kp, ntn, (unsigned)nblocks, ctx->toe_inv_scale);
  cbct_prof_mark(ctx, PROF_FILTER, 1, s);
  CBCT_LAUNCH_CHECK(ctx, "k_toeplitz_mfma");
  return CBCT_OK;
}

// ------------------------------------------------------------- FFT engine
// The convolution evaluated as the reference does it, by FFT, two rows at a
// time.  With c[t] = h[|t|] for |t| < nu (circular, zero elsewhere) and its
// real, even spectrum C over N = 2^L >= 2nu - 1 points, the circular
// convolution of a row zero padded to N is the Toeplitz product on the first
// nu outputs.  C is real, so rows a and b go through one complex transform as
// z = a / s_a + i b / s_b (s: a power of two near the row's L2 norm, so that
// neither row's rounding swamps the other's), and, since
// IFFT(Z C) = conj(FFT(conj(Z) C / N)), both transforms are forward ones:
// v = FFT(conj(FFT(z)) C / N), y_a = s_a Re v, y_b = -s_b Im v.
//
// Each transform is a Stockham FFT: passes of radix P (the points a thread
// holds: 16 for N >= 256, else 8), then one smaller radix pass; pass (R, NS)
// takes butterfly j's inputs X[j + r N/R], multiplies them by
// W_N^(r (j mod NS) N / (NS R)), and writes its R outputs to
// Y[(j / NS) NS R + j mod NS + s NS].  N/P threads per row pair; the passes
// exchange points through LDS (two buffers, padded per exchange).  A
// thread's P points sit at j0 + m N/P both as the first pass's inputs and as
// the last pass's outputs, so the loads, the spectral product between the
// transforms and the stores stay in registers and are coalesced.  One
// workgroup holds max(1, 2048 / N) pairs.
struct cf {
  float x, y;
};
__device__ __forceinline__ cf operator+(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf operator-(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf mul_mi(cf a) { return {a.y, -a.x}; }   // a * (-i)

__device__ __forceinline__ void dft2(cf &a0, cf &a1) {
  const cf t = a0 - a1;
  a0 = a0 + a1;
  a1 = t;
}
// natural order in and out: y_s = sum_r a_r (-i)^(r s)
__device__ __forceinline__ void dft4(cf &a0, cf &a1, cf &a2, cf &a3) {
  const cf t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_mi(a1 - a3);
  a0 = t0 + t2;
  a1 = t1 + t3;
  a2 = t0 - t2;
  a3 = t1 - t3;
}
// y_s = sum_r a_r W8^(r s), W8 = exp(-i pi / 4): even and odd halves by dft4
__device__ __forceinline__ void dft8(cf *a) {
  cf e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6], o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  const float r = 0.70710678118654752f;
  o1 = {r * (o1.x + o1.y), r * (o1.y - o1.x)};    // * W8   = (1 - i) / sqrt 2
  o2 = mul_mi(o2);                                 // * W8^2 = -i
  o3 = {r * (o3.y - o3.x), -r * (o3.x + o3.y)};   // * W8^3 = (-1 - i) / sqrt 2
  a[0] = e0 + o0;
  a[1] = e1 + o1;
  a[2] = e2 + o2;
  a[3] = e3 + o3;
  a[4] = e0 - o0;
  a[5] = e1 - o1;
  a[6] = e2 - o2;
  a[7] = e3 - o3;
}

// y_s = sum_r a_r W16^(r s) as 4 x 4: t[n2][k1] = DFT4 over n1 of a[4 n1 + n2],
// times W16^(n2 k1), then y[k1 + 4 k2] = DFT4 over n2 of t[n2][k1]
__device__ __forceinline__ void dft16(cf *a) {
  constexpr float C1 = 0.92387953251128676f, S1 = 0.38268343236508977f, R2 = 0.70710678118654752f;
  cf t[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cf x0 = a[n2], x1 = a[4 + n2], x2 = a[8 + n2], x3 = a[12 + n2];
    dft4(x0, x1, x2, x3);
    t[n2][0] = x0;
    t[n2][1] = x1;
    t[n2][2] = x2;
    t[n2][3] = x3;
  }
  // W16^j for j = n2 k1: 1, 2, 3 (n2 = 1), 2, 4, 6 (n2 = 2), 3, 6, 9 (n2 = 3)
  t[1][1] = cmul(t[1][1], {C1, -S1});
  t[1][2] = cmul(t[1][2], {R2, -R2});
  t[1][3] = cmul(t[1][3], {S1, -C1});
  t[2][1] = cmul(t[2][1], {R2, -R2});
  t[2][2] = mul_mi(t[2][2]);
  t[2][3] = cmul(t[2][3], {-R2, -R2});
  t[3][1] = cmul(t[3][1], {S1, -C1});
  t[3][2] = cmul(t[3][2], {-R2, -R2});
  t[3][3] = cmul(t[3][3], {-C1, S1});
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cf y0 = t[0][k1], y1 = t[1][k1], y2 = t[2][k1], y3 = t[3][k1];
    dft4(y0, y1, y2, y3);
    a[k1] = y0;
    a[k1 + 4] = y1;
    a[k1 + 8] = y2;
    a[k1 + 12] = y3;
  }
}

// LDS barrier: waits for this wave's LDS accesses only, so global loads
// issued earlier (the next pair's rows) stay in flight across it
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Each thread holds P points (16 for N >= 256, else 8); passes of radix P,
// then one smaller radix pass
template <int N, int NS, int P>
struct FftRadix {
  static constexpr int R = N / NS >= P ? P : N / NS;
};

// twiddles a butterfly keeps: all R - 1 (the kernel is bound by vector
// instruction issue and holds two waves per SIMD, so registers are cheaper
// than forming twiddles as products in every pass)
template <int R>
struct FftKeep {
  static constexpr int value = R - 1;
};

// twiddles one transform keeps per thread: (P/R)(R - 1) per pass with NS > 1
template <int N, int NS, int P>
struct FftTw {
  static constexpr int R = FftRadix<N, NS, P>::R;
  static constexpr int here = NS > 1 ? (P / R) * FftKeep<R>::value : 0;
  static constexpr int value = here + (NS * R < N ? FftTw<N, NS * R, P>::value : 0);
};
template <int N, int P>
struct FftTw<N, N, P> {
  static constexpr int here = 0, value = 0;
};

// the thread's twiddles from the table, in the order the passes use them:
// pass (R, NS), butterfly j = j0 + q N/P, input r: W_N^(r (j mod NS) N/(NS R))
template <int N, int NS, int P, int OFF, int NT>
__device__ __forceinline__ void fft_tw_load(cf (&w)[NT], int j0, const cf *__restrict__ tw) {
  constexpr int R = FftRadix<N, NS, P>::R, K = FftKeep<R>::value;
  if constexpr (NS > 1) {
#pragma unroll
    for (int q = 0; q < P / R; ++q) {
      const int k = (j0 + q * (N / P)) & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) w[OFF + q * K + r - 1] = tw[k * r * (N / (NS * R))];
    }
  }
  if constexpr (NS * R < N) fft_tw_load<N, NS * R, P, OFF + FftTw<N, NS, P>::here, NT>(w, j0, tw);
}

// pass (R, NS) on the thread's butterflies j = j0 + q N/P, q < P/R, whose
// inputs are a[q R + r]; the outputs replace them as a[q R + s]
template <int N, int NS, int P, int OFF, int NT>
__device__ __forceinline__ void fft_pass(cf (&a)[P], const cf (&w)[NT]) {
  constexpr int R = FftRadix<N, NS, P>::R, K = FftKeep<R>::value;
#pragma unroll
  for (int q = 0; q < P / R; ++q) {
    cf *v = &a[q * R];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[OFF + q * K + r - 1]);
    }
    if constexpr (R == 16) dft16(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
    else dft2(v[0], v[1]);
  }
}

// LDS padding of the exchange after pass NS (positions p < N, padded below
// N + N/8): the pass's writes and the next pass's reads both fall on
// distinct banks for every 32 lanes (64 banks of 4 bytes, 8-byte points).
// After the first pass (writes P apart): p + p/32; after the second (groups
// of P consecutive, P^2 apart): p + P (p / P^2); later passes write
// consecutive positions: none.
template <int NS, int P>
__device__ __forceinline__ int fft_padx(int p) {
  if constexpr (NS == 1) return p + (p >> 5);
  else if constexpr (NS == P) return p + (p / (P * P)) * P;
  else return p;
}
template <int NS, int P>
struct FftPadBlock {   // offsets that are multiples of this pad to a constant
  static constexpr int value = NS == 1 ? 32 : (NS == P ? P * P : 1);
};

// passes from sub-transform size NS on; a[] holds that pass's inputs, and on
// return the last pass's outputs.  Writes alternate between X and Y.  The
// padded positions are written as one base register plus constant offsets
// wherever the offsets allow it.
template <int N, int NS, int P, int OFF, int NT>
__device__ __forceinline__ void fft_from(cf (&a)[P], int j0, const cf (&w)[NT], cf *X, cf *Y) {
  constexpr int R = FftRadix<N, NS, P>::R;
  fft_pass<N, NS, P, OFF, NT>(a, w);
  if constexpr (NS * R < N) {
    constexpr int R2 = FftRadix<N, NS * R, P>::R;
    if constexpr (NS == 1) {   // R = P, one butterfly j0: positions P j0 + s
      cf *x = X + fft_padx<1, P>(P * j0);
#pragma unroll
      for (int s = 0; s < P; ++s) x[s] = a[s];
    } else if constexpr (NS == P) {   // R = P: positions P^2 (j0 / P) + j0 % P + P s
      cf *x = X + (P * P + P) * (j0 / P) + (j0 % P);
#pragma unroll
      for (int s = 0; s < P; ++s) x[P * s] = a[s];
    } else {
#pragma unroll
      for (int q = 0; q < P / R; ++q) {
        const int j = j0 + q * (N / P);
        cf *x = X + (j & ~(NS - 1)) * R + (j & (NS - 1));
#pragma unroll
        for (int s = 0; s < R; ++s) x[s * NS] = a[q * R + s];
      }
    }
    lds_barrier();
    const cf *y = X + fft_padx<NS, P>(j0);
    constexpr int B = FftPadBlock<NS, P>::value;
#pragma unroll
    for (int q = 0; q < P / R2; ++q)
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const int c = q * (N / P) + r * (N / R2);
        a[q * R2 + r] = c % B == 0 ? y[fft_padx<NS, P>(c)] : X[fft_padx<NS, P>(j0 + c)];
      }
    fft_from<N, NS * R, P, OFF + FftTw<N, NS, P>::here, NT>(a, j0, w, Y, X);
  }
}

template <int N, int NS, int P>
struct FftWrites {   // LDS writes (passes - 1) of one transform
  static constexpr int R = FftRadix<N, NS, P>::R;
  static constexpr int value = NS * R < N ? 1 + FftWrites<N, NS * R, P>::value : 0;
};
template <int N, int P>
struct FftWrites<N, N, P> {
  static constexpr int value = 0;
};

template <int LOG2N>
struct FftShape {
  static constexpr int N = 1 << LOG2N;
  static constexpr int LP = LOG2N >= 8 ? 4 : 3;          // log2 of the points per thread
  static constexpr int P = 1 << LP;
  static constexpr int G = N >= 2048 ? 1 : 2048 / N;   // row pairs per workgroup
  static constexpr int T = G * N / P;                  // threads
  static constexpr int RL = N >> (LP * ((LOG2N - 1) / LP));   // radix of the last pass
  static constexpr int PN = N + N / 8;                 // padded length of one pair's buffer
};

// Persistent: workgroup b takes pair groups b, b + gridDim.x, ...; the next
// group's rows (and cosine weights) are loaded while this one transforms.
// PARKER: the short-scan weights, in an instantiation of their own (their
// range reductions would otherwise set the common kernel's register count).
template <int LOG2N, bool PARKER>
__global__ void __launch_bounds__(FftShape<LOG2N>::T) k_ramp_fft(
    const float *in, float *out, size_t nrows, int nu, const cf *__restrict__ tw,
    const float *__restrict__ cs, const float *__restrict__ cosw, int nv,
    const float *__restrict__ angles, float delta, float tan_delta) {
  using S = FftShape<LOG2N>;
  constexpr int N = S::N, P = S::P, NP = N / P, G = S::G, T = S::T, RL = S::RL, PN = S::PN;
  constexpr int H = P / 2;   // points j0 + m N/P with m >= P/2 are >= N/2 >= nu: never data
  constexpr int NW = T / 64 > 0 ? T / 64 : 1;
  constexpr int NT = FftTw<N, 1, P>::value > 0 ? FftTw<N, 1, P>::value : 1;
  __shared__ cf buf[2][G * PN];
  __shared__ double part[2][NW];
  __shared__ float partf[2][NW];
  const int t = threadIdx.x, g = t / NP, j0 = t & (NP - 1);
  const size_t ngroups = (nrows + 2 * G - 1) / (2 * G);

  cf w[NT];
  fft_tw_load<N, 1, P, 0, NT>(w, j0, tw);
  float csr[P];
#pragma unroll
  for (int m = 0; m < P; ++m) csr[m] = cs[j0 + m * NP];

  // Rows go through buffer resources over one group's 2G rows (base and
  // size uniform, 32-bit offsets): reads past the group's last row or past
  // nu return 0 and such writes are dropped, so no load or store sits under
  // a branch and the memory counters stay exact across iterations (loads
  // and stores share one counter, so an uncertain count would make every
  // iteration wait for its stores).
  auto rsrc = [](const float *p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
  };
  auto group_rows = [&](size_t gi) -> size_t {
    const size_t r0 = gi * 2 * G;
    return r0 >= nrows ? 0 : (nrows - r0 < 2 * (size_t)G ? nrows - r0 : 2 * (size_t)G);
  };
  constexpr unsigned OOB = 0x80000000u;   // beyond any group's bytes
  const __amdgpu_buffer_rsrc_t rcos = rsrc(cosw, cosw ? (size_t)nv * nu * sizeof(float) : 0);
  // raw rows and cosine weights of group gi
  float x[2][H], cwv[2][H];
  auto fetch = [&](size_t gi) {
    const __amdgpu_buffer_rsrc_t rin =
        rsrc(in + gi * 2 * G * (size_t)nu, group_rows(gi) * nu * sizeof(float));
    const unsigned v0 = (unsigned)((gi * 2 * G) % (size_t)nv);   // uniform
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int rl = 2 * g + h;   // row within the group
      // v0 < nv and rl < 2G: one subtraction unless a view has fewer rows
      unsigned v = v0 + (unsigned)rl;
      if (2 * G <= nv) {
        if (v >= (unsigned)nv) v -= (unsigned)nv;
      } else {
        v %= (unsigned)nv;
      }
#pragma unroll
      for (int m = 0; m < H; ++m) {
        const int e = j0 + m * NP;
        const unsigned off = e < nu ? (unsigned)(rl * nu + e) * 4u : OOB;
        const unsigned coff = e < nu ? (v * (unsigned)nu + (unsigned)e) * 4u : OOB;
        x[h][m] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, off, 0, 0));
        cwv[h][m] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rcos, coff, 0, 0));
      }
    }
  };
  // one group: consume its prefetched rows, fetch the next group's, transform,
  // store.  The first group is peeled off the loop, so that every entry to
  // the loop follows a group's stores, as the back edge does: the memory
  // counters then agree at the loop head and waiting for the prefetched
  // rows never waits for the previous group's stores (loads and stores
  // share one in-order counter).
  auto process = [&](size_t gi) __attribute__((always_inline)) {
    const size_t ra = (gi * G + g) * 2;
    const __amdgpu_buffer_rsrc_t rout =
        rsrc(out + gi * 2 * G * (size_t)nu, group_rows(gi) * nu * sizeof(float));
    // the pair's weighted rows
    float xv[2][H];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const size_t r = ra + h;
      float beta = 0.0f;
      if (PARKER && r < nrows) beta = angles[r / (size_t)nv] * (3.14159265358979f / 180.0f);
#pragma unroll
      for (int m = 0; m < H; ++m) {
        float v = x[h][m];
        if (cosw) v *= cwv[h][m];
        if (PARKER) v *= parker_weight(j0 + m * NP, nu, beta, delta, tan_delta);
        xv[h][m] = v;
      }
    }
    fetch(gi + gridDim.x);
    // their sums of squares over the pair's N/P threads (within one wave up
    // to 64): in FP32, and again in FP64 for a pair whose FP32 sums leave
    // [2^-100, 2^120] (tiny, huge, zero or not finite rows).  Only a power of
    // two near the norm is needed, so FP32 rounding does not matter; a pair
    // that spans several waves is alone in its workgroup, so the FP64 pass's
    // barrier is taken by the whole workgroup or not at all.
    double qs[2];
    {
      float qf[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float q = 0.0f;
#pragma unroll
        for (int m = 0; m < H; ++m) q = fmaf(xv[h][m], xv[h][m], q);
        qf[h] = q;
      }
#pragma unroll
      for (int o = (NP < 64 ? NP : 64) / 2; o > 0; o >>= 1) {
        qf[0] += __shfl_xor(qf[0], o);
        qf[1] += __shfl_xor(qf[1], o);
      }
      if constexpr (NP > 64) {
        if ((t & 63) == 0) {
          partf[0][t >> 6] = qf[0];
          partf[1][t >> 6] = qf[1];
        }
        lds_barrier();
        qf[0] = qf[1] = 0.0f;
#pragma unroll
        for (int wv = 0; wv < NP / 64; ++wv) {
          qf[0] += partf[0][g * (NP / 64) + wv];
          qf[1] += partf[1][g * (NP / 64) + wv];
        }
      }
      qs[0] = qf[0];
      qs[1] = qf[1];
      const bool in_range = qf[0] >= 0x1p-100f && qf[0] <= 0x1p120f && qf[1] >= 0x1p-100f &&
                            qf[1] <= 0x1p120f;
      if (!in_range) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          double q = 0.0;
#pragma unroll
          for (int m = 0; m < H; ++m) q += (double)xv[h][m] * (double)xv[h][m];
          qs[h] = q;
        }
#pragma unroll
        for (int o = (NP < 64 ? NP : 64) / 2; o > 0; o >>= 1) {
          qs[0] += __shfl_xor(qs[0], o);
          qs[1] += __shfl_xor(qs[1], o);
        }
        if constexpr (NP > 64) {
          if ((t & 63) == 0) {
            part[0][t >> 6] = qs[0];
            part[1][t >> 6] = qs[1];
          }
          lds_barrier();
          qs[0] = qs[1] = 0.0;
#pragma unroll
          for (int wv = 0; wv < NP / 64; ++wv) {
            qs[0] += part[0][g * (NP / 64) + wv];
            qs[1] += part[1][g * (NP / 64) + wv];
          }
        }
      }
    }
    // power-of-two scales: |row / s| in [1, 2); zero rows give zero.  With
    // q = f 2^E (f in [1/2, 1)), s = 2^((E - 1) >> 1) puts q / s^2 in [1, 4)
    float sc[2], inv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (qs[h] > 0.0) {
        int e = 0;
        (void)frexp(qs[h], &e);
        e = (e - 1) >> 1;
        e = e < -120 ? -120 : (e > 120 ? 120 : e);
        sc[h] = ldexpf(1.0f, e);
        inv[h] = ldexpf(1.0f, -e);
      } else {
        sc[h] = inv[h] = 0.0f;
      }
    }
    cf a[P];
#pragma unroll
    for (int m = 0; m < H; ++m) a[m] = {xv[0][m] * inv[0], xv[1][m] * inv[1]};
#pragma unroll
    for (int m = H; m < P; ++m) a[m] = {0.0f, 0.0f};
    cf *X = &buf[0][g * PN], *Y = &buf[1][g * PN];
    fft_from<N, 1, P, 0, NT>(a, j0, w, X, Y);
    // a[q RL + s] = Z[j0 + (q + s P/RL) N/P]: the spectral product, conjugated,
    // into the first pass's order b[m] = V[j0 + m N/P]
    cf b[P];
#pragma unroll
    for (int q = 0; q < P / RL; ++q)
#pragma unroll
      for (int s = 0; s < RL; ++s) {
        const int m = q + s * (P / RL);
        const cf z = a[q * RL + s];
        b[m] = {z.x * csr[m], -z.y * csr[m]};
      }
    // the next write goes to the buffer written last but one (a slower wave
    // may still read the last one); each transform writes FftWrites times,
    // so every group starts on X again
    constexpr bool odd = FftWrites<N, 1, P>::value & 1;
    fft_from<N, 1, P, 0, NT>(b, j0, w, odd ? Y : X, odd ? X : Y);
    // outputs at j0 + m N/P; only m < P/2 can be below nu
#pragma unroll
    for (int q = 0; q < P / RL; ++q)
#pragma unroll
      for (int s = 0; s < RL; ++s) {
        const int m = q + s * (P / RL);
        if (m >= H) continue;
        const int e = j0 + m * NP;
        const cf v = b[q * RL + s];
        const unsigned oa = e < nu ? (unsigned)(2 * g * nu + e) * 4u : OOB;
        const unsigned ob = e < nu ? (unsigned)((2 * g + 1) * nu + e) * 4u : OOB;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v.x * sc[0]), rout, oa, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, -v.y * sc[1]), rout, ob, 0, 0);
      }
  };
  fetch(blockIdx.x);
  if ((size_t)blockIdx.x < ngroups) {
    process(blockIdx.x);
    for (size_t gi = (size_t)blockIdx.x + gridDim.x; gi < ngroups; gi += gridDim.x) process(gi);
  }
}

// Host: one-sided response F[0..nu] -> spatial kernel h[0..nu-1] (FP64).
std::vector<double> spatial_kernel(int nu, const float *F) {
  std::vector<double> h(nu);
  const double pi = 3.14159265358979323846;
  const double inv = 1.0 / (2.0 * nu);
  for (int m = 0; m < nu; ++m) {
    double acc = (double)F[0] + (double)F[nu] * ((m & 1) ? -1.0 : 1.0);
    for (int k = 1; k < nu; ++k) {
      // reduce k*m mod 2nu exactly before taking the cosine
      const long long km = ((long long)k * m) % (2LL * nu);
      acc += 2.0 * (double)F[k] * std::cos(pi * (double)km / nu);
    }
    h[m] = acc * inv;
  }
  return h;
}

int ensure_cos_weights(cbct_ctx ctx, const cbct_geom *g) {
  if (ctx->cosw_valid && ctx->cosw_geom.nu == g->nu && ctx->cosw_geom.nv == g->nv &&
      ctx->cosw_geom.fov_u == g->fov_u && ctx->cosw_geom.fov_v == g->fov_v &&
      ctx->cosw_geom.sad == g->sad && ctx->cosw_geom.sdd == g->sdd)
    return CBCT_OK;
  std::vector<float> w((size_t)g->nu * g->nv);
  int rc = cbct_cosine_weights(g, w.data());
  if (rc) return cbct_fail(ctx, rc, "compute_cosine_weights: invalid geometry");
  if (ctx->d_cosw) (void)hipFree(ctx->d_cosw);
  ctx->d_cosw = nullptr;
  ctx->cosw_valid = false;
  CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_cosw, sizeof(float) * w.size()));
  rc = cbct_upload(ctx, ctx->d_cosw, w.data(), sizeof(float) * w.size());
  if (rc) return rc;
  ctx->cosw_geom = *g;
  ctx->cosw_valid = true;
  return CBCT_OK;
}

constexpr int FFT_MIN_LOG2N = 6, FFT_MAX_LOG2N = 13;   // N = 64 .. 8192, nu <= 4096

// log2 of the FFT engine's transform length for rows of nu: N >= 2nu - 1
int fft_log2n(int nu) {
  int l = FFT_MIN_LOG2N;
  while ((1 << l) < 2 * nu - 1) ++l;
  return l;
}

// Host: the FFT engine's tables for the current filter, from its FP64
// kernel: twiddles W_N^m = exp(-2 pi i m / N), m < N, and the spectrum of
// the circular kernel c[t] = h[|t|] (|t| < nu), C[k] = h[0] + 2 sum_t h[t]
// cos(2 pi k t / N), divided by N.  Built once per filter on first use.
int ensure_fft_tables(cbct_ctx ctx, int log2n) {
  if (ctx->d_fft && ctx->fft_log2n == log2n) return CBCT_OK;
  const int nu = ctx->filter_nu;
  const size_t N = size_t(1) << log2n;
  const double pi = 3.14159265358979323846;
  std::vector<double> cs(N);
  std::vector<float> tab(3 * N);
  for (size_t m = 0; m < N; ++m) {
    cs[m] = std::cos(2.0 * pi * (double)m / (double)N);
    tab[2 * m] = (float)cs[m];
    tab[2 * m + 1] = (float)-std::sin(2.0 * pi * (double)m / (double)N);
  }
  const std::vector<double> &h = ctx->h64;
  for (size_t k = 0; k <= N / 2; ++k) {
    double acc = h[0];
    for (int t = 1; t < nu; ++t) acc += 2.0 * h[t] * cs[(k * (size_t)t) & (N - 1)];
    tab[2 * N + k] = (float)(acc / (double)N);
    tab[2 * N + ((N - k) & (N - 1))] = tab[2 * N + k];
  }
  if (ctx->d_fft) {
    int rc = cbct_wait_tables(ctx);
    if (rc) return rc;
    (void)hipFree(ctx->d_fft);
    ctx->d_fft = nullptr;
  }
  CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_fft, sizeof(float) * tab.size()));
  int rc = cbct_upload(ctx, ctx->d_fft, tab.data(), sizeof(float) * tab.size());
  if (rc) return rc;
  ctx->fft_log2n = log2n;
  return CBCT_OK;
}

// workgroups of k_ramp_fft<L> that fit on the device at once (its grid)
std::atomic<int> g_fft_resident[2][FFT_MAX_LOG2N + 1];   // [parker][log2 N]

template <int L>
int launch_ramp_fft(cbct_ctx ctx, size_t ngroups, hipStream_t s, const float *in, float *out,
                    size_t nrows, int nu, const float *tab, const float *cosw, int nv,
                    const float *angles, int parker, float delta, float tan_delta) {
  const size_t N = size_t(1) << L;
  const int pk = parker ? 1 : 0;
  int resident = g_fft_resident[pk][L].load();
  if (resident == 0) {
    int per_cu = 0, cus = 0;
    if (parker)
      CBCT_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ramp_fft<L, true>,
                                                                        FftShape<L>::T, 0));
    else
      CBCT_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ramp_fft<L, false>,
                                                                        FftShape<L>::T, 0));
    CBCT_HIP_CHECK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    resident = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    g_fft_resident[pk][L].store(resident);
  }
  const unsigned nb = (unsigned)(ngroups < (size_t)resident ? ngroups : (size_t)resident);
  // the short-scan weights (Parker) only in their own instantiation: their
  // range reductions would otherwise set the common kernel's register count
  if (parker)
    hipLaunchKernelGGL((k_ramp_fft<L, true>), dim3(nb), dim3(FftShape<L>::T), 0, s, in, out, nrows,
                       nu, reinterpret_cast<const cf *>(tab), tab + 2 * N, cosw, nv, angles, delta,
                       tan_delta);
  else
    hipLaunchKernelGGL((k_ramp_fft<L, false>), dim3(nb), dim3(FftShape<L>::T), 0, s, in, out, nrows,
                       nu, reinterpret_cast<const cf *>(tab), tab + 2 * N, cosw, nv, angles, delta,
                       tan_delta);
  return CBCT_OK;
}

// Weighted (cosine plane, Parker; none when cosw is null) rows -> FFT-engine
// filter into out.  Each workgroup reads its rows whole before it writes
// them, so in may alias out.  Rows longer than 4096 go to the MFMA engine.
int run_ramp_fft(cbct_ctx ctx, int nu, int nv, size_t nrows, const float *in, float *out,
                 const float *cosw, const float *angles, int parker, float delta, float tan_delta,
                 hipStream_t s) {
  if (nrows == 0) return CBCT_OK;
  const int L = fft_log2n(nu);
  if (L > FFT_MAX_LOG2N)
    return run_toeplitz_mfma(ctx, nu, nv, nrows, in, out, cosw, angles, parker, delta, tan_delta, s);
  int rc = ensure_fft_tables(ctx, L);
  if (rc) return rc;
  const size_t rows_per_group = 2 * (size_t)(L >= 11 ? 1 : 2048 >> L);
  const size_t ng = (nrows + rows_per_group - 1) / rows_per_group;
  const float *tab = ctx->d_fft;
  cbct_prof_mark(ctx, PROF_FILTER, 0, s);
  switch (L) {
    case 6: rc = launch_ramp_fft<6>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    case 7: rc = launch_ramp_fft<7>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    case 8: rc = launch_ramp_fft<8>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    case 9: rc = launch_ramp_fft<9>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    case 10: rc = launch_ramp_fft<10>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    case 11: rc = launch_ramp_fft<11>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    case 12: rc = launch_ramp_fft<12>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
    default: rc = launch_ramp_fft<13>(ctx, ng, s, in, out, nrows, nu, tab, cosw, nv, angles, parker, delta, tan_delta); break;
  }
  if (rc) return rc;
  cbct_prof_mark(ctx, PROF_FILTER, 1, s);
  CBCT_LAUNCH_CHECK(ctx, "k_ramp_fft");
  return CBCT_OK;
}

int run_toeplitz(cbct_ctx ctx, int nu, size_t nrows, const float *in, float *out, hipStream_t s) {
  if (nrows == 0) return CBCT_OK;
  const float *src = in;
  if (in == out) {
    // the tiled kernel reads whole rows while writing: stage the input
    int rc = cbct_reserve(ctx, ctx->ws->filt_tmp, sizeof(float) * nrows * nu);
    if (rc) return rc;
    CBCT_HIP_CHECK(ctx, hipMemcpyAsync(ctx->ws->filt_tmp.ptr, in, sizeof(float) * nrows * nu,
                                       hipMemcpyDeviceToDevice, s));
    src = static_cast<const float *>(ctx->ws->filt_tmp.ptr);
  }
  const size_t ytiles = (nrows + TT - 1) / TT;
  if (ytiles > 0x7fffffffull)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_ramp_filter: too many rows");
  dim3 grid(ceil_div(nu, TT), (unsigned)ytiles);
  cbct_prof_mark(ctx, PROF_FILTER, 0, s);
  hipLaunchKernelGGL(k_toeplitz_f32, grid, dim3(256), 0, s, src, out, ctx->d_h, nu, nrows);
  cbct_prof_mark(ctx, PROF_FILTER, 1, s);
  CBCT_LAUNCH_CHECK(ctx, "k_toeplitz_f32");
  return CBCT_OK;
}

}  // namespace

extern "C" {

int cbct_cosine_weights(const cbct_geom *g, float *w) {
  if (!g || !w || g->nu <= 0 || g->nv <= 0) return CBCT_ERR_ARG;
  const double SAD = g->sad, SDD = g->sdd;
  for (int y = 0; y < g->nv; ++y)
    for (int x = 0; x < g->nu; ++x) {
      const double xx = ((double)x / (double)g->nu - 0.5) * (double)g->fov_u;
      const double yy = ((double)y / (double)g->nv - 0.5) * (double)g->fov_v;
      const double s = SAD * xx / SDD;
      const double v = SAD * yy / SDD;
      w[x + (size_t)g->nu * y] = (float)(SAD / std::sqrt(SAD * SAD + s * s + v * v));
    }
  return CBCT_OK;
}

int cbct_default_filter(int nu, float *F) {
  if (nu <= 0 || !F) return CBCT_ERR_ARG;
  const int n = nu + 1;
  const float A2 = (float)(0.2 * n * n);
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    const float k = (float)i;
    F[i] = (k * k < A2) ? k * A2 / (A2 - k * k) * std::exp(-A2 / (A2 - k * k)) : 0.0f;
    sum += std::fabs((double)F[i]);
  }
  const float scale = (float)(n / sum);
  for (int i = 0; i < n; ++i) F[i] *= scale;
  return CBCT_OK;
}

int cbct_filter_set(cbct_ctx ctx, int nu, const float *F) {
  if (!ctx) return CBCT_ERR_ARG;
  if (nu <= 0 || nu > 8192 || !F)
    return cbct_fail(ctx, CBCT_ERR_ARG, "set_frequency_filter: filter must hold nu+1 taps");
  int rc = cbct_set_device(ctx);
  if (rc) return rc;
  std::vector<double> h = spatial_kernel(nu, F);
  std::vector<float> hm(2 * (size_t)nu - 1);
  for (int d = -(nu - 1); d <= nu - 1; ++d) hm[d + nu - 1] = (float)h[d < 0 ? -d : d];
  // calls still reading the old tables finish first
  rc = cbct_wait_tables(ctx);
  if (rc) return rc;
  if (ctx->d_h) (void)hipFree(ctx->d_h);
  if (ctx->d_thi) (void)hipFree(ctx->d_thi);
  if (ctx->d_tlo) (void)hipFree(ctx->d_tlo);
  if (ctx->d_fft) (void)hipFree(ctx->d_fft);
  ctx->d_h = nullptr;
  ctx->d_thi = ctx->d_tlo = nullptr;
  ctx->d_fft = nullptr;
  ctx->filter_nu = 0;
  ctx->h64 = h;
  CBCT_HIP_CHECK(ctx, hipMalloc(&ctx->d_h, sizeof(float) * hm.size()));
  rc = cbct_upload(ctx, ctx->d_h, hm.data(), sizeof(float) * hm.size());
  if (rc) return rc;
  ctx->filter_nu = nu;
  return CBCT_OK;
}

int cbct_filter_engine(cbct_ctx ctx, int engine) {
  if (!ctx) return CBCT_ERR_ARG;
  if (engine < 0 || engine > 2)
    return cbct_fail(ctx, CBCT_ERR_UNSUPPORTED, "cbct_filter_engine: engine %d not available", engine);
  ctx->filter_engine = engine;
  return CBCT_OK;
}

int cbct_ramp_filter(cbct_ctx ctx, int nu, size_t nrows, const float *in, float *out, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  if (!in || !out) return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_ramp_filter: illegal array pointer provided");
  if (ctx->filter_nu == 0)
    return cbct_fail(ctx, CBCT_ERR_STATE, "Error: conebeam_backwards_projection: for _filtered_ backprojection both cosine weights and a filter must be provided");
  if (nu != ctx->filter_nu)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_ramp_filter: filter was set for nu=%d, called with nu=%d", ctx->filter_nu, nu);
  int rc = cbct_set_device(ctx);
  if (rc) return rc;
  CallScope scope(ctx, as_stream(stream));
  if (ctx->filter_engine == 2)
    return run_ramp_fft(ctx, nu, 1, nrows, in, out, nullptr, nullptr, 0, 0.0f, 0.0f,
                        as_stream(stream));
  if (ctx->filter_engine == 1)
    return run_toeplitz_mfma(ctx, nu, 1, nrows, in, out, nullptr, nullptr, 0, 0.0f, 0.0f,
                             as_stream(stream));
  return run_toeplitz(ctx, nu, nrows, in, out, as_stream(stream));
}

int cbct_fdk_prefilter(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const float *in,
                       float *out, int short_scan, int offset_correct, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "conebeam_backwards_projection";
  if (!in || !out) return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (ctx->filter_nu == 0)
    return cbct_fail(ctx, CBCT_ERR_STATE, "Error: %s: for _filtered_ backprojection both cosine weights and a filter must be provided", who);
  if (ctx->filter_nu != g->nu)
    return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: filter length %d does not match nu+1=%d", who, ctx->filter_nu + 1, g->nu + 1);
  if (nviews == 0) return CBCT_OK;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  CallScope scope(ctx, s);
  rc = ensure_cos_weights(ctx, g);
  if (rc) return rc;
  const size_t total = (size_t)g->nu * g->nv * nviews;
  const float *d_angles = ctx->d_angles + view0;
  const float delta = std::atan(g->fov_u / (2.0f * g->sdd));
  const float tan_delta = std::tan(delta);
  const size_t nrows = (size_t)g->nv * nviews;
  if (ctx->filter_engine == 2) {
    // weights fused into the transform's loads
    rc = run_ramp_fft(ctx, g->nu, g->nv, nrows, in, out, ctx->d_cosw, d_angles,
                      short_scan ? 1 : 0, delta, tan_delta, s);
    if (rc) return rc;
  } else if (ctx->filter_engine == 1) {
    // weights fused into the fp16 split, then the MFMA product into out
    rc = run_toeplitz_mfma(ctx, g->nu, g->nv, nrows, in, out, ctx->d_cosw, d_angles,
                           short_scan ? 1 : 0, delta, tan_delta, s);
    if (rc) return rc;
  } else {
  // stage 1: weights into the staging buffer
  rc = cbct_reserve(ctx, ctx->ws->filt_tmp, sizeof(float) * total);
  if (rc) return rc;
  float *tmp = static_cast<float *>(ctx->ws->filt_tmp.ptr);
  hipLaunchKernelGGL(k_cos_parker, stride_grid(total), dim3(256), 0, s, in, tmp, ctx->d_cosw, g->nu,
                     g->nv, nviews, d_angles, short_scan ? 1 : 0, delta, tan_delta);
  CBCT_LAUNCH_CHECK(ctx, "k_cos_parker");
  // stage 2: the Toeplitz filter from the staging buffer into out
  dim3 grid(ceil_div(g->nu, TT), (unsigned)((nrows + TT - 1) / TT));
  cbct_prof_mark(ctx, PROF_FILTER, 0, s);
  hipLaunchKernelGGL(k_toeplitz_f32, grid, dim3(256), 0, s, tmp, out, ctx->d_h, g->nu, nrows);
  cbct_prof_mark(ctx, PROF_FILTER, 1, s);
  CBCT_LAUNCH_CHECK(ctx, "k_toeplitz_f32");
  }
  // stage 3: full offset weights
  if (offset_correct) {
    hipLaunchKernelGGL(k_offset_weights, stride_grid(total), dim3(256), 0, s, out, g->nu, g->nv,
                       nviews, ctx->d_views_f + view0, g->fov_u, 0);
    CBCT_LAUNCH_CHECK(ctx, "k_offset_weights");
  }
  return CBCT_OK;
}

int cbct_offset_weights(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, float *proj,
                        int sqrt_variant, void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "offset_correct";
  if (!proj) return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (nviews == 0) return CBCT_OK;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  const size_t total = (size_t)g->nu * g->nv * nviews;
  CallScope scope(ctx, as_stream(stream));
  hipLaunchKernelGGL(k_offset_weights, stride_grid(total), dim3(256), 0, as_stream(stream), proj,
                     g->nu, g->nv, nviews, ctx->d_views_f + view0, g->fov_u, sqrt_variant ? 1 : 0);
  CBCT_LAUNCH_CHECK(ctx, "k_offset_weights");
  return CBCT_OK;
}

int cbct_parker_weights(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, float *proj,
                        void *stream) {
  if (!ctx) return CBCT_ERR_ARG;
  const char *who = "redundancy_correct";
  if (!proj) return cbct_fail(ctx, CBCT_ERR_ARG, "Error: %s: illegal array pointer provided", who);
  int rc = cbct_check_geom(ctx, g, who);
  if (rc) return rc;
  rc = cbct_check_views(ctx, view0, nviews, who);
  if (rc) return rc;
  if (nviews == 0) return CBCT_OK;
  rc = cbct_set_device(ctx);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  CallScope scope(ctx, s);
  // unit "cosine weights": reuse k_cos_parker with a ones plane
  const size_t plane = (size_t)g->nu * g->nv;
  rc = cbct_reserve(ctx, ctx->ws->filt_tmp, sizeof(float) * plane);
  if (rc) return rc;
  float *ones = static_cast<float *>(ctx->ws->filt_tmp.ptr);
  rc = cbct_fill(ctx, plane, ones, 1.0f, stream);
  if (rc) return rc;
  const float *d_angles = ctx->d_angles + view0;
  const float delta = std::atan(g->fov_u / (2.0f * g->sdd));
  const size_t total = plane * nviews;
  hipLaunchKernelGGL(k_cos_parker, stride_grid(total), dim3(256), 0, s, proj, proj, ones, g->nu,
                     g->nv, nviews, d_angles, 1, delta, std::tan(delta));
  CBCT_LAUNCH_CHECK(ctx, "k_cos_parker");
  return CBCT_OK;
}

}  // extern "C"
