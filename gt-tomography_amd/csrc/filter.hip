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
// That is a dense (rows x Nu) * (Nu x Nu) product -- GEMM-shaped work, so it
// runs as a tiled contraction (FP32 VALU engine, or FP16 hi/lo split MFMA
// engine with FP32 accumulation).
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
  ctx->d_h = nullptr;
  ctx->d_thi = ctx->d_tlo = nullptr;
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
  if (engine != 0 && engine != 1)
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
  if (ctx->filter_engine == 1) {
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
