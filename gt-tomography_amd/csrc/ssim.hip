// Multi-scale SSIM of a volume against a reference (xray/mssim.cpp:14-121).
//
//   S = prod_{i < scales} CS(sigma 2^i) * L(sigma 2^(scales-1)),   result = mean(S)
//   CS = (2 cov_xy + C2) / (var_x + var_y + C2)           (mssim.cpp:14-74)
//   L  = (2 mu_x mu_y + C1) / (mu_x^2 + mu_y^2 + C1)      (mssim.cpp:88-117)
//   C1 = (0.01 R)^2, C2 = (0.03 R)^2, R = max(reference) - min(reference)
//
// The local moments are Gaussian-weighted means (the reference calls
// Gadgetron's cuGaussianFilterOperator<float,3>, which is not part of the
// reference sources).  The filter here is separable with taps
// exp(-k^2 / (2 sigma^2)), |k| <= int(3.5 sigma + 0.5), normalised to sum 1,
// and half-sample symmetric borders (index -1 reads 0, n reads n-1).
//
// Layout of one scale (x fastest, n = nx*ny*nz):
//   pass x:  image, reference -> F[5]: the five fields x, y, x^2, y^2, xy,
//            filtered along x (products formed on the fly);
//   pass y:  F[5] -> G[5];
//   pass z:  G[5] -> S, with the CS (and, at the last scale, L) combine fused.
// Image and reference are shifted by the midpoint of the reference's range
// before the products are formed: the covariances do not change, and the
// FP32 sums then lose far less to cancellation than with raw values.
#include <cmath>
#include <vector>

#include "cbct_internal.h"

namespace {

constexpr int SSIM_BLOCK = 256;
constexpr int SSIM_MAX_RADIUS = 4096;
constexpr int MM_BLOCK = 256;
constexpr int MM_MAX_BLOCKS = 1024;

__device__ __forceinline__ int reflect_index(int i, int m) {
  if ((unsigned)i < (unsigned)m) return i;
  const int period = 2 * m;
  i %= period;
  if (i < 0) i += period;
  return i < m ? i : period - 1 - i;
}

struct SsimPass {
  int nx, ny, nz, radius;
  const float *taps;   // 2*radius+1 weights
  float shift, c1, c2;
  int first, lum;      // pass z: S starts at 1; multiply the luminance term
};

// MODE 0: in = (image, reference) -> out = F[5]
// MODE 1: in = F[5]               -> out = G[5]
// MODE 2: in = G[5]               -> out = S (combine)
template <int AXIS, int MODE>
__global__ void __launch_bounds__(SSIM_BLOCK) k_ssim_pass(SsimPass a, size_t n,
                                                          const float *__restrict__ in0,
                                                          const float *__restrict__ in1,
                                                          float *__restrict__ out) {
  const size_t plane = (size_t)a.nx * a.ny;
  for (size_t t = blockIdx.x * (size_t)SSIM_BLOCK + threadIdx.x; t < n;
       t += (size_t)gridDim.x * SSIM_BLOCK) {
    const size_t z = t / plane;
    const size_t rem = t - z * plane;
    const size_t y = rem / a.nx;
    const size_t x = rem - y * a.nx;
    int q, m;
    size_t st;
    if (AXIS == 0) {
      q = (int)x; m = a.nx; st = 1;
    } else if (AXIS == 1) {
      q = (int)y; m = a.ny; st = a.nx;
    } else {
      q = (int)z; m = a.nz; st = plane;
    }
    const size_t base = t - (size_t)q * st;
    float acc[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = -a.radius; k <= a.radius; ++k) {
      const float w = a.taps[k + a.radius];
      const size_t idx = base + (size_t)reflect_index(q + k, m) * st;
      if (MODE == 0) {
        const float xv = in0[idx] - a.shift;
        const float yv = in1[idx] - a.shift;
        acc[0] += w * xv;
        acc[1] += w * yv;
        acc[2] += w * (xv * xv);
        acc[3] += w * (yv * yv);
        acc[4] += w * (xv * yv);
      } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) acc[f] += w * in0[idx + f * n];
      }
    }
    if (MODE < 2) {
#pragma unroll
      for (int f = 0; f < 5; ++f) out[t + f * n] = acc[f];
    } else {
      const float mx = acc[0], my = acc[1];          // shifted means
      const float vx = acc[2] - mx * mx;
      const float vy = acc[3] - my * my;
      const float cxy = acc[4] - mx * my;
      float s = (2.0f * cxy + a.c2) / (vx + vy + a.c2);
      if (a.lum) {
        const float ux = mx + a.shift, uy = my + a.shift;
        s *= (2.0f * ux * uy + a.c1) / (ux * ux + uy * uy + a.c1);
      }
      out[t] = a.first ? s : out[t] * s;
    }
  }
}

__global__ void __launch_bounds__(MM_BLOCK) k_minmax_partial(size_t n, const float *__restrict__ v,
                                                             float *__restrict__ part) {
  float lo = INFINITY, hi = -INFINITY;
  for (size_t t = blockIdx.x * (size_t)MM_BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * MM_BLOCK) {
    const float x = v[t];
    lo = fminf(lo, x);
    hi = fmaxf(hi, x);
  }
  __shared__ float slo[MM_BLOCK], shi[MM_BLOCK];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int s = MM_BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      slo[threadIdx.x] = fminf(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = fmaxf(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = slo[0];
    part[2 * blockIdx.x + 1] = shi[0];
  }
}

__global__ void __launch_bounds__(MM_BLOCK) k_minmax_final(int nparts, const float *__restrict__ part,
                                                           float *__restrict__ out) {
  float lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < nparts; i += MM_BLOCK) {
    lo = fminf(lo, part[2 * i]);
    hi = fmaxf(hi, part[2 * i + 1]);
  }
  __shared__ float slo[MM_BLOCK], shi[MM_BLOCK];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int s = MM_BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      slo[threadIdx.x] = fminf(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = fmaxf(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = slo[0];
    out[1] = shi[0];
  }
}

inline dim3 pass_grid(size_t n) {
  size_t b = (n + SSIM_BLOCK - 1) / SSIM_BLOCK;
  if (b > 262144) b = 262144;
  if (b == 0) b = 1;
  return dim3((unsigned)b);
}

inline int tap_radius(float sigma) {
  return sigma > 0.0f ? (int)(3.5 * (double)sigma + 0.5) : 0;
}

void gaussian_taps(float sigma, int r, std::vector<float> &out) {
  std::vector<double> w(2 * r + 1);
  double sum = 0.0;
  for (int k = -r; k <= r; ++k) {
    w[k + r] = r ? std::exp(-0.5 * (double)k * k / ((double)sigma * sigma)) : 1.0;
    sum += w[k + r];
  }
  for (double v : w) out.push_back((float)(v / sum));
}

}  // namespace

extern "C" int cbct_mssim(cbct_ctx ctx, int nx, int ny, int nz, const float *image,
                          const float *reference, float sigma_x, float sigma_y, float sigma_z,
                          int scales, double *out, void *st) {
  if (!ctx) return CBCT_ERR_ARG;
  if (!image || !reference || !out)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_mssim: illegal array pointer provided");
  if (nx <= 0 || ny <= 0 || nz <= 0)
    return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_mssim: image dimensions must be positive");
  if (scales < 1 || scales > 16) return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_mssim: scales must be in [1, 16]");
  const float sig[3] = {sigma_x, sigma_y, sigma_z};
  for (float s : sig)
    if (!(s >= 0.0f) || !std::isfinite(s))
      return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_mssim: sigma must be finite and non-negative");
  // per scale and axis: sigma * 2^i in float, as mssim.cpp:83,89 scale it
  std::vector<float> taps;
  std::vector<size_t> tap_off;
  std::vector<int> radius;
  for (int i = 0; i < scales; ++i) {
    const float f = (float)std::pow(2.0f, (float)i);
    for (int ax = 0; ax < 3; ++ax) {
      const float s = sig[ax] * f;
      const int r = tap_radius(s);
      if (r > SSIM_MAX_RADIUS)
        return cbct_fail(ctx, CBCT_ERR_ARG, "cbct_mssim: sigma %g at scale %d needs more than %d taps a side",
                         (double)s, i, SSIM_MAX_RADIUS);
      tap_off.push_back(taps.size());
      radius.push_back(r);
      gaussian_taps(s, r, taps);
    }
  }
  int rc = cbct_set_device(ctx);
  if (rc) return rc;
  CallScope scope(ctx, as_stream(st));
  const size_t n = (size_t)nx * ny * nz;
  const size_t fields = 11 * n;   // F[5], G[5], S
  const size_t bytes = sizeof(float) * (fields + taps.size() + 2 * MM_MAX_BLOCKS + 64);
  rc = cbct_reserve(ctx, ctx->ws->ssim_ws, bytes);
  if (rc) return rc;
  float *F = static_cast<float *>(ctx->ws->ssim_ws.ptr);
  float *G = F + 5 * n;
  float *S = G + 5 * n;
  float *d_taps = S + n;
  float *part = d_taps + taps.size();
  float *range = part + 2 * MM_MAX_BLOCKS;
  hipStream_t s = as_stream(st);
  CBCT_HIP_CHECK(ctx, hipMemcpyAsync(d_taps, taps.data(), sizeof(float) * taps.size(),
                                     hipMemcpyHostToDevice, s));
  // data range of the reference (mssim.cpp:19, :93)
  size_t mmb = (n + MM_BLOCK - 1) / MM_BLOCK;
  if (mmb > MM_MAX_BLOCKS) mmb = MM_MAX_BLOCKS;
  hipLaunchKernelGGL(k_minmax_partial, dim3((unsigned)mmb), dim3(MM_BLOCK), 0, s, n, reference, part);
  CBCT_LAUNCH_CHECK(ctx, "k_minmax_partial");
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(MM_BLOCK), 0, s, (int)mmb, part, range);
  CBCT_LAUNCH_CHECK(ctx, "k_minmax_final");
  float *h_range = reinterpret_cast<float *>(ctx->h_red);
  CBCT_HIP_CHECK(ctx, hipMemcpyAsync(h_range, range, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
  CBCT_HIP_CHECK(ctx, hipStreamSynchronize(s));
  const float lo = h_range[0], hi = h_range[1];
  const float data_range = hi - lo;
  // std::pow(float, int) promotes to double (mssim.cpp:65, :108)
  const float c1 = (float)std::pow((double)(0.01f * data_range), 2.0);
  const float c2 = (float)std::pow((double)(0.03f * data_range), 2.0);
  SsimPass a;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.shift = 0.5f * lo + 0.5f * hi;
  a.c1 = c1;
  a.c2 = c2;
  const dim3 grid = pass_grid(n);
  for (int i = 0; i < scales; ++i) {
    a.first = i == 0;
    a.lum = i == scales - 1;
    a.radius = radius[3 * i + 0];
    a.taps = d_taps + tap_off[3 * i + 0];
    hipLaunchKernelGGL((k_ssim_pass<0, 0>), grid, dim3(SSIM_BLOCK), 0, s, a, n, image, reference, F);
    CBCT_LAUNCH_CHECK(ctx, "k_ssim_pass<x>");
    a.radius = radius[3 * i + 1];
    a.taps = d_taps + tap_off[3 * i + 1];
    hipLaunchKernelGGL((k_ssim_pass<1, 1>), grid, dim3(SSIM_BLOCK), 0, s, a, n, F, nullptr, G);
    CBCT_LAUNCH_CHECK(ctx, "k_ssim_pass<y>");
    a.radius = radius[3 * i + 2];
    a.taps = d_taps + tap_off[3 * i + 2];
    hipLaunchKernelGGL((k_ssim_pass<2, 2>), grid, dim3(SSIM_BLOCK), 0, s, a, n, G, nullptr, S);
    CBCT_LAUNCH_CHECK(ctx, "k_ssim_pass<z>");
  }
  double sum = 0.0;
  rc = cbct_reduce_f64(ctx, CBCT_SUM, n, S, nullptr, &sum, st);
  if (rc) return rc;
  *out = sum / (double)n;
  return CBCT_OK;
}
