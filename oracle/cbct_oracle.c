/*
 * FP64 CPU oracle for the cone-beam CT projector path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness checker: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product path (libcbct.so) never links or calls it.
 *
 * Every function restates the reference algorithm in double precision with
 * exact mathematical interpolation (no 8-bit texture weights, no fast-math).
 * Citations are to dchansen/gt-tomography.
 *
 * Array layout is first-index-fastest (Gadgetron convention):
 *   volume      V[x + nx*(y + ny*z)]
 *   projections P[u + nu*(v + nv*k)]
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nx, ny, nz;       /* volume voxels                                   */
  double lx, ly, lz;    /* volume extent in mm  (is_dims_in_mm)           */
  int nu, nv;           /* detector pixels                                 */
  double fu, fv;        /* detector extent in mm (CBCT_geometry FOV)      */
  double sad, sdd;      /* source-axis / source-detector distance in mm   */
} orc_geom;

static const double ORC_PI = 3.14159265358979323846;

/* Trilinear sample, texel coordinates (texture normalised coordinate * N - 0.5),
 * zero outside the array: xray/conebeam_projection.cu:35-36 (linear filter,
 * border address mode). */
static double tri_sample(const double *V, int nx, int ny, int nz,
                         double tx, double ty, double tz) {
  if (!(fabs(tx) < 1e9 && fabs(ty) < 1e9 && fabs(tz) < 1e9)) return 0.0;
  double fx = floor(tx), fy = floor(ty), fz = floor(tz);
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  double ax = tx - fx, ay = ty - fy, az = tz - fz;
  double acc = 0.0;
  for (int dz = 0; dz < 2; ++dz) {
    int z = iz + dz;
    if (z < 0 || z >= nz) continue;
    double wz = dz ? az : 1.0 - az;
    for (int dy = 0; dy < 2; ++dy) {
      int y = iy + dy;
      if (y < 0 || y >= ny) continue;
      double wy = dy ? ay : 1.0 - ay;
      for (int dx = 0; dx < 2; ++dx) {
        int x = ix + dx;
        if (x < 0 || x >= nx) continue;
        double wx = dx ? ax : 1.0 - ax;
        acc += wx * wy * wz * V[x + (size_t)nx * (y + (size_t)ny * z)];
      }
    }
  }
  return acc;
}

/* Bilinear sample of one projection (layered texture), zero outside:
 * xray/conebeam_projection.cu:38-39. */
static double bi_sample(const double *P, int nu, int nv, double tu, double tv) {
  if (!(fabs(tu) < 1e9 && fabs(tv) < 1e9)) return 0.0;
  double fu = floor(tu), fv = floor(tv);
  int iu = (int)fu, iv = (int)fv;
  double au = tu - fu, av = tv - fv;
  double acc = 0.0;
  for (int dv = 0; dv < 2; ++dv) {
    int v = iv + dv;
    if (v < 0 || v >= nv) continue;
    double wv = dv ? av : 1.0 - av;
    for (int du = 0; du < 2; ++du) {
      int u = iu + du;
      if (u < 0 || u >= nu) continue;
      double wu = du ? au : 1.0 - au;
      acc += wu * wv * P[u + (size_t)nu * v];
    }
  }
  return acc;
}

/* Number of samples per ray: the float product samples_per_pixel*float(Nx)
 * is passed to an int parameter (truncation): conebeam_projection.cu:625,782. */
int orc_num_samples(float samples_per_pixel, int nx) {
  float f = samples_per_pixel * (float)nx;
  return (int)f;
}

/* Per-ray geometry of the forward projector: conebeam_projection.cu:441-500.
 * Outputs the normalised start point p0, the per-sample step q (both in
 * normalised texture coordinates) and the step length in mm. */
static void fp_ray(const orc_geom *g, double angle_deg, double off_u, double off_v,
                   int i, int j, int n, double p0[3], double q[3], double *step_mm) {
  double th = angle_deg * (ORC_PI / 180.0);
  double c = cos(th), s = sin(th);
  /* startPoint = R(theta) * (0,-SAD,0)   (float3x3.h:54-63) */
  double S[3] = { s * g->sad, -c * g->sad, 0.0 };
  double pu = ((i + 0.5) / g->nu - 0.5) * g->fu + off_u;
  double pv = ((j + 0.5) / g->nv - 0.5) * g->fv + off_v;
  double add = g->sdd - g->sad;
  double E[3] = { c * pu - s * add, s * pu + c * add, pv };
  double d[3] = { E[0] - S[0], E[1] - S[1], E[2] - S[2] };
  double L[3] = { g->lx, g->ly, g->lz };
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    double tA = (L[a] / 2 - S[a]) / d[a];
    double tB = (-L[a] / 2 - S[a]) / d[a];
    lo[a] = fmin(tB, tA);
    hi[a] = fmax(tB, tA);
  }
  double a1 = fmax(fmax(fmax(lo[0], lo[1]), lo[2]), 0.0);
  double a2 = fmin(fmin(fmin(hi[0], hi[1]), hi[2]), 1.0);
  double seg[3] = { (a2 - a1) * d[0], (a2 - a1) * d[1], (a2 - a1) * d[2] };
  *step_mm = sqrt(seg[0] * seg[0] + seg[1] * seg[1] + seg[2] * seg[2]) / n;
  for (int a = 0; a < 3; ++a) {
    p0[a] = (S[a] + a1 * d[a]) / L[a] + 0.5;
    q[a] = d[a] / L[a] * ((a2 - a1) / n);
  }
}

/* Forward projection, conebeam_projection.cu:412-528 (kernel) and :530-635
 * (device wrapper).  angles in degrees, offsets interleaved (u,v) per view. */
void orc_fp(const orc_geom *g, const float *angles_deg, const float *offsets, int nviews,
            const double *vol, double *proj, float samples_per_pixel, int accumulate) {
  int n = orc_num_samples(samples_per_pixel, g->nx);
  const size_t npix = (size_t)g->nu * g->nv;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int k = 0; k < nviews; ++k) {
    for (int j = 0; j < g->nv; ++j) {
      for (int i = 0; i < g->nu; ++i) {
        double p0[3], q[3], step;
        fp_ray(g, angles_deg[k], offsets[2 * k], offsets[2 * k + 1], i, j, n, p0, q, &step);
        double acc = 0.0;
        for (int m = 0; m < n; ++m) {
          double tx = (p0[0] + q[0] * m) * g->nx - 0.5;
          double ty = (p0[1] + q[1] * m) * g->ny - 0.5;
          double tz = (p0[2] + q[2] * m) * g->nz - 0.5;
          acc += tri_sample(vol, g->nx, g->ny, g->nz, tx, ty, tz);
        }
        size_t idx = i + (size_t)g->nu * j + npix * k;
        proj[idx] = (accumulate ? proj[idx] : 0.0) + acc * step;
      }
    }
  }
}

/* orc_fp evaluated on a list of rays only (view k, column i, row j packed as
 * rays[3r..3r+2]): the same ray setup and sample sum as orc_fp
 * (conebeam_projection.cu:412-528), so out[r] equals orc_fp's value at that
 * detector pixel.  Lets full-size geometries be checked on a subsample. */
void orc_fp_rays(const orc_geom *g, const float *angles_deg, const float *offsets,
                 const double *vol, float samples_per_pixel, size_t nrays, const int *rays,
                 double *out) {
  int n = orc_num_samples(samples_per_pixel, g->nx);
#pragma omp parallel for schedule(dynamic, 16)
  for (long r = 0; r < (long)nrays; ++r) {
    int k = rays[3 * r], i = rays[3 * r + 1], j = rays[3 * r + 2];
    double p0[3], q[3], step;
    fp_ray(g, angles_deg[k], offsets[2 * k], offsets[2 * k + 1], i, j, n, p0, q, &step);
    double acc = 0.0;
    for (int m = 0; m < n; ++m) {
      double tx = (p0[0] + q[0] * m) * g->nx - 0.5;
      double ty = (p0[1] + q[1] * m) * g->ny - 0.5;
      double tz = (p0[2] + q[2] * m) * g->nz - 0.5;
      acc += tri_sample(vol, g->nx, g->ny, g->nz, tx, ty, tz);
    }
    out[r] = acc * step;
  }
}

/* Exact transpose of orc_fp (ray-driven scatter).  vol += FP^T(proj).
 * Not a reference function: it exists so that <FP x, y> = <x, FP^T y> can be
 * checked (the reference FP/BP pair is not adjoint). */
void orc_fp_transpose(const orc_geom *g, const float *angles_deg, const float *offsets,
                      int nviews, const double *proj, double *vol, float samples_per_pixel,
                      int accumulate) {
  int n = orc_num_samples(samples_per_pixel, g->nx);
  const size_t npix = (size_t)g->nu * g->nv;
  const size_t nvox = (size_t)g->nx * g->ny * g->nz;
  if (!accumulate) memset(vol, 0, nvox * sizeof(double));
  for (int k = 0; k < nviews; ++k)
    for (int j = 0; j < g->nv; ++j)
      for (int i = 0; i < g->nu; ++i) {
        double p0[3], q[3], step;
        fp_ray(g, angles_deg[k], offsets[2 * k], offsets[2 * k + 1], i, j, n, p0, q, &step);
        double y = proj[i + (size_t)g->nu * j + npix * k] * step;
        for (int m = 0; m < n; ++m) {
          double tx = (p0[0] + q[0] * m) * g->nx - 0.5;
          double ty = (p0[1] + q[1] * m) * g->ny - 0.5;
          double tz = (p0[2] + q[2] * m) * g->nz - 0.5;
          if (!(fabs(tx) < 1e9 && fabs(ty) < 1e9 && fabs(tz) < 1e9)) continue;
          double fx = floor(tx), fy = floor(ty), fz = floor(tz);
          int ix = (int)fx, iy = (int)fy, iz = (int)fz;
          double ax = tx - fx, ay = ty - fy, az = tz - fz;
          for (int dz = 0; dz < 2; ++dz) {
            int z = iz + dz;
            if (z < 0 || z >= g->nz) continue;
            double wz = dz ? az : 1.0 - az;
            for (int dy = 0; dy < 2; ++dy) {
              int yy = iy + dy;
              if (yy < 0 || yy >= g->ny) continue;
              double wy = dy ? ay : 1.0 - ay;
              for (int dx = 0; dx < 2; ++dx) {
                int x = ix + dx;
                if (x < 0 || x >= g->nx) continue;
                double wx = dx ? ax : 1.0 - ax;
                vol[x + (size_t)g->nx * (yy + (size_t)g->ny * z)] += wx * wy * wz * y;
              }
            }
          }
        }
      }
}

/* Detector coordinate of a voxel centre for one view:
 * conebeam_projection.cu:976-1001.  Returns texel coordinates (tu, tv) and the
 * rotated coordinate q.  */
static void bp_project(const orc_geom *g, double angle_deg, double off_u, double off_v,
                       const double pos[3], double *tu, double *tv) {
  double th = angle_deg * (ORC_PI / 180.0);
  double c = cos(th), s = sin(th);
  /* inverseRotation = R(-theta): rows (c, s, 0), (-s, c, 0), (0, 0, 1) */
  double qx = c * pos[0] + s * pos[1];
  double qy = -s * pos[0] + c * pos[1];
  double qz = pos[2];
  double dy = qy + g->sad;            /* dir = q - (0,-SAD,0) */
  double dx = qx / dy, dz = qz / dy;  /* dir /= dir.y        */
  double ex = g->sdd * dx;            /* endPoint = start + dir*SDD */
  double ez = g->sdd * dz;
  double eu = ex - off_u, ev = ez - off_v;
  *tu = (eu / g->fu + 0.5) * g->nu - 0.5;
  *tv = (ev / g->fv + 0.5) * g->nv - 0.5;
}

/* FDK weight 1/U^2, transcribed literally (hypot/atan2/sin form):
 * conebeam_projection.cu:1009-1024. */
static double fbp_weight(const orc_geom *g, double angle_deg, const double pos[3]) {
  double beta = angle_deg * (ORC_PI / 180.0);
  double r = hypot(pos[0], pos[1]);
  double phi = atan2(pos[1], pos[0]);
  double ym = r * sin(beta - phi);
  double U = (g->sad + ym) / g->sad;
  return 1.0 / (U * U);
}

/* Voxel-driven backprojection, conebeam_projection.cu:922-1037.
 * fbp != 0 selects the FDK-weighted template instance. */
void orc_bp(const orc_geom *g, const float *angles_deg, const float *offsets, int nviews,
            const double *proj, double *vol, int fbp, int accumulate) {
  const size_t npix = (size_t)g->nu * g->nv;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        double pos[3] = { ((x + 0.5) / g->nx - 0.5) * g->lx,
                          ((y + 0.5) / g->ny - 0.5) * g->ly,
                          ((z + 0.5) / g->nz - 0.5) * g->lz };
        double acc = 0.0;
        for (int k = 0; k < nviews; ++k) {
          double tu, tv;
          bp_project(g, angles_deg[k], offsets[2 * k], offsets[2 * k + 1], pos, &tu, &tv);
          double w = fbp ? fbp_weight(g, angles_deg[k], pos) : 1.0;
          acc += w * bi_sample(proj + npix * k, g->nu, g->nv, tu, tv);
        }
        size_t idx = x + (size_t)g->nx * (y + (size_t)g->ny * z);
        vol[idx] = (accumulate ? vol[idx] : 0.0) + acc;
      }
}

/* orc_bp evaluated at a list of voxels only (linear indices x + nx*(y + ny*z)):
 * the same projection, weight and sum over views as orc_bp
 * (conebeam_projection.cu:922-1037), so out[r] equals orc_bp's value there. */
void orc_bp_voxels(const orc_geom *g, const float *angles_deg, const float *offsets, int nviews,
                   const double *proj, int fbp, size_t nvox, const long long *voxels,
                   double *out) {
  const size_t npix = (size_t)g->nu * g->nv;
#pragma omp parallel for schedule(dynamic, 64)
  for (long r = 0; r < (long)nvox; ++r) {
    long long idx = voxels[r];
    int x = (int)(idx % g->nx);
    int y = (int)((idx / g->nx) % g->ny);
    int z = (int)(idx / ((long long)g->nx * g->ny));
    double pos[3] = { ((x + 0.5) / g->nx - 0.5) * g->lx,
                      ((y + 0.5) / g->ny - 0.5) * g->ly,
                      ((z + 0.5) / g->nz - 0.5) * g->lz };
    double acc = 0.0;
    for (int k = 0; k < nviews; ++k) {
      double tu, tv;
      bp_project(g, angles_deg[k], offsets[2 * k], offsets[2 * k + 1], pos, &tu, &tv);
      double w = fbp ? fbp_weight(g, angles_deg[k], pos) : 1.0;
      acc += w * bi_sample(proj + npix * k, g->nu, g->nv, tu, tv);
    }
    out[r] = acc;
  }
}

/* Exact transpose of orc_bp (voxel-driven splat into the projections). */
void orc_bp_transpose(const orc_geom *g, const float *angles_deg, const float *offsets,
                      int nviews, const double *vol, double *proj, int fbp, int accumulate) {
  const size_t npix = (size_t)g->nu * g->nv;
  if (!accumulate) memset(proj, 0, npix * nviews * sizeof(double));
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        double pos[3] = { ((x + 0.5) / g->nx - 0.5) * g->lx,
                          ((y + 0.5) / g->ny - 0.5) * g->ly,
                          ((z + 0.5) / g->nz - 0.5) * g->lz };
        double val = vol[x + (size_t)g->nx * (y + (size_t)g->ny * z)];
        for (int k = 0; k < nviews; ++k) {
          double tu, tv;
          bp_project(g, angles_deg[k], offsets[2 * k], offsets[2 * k + 1], pos, &tu, &tv);
          if (!(fabs(tu) < 1e9 && fabs(tv) < 1e9)) continue;
          double w = fbp ? fbp_weight(g, angles_deg[k], pos) : 1.0;
          double fu = floor(tu), fv = floor(tv);
          int iu = (int)fu, iv = (int)fv;
          double au = tu - fu, av = tv - fv;
          for (int dv = 0; dv < 2; ++dv) {
            int v = iv + dv;
            if (v < 0 || v >= g->nv) continue;
            double wv = dv ? av : 1.0 - av;
            for (int du = 0; du < 2; ++du) {
              int u = iu + du;
              if (u < 0 || u >= g->nu) continue;
              double wu = du ? au : 1.0 - au;
              proj[u + (size_t)g->nu * v + npix * k] += w * wu * wv * val;
            }
          }
        }
      }
}

/* Space-carving mask, conebeam_spacecarver_kernel (conebeam_projection.cu:822-921):
 * the voxel centre with its z coordinate negated (:854) is projected onto
 * every view; the projection is sampled at the nearest texel (point filter,
 * normalised coordinates, zero outside the detector: the mask texture,
 * :41-42); views whose value is below `limit` and not zero are counted;
 * mask = count < 3.  ambiguous[v] = 1 when, for some view, the voxel's
 * detector coordinate lies within 1e-3 texel of a texel edge (FP32 rounding
 * may pick the other texel there). */
void orc_spacecarver(const orc_geom *g, const float *angles_deg, const float *offsets, int nviews,
                     const double *proj, double limit, unsigned char *mask,
                     unsigned char *ambiguous) {
  const size_t npix = (size_t)g->nu * g->nv;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        const double px = ((x + 0.5) / g->nx - 0.5) * g->lx;
        const double py = ((y + 0.5) / g->ny - 0.5) * g->ly;
        const double pz = -(((z + 0.5) / g->nz - 0.5) * g->lz);
        int count = 0, amb = 0;
        for (int k = 0; k < nviews; ++k) {
          const double th = angles_deg[k] * (ORC_PI / 180.0);
          const double c = cos(th), sn = sin(th);
          const double qx = c * px + sn * py, qy = -sn * px + c * py;
          const double t = g->sdd / (qy + g->sad);
          const double eu = qx * t - offsets[2 * k], ev = pz * t - offsets[2 * k + 1];
          const double tu = (eu / g->fu + 0.5) * g->nu, tv = (ev / g->fv + 0.5) * g->nv;
          const double fu = floor(tu), fv = floor(tv);
          if (fmin(fmin(tu - fu, fu + 1.0 - tu), fmin(tv - fv, fv + 1.0 - tv)) < 1e-3) amb = 1;
          double val = 0.0;
          if (fu >= 0 && fu < g->nu && fv >= 0 && fv < g->nv)
            val = proj[(size_t)fu + (size_t)g->nu * (size_t)fv + npix * k];
          if (val < limit && val != 0.0) ++count;
        }
        const size_t idx = x + (size_t)g->nx * (y + (size_t)g->ny * z);
        mask[idx] = count < 3;
        ambiguous[idx] = (unsigned char)amb;
      }
}

/* Cosine weights C[u,v] = SAD / sqrt(SAD^2 + s^2 + t^2), sampled at u/Nu (not
 * pixel centres): xray/hoCuConebeamProjectionOperator.cpp:64-102. */
void orc_cosine_weights(const orc_geom *g, double *w) {
  for (int y = 0; y < g->nv; ++y)
    for (int x = 0; x < g->nu; ++x) {
      double xx = ((double)x / (double)g->nu - 0.5) * g->fu;
      double yy = ((double)y / (double)g->nv - 0.5) * g->fv;
      double s = g->sad * xx / g->sdd;
      double v = g->sad * yy / g->sdd;
      w[x + (size_t)g->nu * y] = g->sad / sqrt(g->sad * g->sad + s * s + v * v);
    }
}

/* Default (Guo et al.) frequency filter, Nu+1 taps normalised to sum Nu+1:
 * xray/hoCuConebeamProjectionOperator.cpp:26-62. */
void orc_guo_filter(int nu, double *F) {
  int n = nu + 1;
  double A2 = 0.2 * n * n;
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    double k = i;
    F[i] = (k * k < A2) ? k * A2 / (A2 - k * k) * exp(-A2 / (A2 - k * k)) : 0.0;
    sum += fabs(F[i]);
  }
  for (int i = 0; i < n; ++i) F[i] *= n / sum;
}

/* Spatial kernel of the frequency filter as applied by the pad -> R2C -> *F ->
 * C2R -> /(2Nu) -> crop chain (conebeam_projection.cu:87-139, :1504-1511):
 * h[m] = (1/2Nu) * (F0 + 2 sum_{k=1}^{Nu-1} Fk cos(pi k m / Nu) + F_Nu (-1)^m),
 * m = 0..Nu-1 (the kernel is even; |i-j| < Nu so no wrap-around). */
void orc_filter_kernel(int nu, const double *F, double *h) {
  double inv = 1.0 / (2.0 * nu);
  for (int m = 0; m < nu; ++m) {
    double acc = F[0] + F[nu] * ((m & 1) ? -1.0 : 1.0);
    for (int k = 1; k < nu; ++k) acc += 2.0 * F[k] * cos(ORC_PI * (double)k * m / nu);
    h[m] = acc * inv;
  }
}

/* Filter every detector row in place: y[i] = sum_j h[|i-j|] x[j]. */
void orc_filter_rows(int nu, size_t nrows, const double *h, double *data) {
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * nu);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nrows; ++r) {
      double *row = data + (size_t)r * nu;
      for (int i = 0; i < nu; ++i) {
        double acc = 0.0;
        for (int j = 0; j < nu; ++j) acc += h[abs(i - j)] * row[j];
        tmp[i] = acc;
      }
      memcpy(row, tmp, sizeof(double) * nu);
    }
    free(tmp);
  }
}

/* Parker-type short-scan weights (Wesarg q=0.1): conebeam_projection.cu:149-194. */
static double parker_S(double b) {
  if (b <= -0.5) return 0.0;
  if (b < 0.5) return 0.5 * (1.0 + sin(ORC_PI * b));
  return 1.0;
}
/* b(alpha, delta) = q * (2 (delta - alpha) + eps); the reference fixes q = 0.1
 * and notes that q = 1 gives the conventional Parker weights. */
static double parker_b(double alpha, double delta, double q, double eps) {
  return q * (2.0 * (delta - alpha) + eps);
}

/* Weight of the ray with source angle beta and fan angle alpha (radians). */
double orc_parker_weight(double beta, double alpha, double delta, double q, double eps) {
  return 0.5 * (parker_S(beta / parker_b(alpha, delta, q, eps) - 0.5) +
                parker_S((beta + 2.0 * (alpha - delta) - eps) / parker_b(alpha, delta, q, eps) + 0.5) -
                parker_S((beta - ORC_PI + 2.0 * alpha) / parker_b(-alpha, delta, q, eps) - 0.5) -
                parker_S((beta - ORC_PI - 2.0 * delta - eps) / parker_b(-alpha, delta, q, eps) + 0.5));
}

void orc_parker_weights(const orc_geom *g, const float *angles_deg, int nviews, double *proj) {
  double delta = atan(g->fu / (2.0 * g->sdd));
  double tan_delta = tan(delta);
  for (int k = 0; k < nviews; ++k) {
    double beta = angles_deg[k] * (ORC_PI / 180.0);
    for (int v = 0; v < g->nv; ++v)
      for (int u = 0; u < g->nu; ++u) {
        double alpha = -atan(((double)u / g->nu - 0.5) * 2.0 * tan_delta);
        double w = orc_parker_weight(beta, alpha, delta, 0.1, 0.001);
        proj[u + (size_t)g->nu * (v + (size_t)g->nv * k)] *= w;
      }
  }
}

/* Half-fan (Wang) offset weights.  sqrt_variant = 0: full weight
 * (conebeam_projection.cu:233-262); 1: square-root weight (:303-331). */
void orc_offset_weights(const orc_geom *g, const float *offsets, int nviews, int sqrt_variant,
                        double *proj) {
  for (int k = 0; k < nviews; ++k) {
    double ou = offsets[2 * k];
    double sign = ou / fabs(ou);   /* NaN for a zero offset: every test below is false */
    double omega = g->fu / 2.0 - fabs(ou);
    for (int v = 0; v < g->nv; ++v)
      for (int u = 0; u < g->nu; ++u) {
        double *p = &proj[u + (size_t)g->nu * (v + (size_t)g->nv * k)];
        if (!sqrt_variant) {
          double t = (g->fu * ((double)u / g->nu - 0.5) + ou) * sign;
          if (fabs(t) <= omega) {
            double sw = sin(ORC_PI * (t + omega) / (4.0 * omega));
            *p *= sw * sw;
          }
          if (t < -omega) *p *= 0.0;
        } else {
          double t = (g->fu * ((u + 0.5) / g->nu - 0.5) + ou) * sign;
          if (fabs(t) <= fabs(omega)) *p *= sin(ORC_PI * (t + omega) / (4.0 * omega));
        }
      }
  }
}

/* Isotropic TV primal-dual step on one 3-D volume:
 * operators/cuTVPrimalDualOperator.cu:10-106.
 * out += weight * sum_c [ p_c(v - e_c) - p_c(v) ],
 * p = Pi_{|.|<=1}( (weight*sigma)/(1+sigma*alpha) * D x ), D periodic forward
 * difference. */
void orc_tv_primal_dual(int nx, int ny, int nz, const double *x, double *out, double sigma,
                        double alpha, double weight, int accumulate) {
  size_t n = (size_t)nx * ny * nz;
  double *p = (double *)malloc(sizeof(double) * 3 * n);
  int dims[3] = { nx, ny, nz };
  double scale = weight * sigma / (1.0 + sigma * alpha);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int xx = 0; xx < nx; ++xx) {
        int co[3] = { xx, y, z };
        size_t idx = xx + (size_t)nx * (y + (size_t)ny * z);
        double r[3];
        for (int c = 0; c < 3; ++c) {
          int co2[3] = { co[0], co[1], co[2] };
          co2[c] = (co[c] + dims[c] + 1) % dims[c];
          r[c] = x[co2[0] + (size_t)nx * (co2[1] + (size_t)ny * co2[2])] - x[idx];
          r[c] *= scale;
        }
        double nrm = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double d = nrm > 1.0 ? nrm : 1.0;
        for (int c = 0; c < 3; ++c) p[idx + c * n] = r[c] / d;
      }
  if (!accumulate) memset(out, 0, sizeof(double) * n);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int xx = 0; xx < nx; ++xx) {
        int co[3] = { xx, y, z };
        size_t idx = xx + (size_t)nx * (y + (size_t)ny * z);
        double res = 0.0;
        for (int c = 0; c < 3; ++c) {
          int co2[3] = { co[0], co[1], co[2] };
          co2[c] = (co[c] + dims[c] - 1) % dims[c];
          res += p[co2[0] + (size_t)nx * (co2[1] + (size_t)ny * co2[2]) + c * n] - p[idx + c * n];
        }
        out[idx] += res * weight;
      }
  free(p);
}
