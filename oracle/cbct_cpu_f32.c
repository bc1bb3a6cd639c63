/*
 * FP32 OpenMP CPU restatement of the cone-beam projector pair.
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU baseline timed beside the GPU
 * path (the reference has no CPU projector; its projector is GPU-only).  It
 * follows the same sample positions, interpolation and operation order as the
 * FP64 oracle (cbct_oracle.c) but in single precision:
 *   forward projection  xray/conebeam_projection.cu:412-528
 *   backprojection      xray/conebeam_projection.cu:922-1037
 */
#include <math.h>
#include <stddef.h>

typedef struct {
  int nx, ny, nz;
  double lx, ly, lz;
  int nu, nv;
  double fu, fv;
  double sad, sdd;
} orc_geom;

static inline float tri_f(const float *V, int nx, int ny, int nz, float tx, float ty, float tz) {
  if (!(fabsf(tx) < 1e9f && fabsf(ty) < 1e9f && fabsf(tz) < 1e9f)) return 0.0f;
  float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  float ax = tx - fx, ay = ty - fy, az = tz - fz;
  float acc = 0.0f;
  for (int dz = 0; dz < 2; ++dz) {
    int z = iz + dz;
    if (z < 0 || z >= nz) continue;
    float wz = dz ? az : 1.0f - az;
    for (int dy = 0; dy < 2; ++dy) {
      int y = iy + dy;
      if (y < 0 || y >= ny) continue;
      float wy = dy ? ay : 1.0f - ay;
      const float *row = V + (size_t)nx * (y + (size_t)ny * z);
      float a = (ix >= 0 && ix < nx) ? row[ix] : 0.0f;
      float b = (ix + 1 >= 0 && ix + 1 < nx) ? row[ix + 1] : 0.0f;
      acc += wz * wy * (a + ax * (b - a));
    }
  }
  return acc;
}

/* views [0, nviews) of the forward projection, FP32. */
void cpu_fp_f32(const orc_geom *g, const float *angles_deg, const float *offsets, int nviews,
                const float *vol, float *proj, float samples_per_pixel) {
  const int n = (int)(samples_per_pixel * (float)g->nx);
  const size_t npix = (size_t)g->nu * g->nv;
  const float Lx = (float)g->lx, Ly = (float)g->ly, Lz = (float)g->lz;
  const float SAD = (float)g->sad, ADD = (float)(g->sdd - g->sad);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int k = 0; k < nviews; ++k)
    for (int j = 0; j < g->nv; ++j) {
      float th = angles_deg[k] * (3.14159265358979f / 180.0f);
      float c = cosf(th), s = sinf(th);
      float S0 = s * SAD, S1 = -c * SAD;
      float pv = ((j + 0.5f) / g->nv - 0.5f) * (float)g->fv + offsets[2 * k + 1];
      for (int i = 0; i < g->nu; ++i) {
        float pu = ((i + 0.5f) / g->nu - 0.5f) * (float)g->fu + offsets[2 * k];
        float d0 = c * pu - s * ADD - S0, d1 = s * pu + c * ADD - S1, d2 = pv;
        float tA0 = (Lx / 2 - S0) / d0, tB0 = (-Lx / 2 - S0) / d0;
        float tA1 = (Ly / 2 - S1) / d1, tB1 = (-Ly / 2 - S1) / d1;
        float tA2 = (Lz / 2) / d2, tB2 = (-Lz / 2) / d2;
        float a1 = fmaxf(fmaxf(fmaxf(fminf(tA0, tB0), fminf(tA1, tB1)), fminf(tA2, tB2)), 0.0f);
        float a2 = fminf(fminf(fminf(fmaxf(tA0, tB0), fmaxf(tA1, tB1)), fmaxf(tA2, tB2)), 1.0f);
        float h = (a2 - a1) / n;
        float step = sqrtf(d0 * d0 + d1 * d1 + d2 * d2) * fabsf(a2 - a1) / n;
        /* texel-space start and increment */
        float px = ((S0 + a1 * d0) / Lx + 0.5f) * g->nx - 0.5f, qx = d0 / Lx * h * g->nx;
        float py = ((S1 + a1 * d1) / Ly + 0.5f) * g->ny - 0.5f, qy = d1 / Ly * h * g->ny;
        float pz = ((a1 * d2) / Lz + 0.5f) * g->nz - 0.5f, qz = d2 / Lz * h * g->nz;
        float acc = 0.0f;
        for (int m = 0; m < n; ++m)
          acc += tri_f(vol, g->nx, g->ny, g->nz, px + qx * m, py + qy * m, pz + qz * m);
        proj[i + (size_t)g->nu * j + npix * k] = acc * step;
      }
    }
}

/* Voxel-driven backprojection (optionally FDK weighted), FP32. */
void cpu_bp_f32(const orc_geom *g, const float *angles_deg, const float *offsets, int nviews,
                const float *proj, float *vol, int fbp) {
  const size_t npix = (size_t)g->nu * g->nv;
  const float SAD = (float)g->sad, SDD = (float)g->sdd;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        float px = ((x + 0.5f) / g->nx - 0.5f) * (float)g->lx;
        float py = ((y + 0.5f) / g->ny - 0.5f) * (float)g->ly;
        float pz = ((z + 0.5f) / g->nz - 0.5f) * (float)g->lz;
        float acc = 0.0f;
        for (int k = 0; k < nviews; ++k) {
          float th = angles_deg[k] * (3.14159265358979f / 180.0f);
          float c = cosf(th), s = sinf(th);
          float qx = c * px + s * py, qy = -s * px + c * py;
          float inv = SDD / (qy + SAD);
          float eu = qx * inv - offsets[2 * k], ev = pz * inv - offsets[2 * k + 1];
          float tu = (eu / (float)g->fu + 0.5f) * g->nu - 0.5f;
          float tv = (ev / (float)g->fv + 0.5f) * g->nv - 0.5f;
          float w = 1.0f;
          if (fbp) {
            float U = (SAD - qy) / SAD;
            w = 1.0f / (U * U);
          }
          if (!(fabsf(tu) < 1e9f && fabsf(tv) < 1e9f)) continue;
          float fu = floorf(tu), fv = floorf(tv);
          int iu = (int)fu, iv = (int)fv;
          float au = tu - fu, av = tv - fv;
          const float *P = proj + npix * k;
          float r[2];
          for (int dv = 0; dv < 2; ++dv) {
            int v = iv + dv;
            float a = 0.0f, b = 0.0f;
            if (v >= 0 && v < g->nv) {
              if (iu >= 0 && iu < g->nu) a = P[iu + (size_t)g->nu * v];
              if (iu + 1 >= 0 && iu + 1 < g->nu) b = P[iu + 1 + (size_t)g->nu * v];
            }
            r[dv] = a + au * (b - a);
          }
          acc += w * (r[0] + av * (r[1] - r[0]));
        }
        vol[x + (size_t)g->nx * (y + (size_t)g->ny * z)] = acc;
      }
}
