/*
 * libcbct -- MI355X (gfx950) cone-beam CT projector: the C ABI.
 *
 * This is the drop-in boundary for the hot path of dchansen/gt-tomography.
 * Each entry point names the reference interface it replaces.  All arrays
 * are first-index-fastest (Gadgetron layout):
 *   volume       V[x + nx*(y + ny*z)]            (nx, ny, nz)
 *   projections  P[u + nu*(v + nv*k)]            (nu, nv, nviews)
 *
 * Conventions
 *  - Device pointers are caller-owned (hipMalloc / torch); host pointers are
 *    named *_host.  The context owns its scratch (padded copies, filter
 *    tables, reduction partials) and never allocates per call once sized.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  Work
 *    is enqueued asynchronously unless the function returns a host value.
 *  - Every function returns CBCT_OK (0) or a negative CBCT_ERR_* code and
 *    never throws; cbct_last_error() returns the message.  The C++ operator
 *    layer rethrows these as std::runtime_error with the reference's text.
 *  - A context is single-threaded; distinct contexts may be used concurrently.
 *    Library-wide state is limited to one upload stream per device (table
 *    uploads, serialised), the autotuner's table and cbct_fp_noalloc's
 *    scratch; none of it touches the legacy default stream.
 *  - Calls on one context may use different streams: a call that uses the
 *    context's scratch first makes its stream wait for the previous such call
 *    when that was enqueued on another stream, so the scratch is never
 *    overwritten while a kernel still reads it.  cbct_views_set and
 *    cbct_filter_set wait (on the host) for this context's earlier calls
 *    only, not for the whole device.
 */
#ifndef CBCT_H
#define CBCT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cbct_ctx_s *cbct_ctx;

/* Volume + detector geometry.  Mirrors CBCT_geometry (SAD, SDD, FOV) plus the
 * operator's is_dims_in_mm (xray/CBCT_acquisition.h:24-120). */
typedef struct {
  int nx, ny, nz;      /* volume voxels                                   */
  float lx, ly, lz;    /* volume extent in mm (is_dims_in_mm)            */
  int nu, nv;          /* detector pixels                                 */
  float fov_u, fov_v;  /* detector extent in mm (CBCT_geometry::get_FOV) */
  float sad, sdd;      /* source-axis, source-detector distance in mm    */
} cbct_geom;

enum { CBCT_F32 = 0, CBCT_BF16 = 1 };

enum {
  CBCT_OK = 0,
  CBCT_ERR_ARG = -1,         /* illegal pointer / size / index                  */
  CBCT_ERR_STATE = -2,       /* e.g. no views or filter set ("setup not performed") */
  CBCT_ERR_HIP = -3,         /* HIP runtime error                                */
  CBCT_ERR_NOMEM = -4,       /* workspace could not be allocated                 */
  CBCT_ERR_UNSUPPORTED = -5  /* option not available for this dtype/shape        */
};

/* ---------------------------------------------------------------- context */
int cbct_ctx_create(int device, cbct_ctx *out);
int cbct_ctx_destroy(cbct_ctx ctx);
/* Message of the last failing call on ctx (ctx == NULL: last create failure). */
const char *cbct_last_error(cbct_ctx ctx);
/* Bytes of device scratch the context currently holds. */
size_t cbct_workspace_bytes(cbct_ctx ctx);
/* Kernel timing (no reference counterpart; the reference wraps whole driver
 * phases in GPUTimer).  When enabled, the context brackets the main kernel of
 * the next cbct_fp / cbct_bp / cbct_ramp_filter launches with HIP events on
 * the launch stream; cbct_profile_read waits for them and returns the last
 * durations in ms (negative if that kernel has not run since enabling). */
int cbct_profile_enable(cbct_ctx ctx, int on);
int cbct_profile_read(cbct_ctx ctx, double *fp_ms, double *bp_ms, double *filter_ms);
/* Frees the context's scratch (it regrows on demand), after the last call
 * that used it has finished. */
int cbct_ctx_trim(cbct_ctx ctx);
/* ctx drops its own scratch and uses owner's from now on (same device).  The
 * scratch lives until the last context holding it is destroyed.  Contexts
 * that share scratch are used from one host thread at a time; their calls
 * are ordered as calls on one context are (above).  Lets the many per-subset
 * operator contexts of one reconstruction hold one padded volume and one
 * padded sinogram copy between them instead of one each.  No reference
 * counterpart (the reference allocates these per call). */
int cbct_ctx_share_workspace(cbct_ctx ctx, cbct_ctx owner);
/* Settings back to their defaults (engines, tile, view group, z-depth,
 * storage type, profiling); the view table, filter and scratch are kept.
 * For callers that reuse contexts.  No reference counterpart. */
int cbct_ctx_reset(cbct_ctx ctx);

/* View table: n views, angles in degrees, offsets (u, v) in mm interleaved
 * (offsets_host may be NULL = all zero).  Replaces the per-call
 * thrust::device_vector uploads of angles/offsets
 * (xray/conebeam_projection.cu:600-601, :1225-1226); subsequent projector
 * calls address views [view0, view0 + nviews) of this table. */
int cbct_views_set(cbct_ctx ctx, const float *angles_deg_host, const float *offsets_host, int n);
int cbct_views_count(cbct_ctx ctx);

/* ------------------------------------------------------------- projectors */
/* Forward projection, ray-driven, floor(samples_per_pixel*nx) trilinear
 * samples per ray scaled by the step length in mm.  proj = (accumulate ? proj
 * : 0) + A vol for views [view0, view0+nviews).  vol_dtype CBCT_F32 or
 * CBCT_BF16 (bf16 storage, fp32 arithmetic).
 * Replaces conebeam_forwards_projection(cuNDArray*, ...)
 *   (xray/conebeam_projection.h:40-51; kernel xray/conebeam_projection.cu:412-528). */
int cbct_fp(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *vol,
            int vol_dtype, float *proj, float samples_per_pixel, int accumulate, void *stream);

/* cbct_fp that allocates nothing: the same kernels and the same bits, with
 * the volume's padded working copy in library-owned device memory of 64 MiB
 * per GPU (held until the process ends) instead of the context's workspace.
 * Volumes whose working copy does not fit (about 250^3 FP32 or 320^3 BF16
 * working copies) are refused with CBCT_ERR_UNSUPPORTED.  Every call on one
 * GPU shares that memory: calls that overlap on one GPU (other streams,
 * other threads, other contexts) race and must be ordered by the caller.
 * The device's first call looks the memory's address up; later calls make
 * no allocation of any kind.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  cbct_views_set must have been called
 * on the context (it allocates the view table). */
int cbct_fp_noalloc(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *vol,
                    int vol_dtype, float *proj, float samples_per_pixel, int accumulate,
                    void *stream);

/* Voxel-driven backprojection with bilinear detector interpolation.
 * vol = (accumulate ? vol : 0) + sum_views w * P(project(voxel)), w = 1 or the
 * FDK weight 1/U^2 when fbp != 0.  proj_dtype CBCT_F32 or CBCT_BF16.
 * Replaces conebeam_backwards_projection(cuNDArray*, ...) and the
 * conebeam_backwards_projection_kernel<FBP> template
 *   (xray/conebeam_projection.h:57-69; kernel xray/conebeam_projection.cu:922-1037).
 * Offset correction is a separate call (cbct_offset_weights). */
int cbct_bp(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const void *proj,
            int proj_dtype, float *vol, int fbp, int accumulate, void *stream);

/* Forward-projector engine: 1 = detector-column sharing (default; one wave per
 * detector column interpolates the shared z-line once per step: the variant
 * that keeps the column's four corner z-lines in registers across steps when
 * every row group's z window fits two chunks of 128 z, else the per-step
 * variant measured fastest for the storage type), 2 = the other per-step
 * variant, 3 / 4 = the cached variant with a two / three chunk window (falls
 * back to per-step loads per column when the window is larger), 0 = one
 * thread per ray.  All give bitwise-identical projections.  No reference
 * counterpart (the reference has one texture-based kernel). */
int cbct_fp_engine(cbct_ctx ctx, int engine);

/* Forward-projector autotuner: times candidate column-engine variants and
 * tiles (each once warm, twice timed) for views [view0, view0+nviews) of
 * this geometry on scratch arrays, and records the fastest for the geometry,
 * samples per ray and working-copy type in a process-wide table that engine
 * 1 consults before its built-in choice.  A geometry already in the table
 * returns at once unless retune != 0.  All candidates give the same bits.
 * Time it on an otherwise idle device.  Synchronises the stream; allocates
 * and frees its scratch.  best_ms (nullable): the winner's time.  No
 * reference counterpart. */
int cbct_fp_autotune(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, int vol_dtype,
                     float samples_per_pixel, int retune, void *stream, double *best_ms);
/* Forgets every tuned choice. */
int cbct_fp_autotune_clear(void);

/* Name of the kernel cbct_fp would launch for these arguments under the
 * context's current settings ("k_fp_ray", "k_fp_col" or "k_fp_colc"; "" on
 * bad arguments), for matching profiler output.  No reference counterpart. */
const char *cbct_fp_kernel(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews,
                           int vol_dtype, float samples_per_pixel);

/* Column-engine tile: the most detector rows one lane carries (1..12,
 * default 8).  Columns with more rows than 64 x this are split into row
 * groups, each with its own z-line loads.  Results do not depend on it.  No
 * reference counterpart. */
int cbct_fp_tile(cbct_ctx ctx, int max_rows_per_lane);

/* Column-engine launch order: the blocks of `views` (1..64) adjacent views
 * of one detector column and row group are launched one after another, so
 * that they run together on one XCD and share its L2 (their rays pass close
 * to each other through the volume).  1 = view by view; default 6.
 * cbct_fp_autotune also tries 1, 3, 6 and 8.  Results do not depend on it.
 * No reference counterpart. */
int cbct_fp_view_group(cbct_ctx ctx, int views);

/* Backprojector engine: 2 = default: LDS-staged kernels (per view, each
 * block's detector rectangle is copied into LDS by direct global-to-LDS
 * loads, one view ahead, and read from there): with FP32 working copies from
 * a plain zero-bordered copy; with BF16 working copies from a copy of
 * horizontal tap pairs (two taps per dword, one LDS read per voxel-view),
 * where the rectangles fit the LDS tile at a sample of the views, else as 1;
 * 3 = as 2, with the BF16 pair kernel whatever the rectangles' size (views
 * whose rectangle does not fit read the pair copy directly); 1 = quad-texel
 * copy of the projections (the four bilinear taps of a voxel-view in one
 * 16-byte (FP32) or 8-byte (BF16) load, 4x the padded projection workspace);
 * 0 = plain zero-bordered copy, taps loaded one by one.  All give
 * bitwise-identical volumes.  No reference counterpart (the reference samples
 * a layered texture). */
int cbct_bp_engine(cbct_ctx ctx, int engine);

/* Name of the backprojection kernel cbct_bp would launch for these
 * arguments under the context's current settings ("k_bpp" or "k_bp", either
 * storage type; "" on bad arguments), for matching profiler output.  No
 * reference counterpart. */
const char *cbct_bp_kernel(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews);

/* Backprojector: z-voxels per thread, 0 = chosen by size (16, or 8 when 16
 * would make fewer than 1024 tiles), or 8 / 16 forced (tests run both at
 * every size).  The two differ in rounding only (each thread steps the
 * detector row coordinate from its first voxel).  No reference counterpart. */
int cbct_bp_kz(cbct_ctx ctx, int kz);

/* Storage type of the projectors' working copies (the zero-bordered volume
 * copy the forward projector gathers from, the projection copy the
 * backprojector gathers from) when the caller's arrays are FP32:
 * CBCT_F32 (default) or CBCT_BF16 (the copy kernels round to nearest even;
 * arithmetic stays FP32; relative error ~1e-3 per projection).  BF16 arrays
 * passed by the caller are always used as they are.  No reference
 * counterpart (the reference samples FP32 textures). */
int cbct_storage(cbct_ctx ctx, int dtype);

/* Exact transpose of cbct_fp (ray-driven scatter, fp32 atomics):
 * vol = (accumulate ? vol : 0) + A^T proj.  No reference counterpart: the
 * reference (mult_M, mult_MH) pair is not adjoint; this exists for the
 * adjoint test. */
int cbct_fp_transpose(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const float *proj,
                      float *vol, float samples_per_pixel, int accumulate, void *stream);

/* Space-carving mask: mask[v] = (number of views whose nearest-texel sample
 * p satisfies p < limit && p != 0) < 3.
 * Replaces conebeam_spacecarver (xray/conebeam_projection.h:74-85;
 * kernel xray/conebeam_projection.cu:822-921). */
int cbct_spacecarver(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const float *proj,
                     unsigned char *mask, float limit, void *stream);

/* image *= mask.  Replaces apply_mask (xray/conebeam_projection.cu:1651-1657). */
int cbct_apply_mask(cbct_ctx ctx, size_t n, float *image, const unsigned char *mask, void *stream);

/* ----------------------------------------------------------------- FDK */
/* Cosine weights (nv x nu, fp64 arithmetic then fp32) into a host buffer.
 * Replaces hoCuConebeamProjectionOperator::compute_cosine_weights
 *   (xray/hoCuConebeamProjectionOperator.cpp:64-102). */
int cbct_cosine_weights(const cbct_geom *g, float *weights_host);

/* Default (Guo et al.) frequency filter, nu+1 taps.
 * Replaces compute_default_frequency_filter
 *   (xray/hoCuConebeamProjectionOperator.cpp:26-62). */
int cbct_default_filter(int nu, float *filter_host);

/* Install the one-sided frequency response F[0..nu] (nu+1 floats, host).  The
 * context converts it to the equivalent spatial Toeplitz kernel of the
 * pad(2nu) -> R2C -> *F -> C2R -> /2nu -> crop chain, in fp64, and keeps the
 * fp32 kernel on the device; the engines' own tables (the fp16 hi/lo split
 * for the MFMA engine, twiddles and the kernel's spectrum for the FFT
 * engine) are built from the fp64 kernel on first use.
 * Replaces set_frequency_filter (xray/hoCuConebeamProjectionOperator.h:117-119). */
int cbct_filter_set(cbct_ctx ctx, int nu, const float *filter_host);

/* Ramp-filter engine selection: 0 = fp32 VALU Toeplitz, 1 = fp16 hi/lo split
 * MFMA Toeplitz with fp32 accumulation, 2 = FFT (default; the reference's
 * method: two rows per complex transform of length 2^k >= 2nu - 1, fp32 with
 * fp64-derived twiddles and spectrum; rows longer than 4096 use engine 1).
 * Other values: CBCT_ERR_UNSUPPORTED. */
int cbct_filter_engine(cbct_ctx ctx, int engine);

/* The FDK pre-filter chain for views [view0, view0+nviews):
 * out = [offset_w]( filter( [parker]( in * cosine_w ) ) ).  in and out may
 * alias.  Replaces the FBP branch of the batched backprojection
 *   (xray/conebeam_projection.cu:1483-1521). */
int cbct_fdk_prefilter(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, const float *in,
                       float *out, int short_scan, int offset_correct, void *stream);

/* Row filter alone (no weights): out[r] = T in[r] for nrows rows of nu. */
int cbct_ramp_filter(cbct_ctx ctx, int nu, size_t nrows, const float *in, float *out,
                     void *stream);

/* Half-fan offset weights in place; sqrt_variant selects the square-root form.
 * Replaces offset_correct / offset_correct_sqrt / apply_offset_correct
 *   (xray/conebeam_projection.cu:230-404). */
int cbct_offset_weights(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, float *proj,
                        int sqrt_variant, void *stream);

/* Parker short-scan weights in place.  Replaces redundancy_correct
 *   (xray/conebeam_projection.cu:141-227). */
int cbct_parker_weights(cbct_ctx ctx, const cbct_geom *g, int view0, int nviews, float *proj,
                        void *stream);

/* ------------------------------------------------ TV / solver passes */
/* out = (accumulate ? out : 0) + weight * D^T Pi( (weight*sigma)/(1+sigma*alpha) D x )
 * per 3-D volume, nvol volumes back to back, periodic forward differences;
 * one fused kernel, no 3*Nvox scratch.
 * Replaces cuTVPrimalDualOperator<float>::primalDual
 *   (operators/cuTVPrimalDualOperator.cu:10-106). */
int cbct_tv_primal_dual(cbct_ctx ctx, int nx, int ny, int nz, int nvol, const float *x,
                        float *out, float sigma, float alpha, float weight, int accumulate,
                        void *stream);

/* x = (x - tau*g + (tau/avg_lambda)*s/precon) / (1 + tau/(avg_lambda*precon)).
 * Replaces osMOMSolverD::updateG (solvers/osMOMSolverD.h:233-247, the base
 * class form the cuCBOS_reconstruct driver instantiates). */
int cbct_update_g(cbct_ctx ctx, size_t n, const float *g, float *x, const float *s,
                  const float *precon, float tau, float avg_lambda, void *stream);

/* Fused TV proximal step: x_out = updateG(TV(x_in), x_in, s, precon) with the
 * TV gradient computed in registers (single isotropic TV regulariser).
 * x_out must not alias x_in.  Equivalent to clear(g); cbct_tv_primal_dual(x,
 * g, ...); cbct_update_g(g, x, ...). */
int cbct_tv_update_g(cbct_ctx ctx, int nx, int ny, int nz, int nvol, const float *x_in,
                     float *x_out, const float *s, const float *precon, float sigma,
                     float alpha, float weight, float tau, float avg_lambda, void *stream);

/* OS-SART residual in place: r = -(r - b) * w   (r holds A_s x on entry).
 * Replaces solvers/osSARTSolver.h:92-98. */
int cbct_sart_residual(cbct_ctx ctx, size_t n, float *r, const float *b, const float *w,
                       void *stream);

/* OS-SART update: x = x + step * img * wx; if nonneg x = max(x, 0).
 * Replaces solvers/osSARTSolver.h:100-105. */
int cbct_sart_update(cbct_ctx ctx, size_t n, float *x, const float *img, const float *wx,
                     float step, int nonneg, void *stream);

/* OS-MOM gradient step: z = z + step * img * precon.
 * Replaces solvers/osMOMSolverD.h:156-157. */
int cbct_mom_gradient_step(cbct_ctx ctx, size_t n, float *z, const float *img,
                           const float *precon, float step, void *stream);

/* OS-MOM momentum: if nonneg x = max(x, 0); z = x*(1 + c) - c*xold with
 * c = (told-1)/t.  Replaces solvers/osMOMSolverD.h:166-173. */
int cbct_mom_momentum(cbct_ctx ctx, size_t n, float *x, const float *xold, float *z, float c,
                      int nonneg, void *stream);

/* Gadgetron cuNDArray element-wise primitives used by the solvers. */
int cbct_fill(cbct_ctx ctx, size_t n, float *x, float value, void *stream);
int cbct_scale(cbct_ctx ctx, size_t n, float *x, float a, void *stream);            /* x *= a  */
int cbct_add_scalar(cbct_ctx ctx, size_t n, float *x, float a, void *stream);       /* x += a  */
int cbct_axpy(cbct_ctx ctx, size_t n, float a, const float *x, float *y, void *stream); /* y += a x */
int cbct_mul(cbct_ctx ctx, size_t n, const float *a, float *y, void *stream);       /* y *= a  */
int cbct_div(cbct_ctx ctx, size_t n, const float *a, float *y, void *stream);       /* y /= a  */
int cbct_sub(cbct_ctx ctx, size_t n, const float *a, float *y, void *stream);       /* y -= a  */
int cbct_add(cbct_ctx ctx, size_t n, const float *a, float *y, void *stream);       /* y += a  */
int cbct_clamp_min(cbct_ctx ctx, size_t n, float *x, float lo, void *stream);
int cbct_reciprocal(cbct_ctx ctx, size_t n, float *x, void *stream);
/* Strided broadcast multiply: y[i] *= a[i % period] (cosine weights over views). */
int cbct_mul_periodic(cbct_ctx ctx, size_t n, const float *a, size_t period, float *y,
                      void *stream);

/* fp32 <-> bf16 storage conversion (round to nearest even). */
int cbct_to_bf16(cbct_ctx ctx, size_t n, const float *x, unsigned short *y, void *stream);
int cbct_from_bf16(cbct_ctx ctx, size_t n, const unsigned short *x, float *y, void *stream);

/* FP64 reductions (deterministic two-pass, fp64 accumulation on the device).
 * op: 0 = dot(a,b), 1 = nrm2(a), 2 = asum(a), 3 = sum(a).  The result is
 * written to *out_host after the stream is synchronised.  Replaces the
 * cuBLAS fp32 dot / nrm2 / asum the solvers and the FDK driver use. */
enum { CBCT_DOT = 0, CBCT_NRM2 = 1, CBCT_ASUM = 2, CBCT_SUM = 3 };
int cbct_reduce_f64(cbct_ctx ctx, int op, size_t n, const float *a, const float *b,
                    double *out_host, void *stream);

/* ----------------------------------------------------------- image quality */
/* Multi-scale SSIM of a volume against a reference volume (device arrays,
 * nx*ny*nz, x fastest), written to *out_host:
 *   mean( prod_{i<scales} CS(sigma 2^i) * L(sigma 2^(scales-1)) ),
 *   CS = (2 cov + C2)/(var_x + var_y + C2), L = (2 mu_x mu_y + C1)/(mu_x^2 + mu_y^2 + C1),
 *   C1 = (0.01 R)^2, C2 = (0.03 R)^2, R = max(reference) - min(reference).
 * Local moments are Gaussian-weighted: separable taps exp(-k^2/(2 sigma^2)),
 * |k| <= int(3.5 sigma + 0.5), normalised; half-sample symmetric borders.
 * Workspace: 11 floats per voxel (context-owned).  Synchronises the stream.
 * Replaces Gadgetron::mssim (xray/mssim.h:7, xray/mssim.cpp:77-121). */
int cbct_mssim(cbct_ctx ctx, int nx, int ny, int nz, const float *image, const float *reference,
               float sigma_x, float sigma_y, float sigma_z, int scales, double *out_host,
               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CBCT_H */
