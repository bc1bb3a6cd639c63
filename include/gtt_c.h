/*
 * libgtt -- C entry points to the C++ operator / solver layer (include/gtt/).
 *
 * These drive the reference's reconstruction programs end to end on host
 * arrays (first index fastest) so that any FFI caller -- and the parity
 * tests -- can run them:
 *   gtt_fdk_reconstruct_3d  FDK_reconstruct_3d (xray/FDK_reconstruct_3d.cpp:24-207)
 *   gtt_fdk_reconstruct_4d  FDK_reconstruct_4d, McKinnon-Bates 4-D FDK
 *                           (xray/FDK_reconstruct_4d.cpp:24-177)
 *   gtt_os_sart             osSARTSolver over CBSubsetOperator<cuNDArray>
 *                           (solvers/osSARTSolver.h:33-126)
 *   gtt_os_mom              osMOMSolverD + cuTVPrimalDualOperator over
 *                           CBSubsetOperator<cuNDArray> (xray/cuCBOS_reconstruct.cpp:271-523)
 *   gtt_*_weighted          the same over CBSubsetWeightOperator<cuNDArray>
 *                           (xray/CBSubsetWeightOperator.h)
 *   gtt_subset_op_*         CBSubsetOperator<cuNDArray> (xray/CBSubsetOperator.cpp)
 *                           on device pointers, for operator-level checks.
 * Binning: bin_index holds the concatenated view indices of nbins bins whose
 * sizes are in bin_sizes; nbins == 0 selects one bin of all views.  Volumes
 * with nbins > 1 are (nx, ny, nz, nbins).  Projections passed in are in the
 * acquisition's view order.  All functions return 0 or -1 (message in
 * gtt_last_error()).
 */
#ifndef GTT_C_H
#define GTT_C_H

#include "cbct.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *gtt_last_error(void);

/* Storage type (CBCT_F32 or CBCT_BF16) of the projectors' working copies in
 * every reconstruction the calling thread starts from now on (see
 * cbct_storage).  No reference counterpart. */
int gtt_set_default_storage(int dtype);

/* Whether the OS reconstructions the calling thread starts from now on tune
 * the forward projector for their geometry first (cbct_fp_autotune; tuned
 * choices are kept per geometry for the process).  Default off.  No
 * reference counterpart. */
int gtt_set_autotune(int on);

/* Whether the OS-SART reconstructions the calling thread starts from now on
 * replay their iterations 2.. as captured HIP graphs (one per subset order;
 * with gamma = 0 only, else they run directly).  Bitwise the same volumes.
 * Default off.  No reference counterpart. */
int gtt_set_graphs(int on);
/* HIP graphs the calling thread has launched so far (for checking that a
 * reconstruction ran on the graph path). */
long long gtt_graph_launches(void);

int gtt_fdk_reconstruct_3d(const cbct_geom *g, const float *angles_deg, const float *offsets,
                           int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                           const float *projections, const float *filter /* nu+1 or NULL */,
                           float samples_per_pixel, float *volume_out, double *scale_out);

/* volume_out: nx x ny x nz x nbins (one volume per bin, also for nbins ==
 * 1); fdk3d_out (nullable): the FBP over the union of the bins. */
int gtt_fdk_reconstruct_4d(const cbct_geom *g, const float *angles_deg, const float *offsets,
                           int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                           const float *projections, const float *filter /* nu+1 or NULL */,
                           float samples_per_pixel, float *volume_out, float *fdk3d_out,
                           double *scale_out);

int gtt_os_sart(const cbct_geom *g, const float *angles_deg, const float *offsets, int nviews,
                const unsigned *bin_index, const int *bin_sizes, int nbins,
                const float *projections, int nsubsets, int iterations, float beta, float gamma,
                int nonneg, float *volume_out, double *objective_out /* iterations or NULL */);

int gtt_os_mom(const cbct_geom *g, const float *angles_deg, const float *offsets, int nviews,
               const unsigned *bin_index, const int *bin_sizes, int nbins,
               const float *projections, int nsubsets, int iterations, float tv_weight,
               float tau, int reg_steps, float beta, float gamma, int nonneg, int fuse_tv,
               float *volume_out);

/* The same solvers over CBSubsetWeightOperator (xray/CBSubsetWeightOperator.h:11-63,
 * selected by --projection_weights in xray/cuCBOS_reconstruct.cpp:481-498):
 * weights (host, the projections' shape) multiply the data and both sides of
 * the operator.  weights == NULL is the unweighted solver. */
int gtt_os_sart_weighted(const cbct_geom *g, const float *angles_deg, const float *offsets,
                         int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                         const float *projections, const float *weights, int nsubsets,
                         int iterations, float beta, float gamma, int nonneg, float *volume_out,
                         double *objective_out);
int gtt_os_mom_weighted(const cbct_geom *g, const float *angles_deg, const float *offsets,
                        int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                        const float *projections, const float *weights, int nsubsets,
                        int iterations, float tv_weight, float tau, int reg_steps, float beta,
                        float gamma, int nonneg, int fuse_tv, float *volume_out);

typedef struct gtt_subset_op_s *gtt_subset_op;
int gtt_subset_op_create(int nsubsets, gtt_subset_op *out);
int gtt_subset_op_destroy(gtt_subset_op op);
/* Sets up the operator; writes the permuted projections (nullable) and the
 * permutation (nullable, nviews entries). */
int gtt_subset_op_setup(gtt_subset_op op, const cbct_geom *g, const float *angles_deg,
                        const float *offsets, int nviews, const unsigned *bin_index,
                        const int *bin_sizes, int nbins, const float *projections,
                        float *permuted_out, unsigned *permutation_out);
/* As gtt_subset_op_setup, as a CBSubsetWeightOperator when weights (host,
 * the projections' shape) are given; permuted_out then holds the weighted
 * projections. */
int gtt_subset_op_setup_weighted(gtt_subset_op op, const cbct_geom *g, const float *angles_deg,
                                 const float *offsets, int nviews, const unsigned *bin_index,
                                 const int *bin_sizes, int nbins, const float *projections,
                                 const float *weights, float *permuted_out,
                                 unsigned *permutation_out);
int gtt_subset_op_views(gtt_subset_op op, int subset);
/* subset < 0: the full operator.  Device pointers, default stream. */
int gtt_subset_op_mult_M(gtt_subset_op op, float *vol_dev, float *proj_dev, int subset, int accumulate);
int gtt_subset_op_mult_MH(gtt_subset_op op, float *proj_dev, float *vol_dev, int subset, int accumulate);

/* ------------------------------------------------------------------------
 * cuConebeamProjectionOperator (xray/cuConebeamProjectionOperator.{h,cpp}):
 * the device-resident projector over the bins of an acquisition.  Domain
 * (nx, ny, nz[, nbins]) device volumes; codomain the bins' views in bin
 * order (nu, nv, sum of bin sizes), device.  use_offset_correction: 0 / 1
 * forces it (set_use_offset_correction), -1 lets setup decide (mean u
 * offset above a tenth of the detector width, .h / .cpp:183-189).
 */
typedef struct gtt_cu_op_s *gtt_cu_op;
int gtt_cu_op_create(gtt_cu_op *out);
int gtt_cu_op_destroy(gtt_cu_op op);
/* setup(acquisition, binning, is_dims_in_mm, transform_angles) (:148-215) */
int gtt_cu_op_setup(gtt_cu_op op, const cbct_geom *g, const float *angles_deg, const float *offsets,
                    int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                    int transform_angles, int use_offset_correction);
/* mult_M (:38-80) and mult_MH (:82-116, full half-fan weights on a copy of
 * the projections when offset correction is on, then the mask if set). */
int gtt_cu_op_mult_M(gtt_cu_op op, const float *vol_dev, float *proj_dev, int accumulate);
int gtt_cu_op_mult_MH(gtt_cu_op op, const float *proj_dev, float *vol_dev, int accumulate);
/* calculate_mask (:118-146): per bin, the space-carving mask of the bin's
 * projections (conebeam_spacecarver, xray/conebeam_projection.cu:822-921,
 * 1043-1167) into mask_dev (the domain's size, one byte per voxel). */
int gtt_cu_op_calculate_mask(gtt_cu_op op, const float *proj_dev, float limit,
                             unsigned char *mask_dev);
/* set_mask: mult_MH multiplies its result by the mask (apply_mask,
 * xray/conebeam_projection.cu:1651-1657).  The mask is copied; NULL clears. */
int gtt_cu_op_set_mask(gtt_cu_op op, const unsigned char *mask_dev);
/* The decisions setup made: short scan, offset correction. */
int gtt_cu_op_flags(gtt_cu_op op, int *short_scan, int *use_offset_correction);

/* ------------------------------------------------------------------------
 * hoCuConebeamProjectionOperator (xray/hoCuConebeamProjectionOperator.{h,cpp}):
 * host arrays in and out.  setup shifts the angles to start at zero
 * (transform_angles), detects short scans and, with use_offset_correction
 * = -1, turns offset correction on for a mean u offset above a tenth of the
 * detector width (.h:39-93).
 *   mult_M (.cpp:105-176): per bin, FP of the bin's volume into the bin's
 *     views of proj_host (other views untouched; accumulate adds the
 *     projection, taking the views of no bin as zero); then, unless in FBP
 *     mode, the square-root half-fan weights on every view (.cpp:168-169).
 *   mult_MH (.cpp:179-266): FBP mode: per bin, cosine weights, [Parker],
 *     ramp filter (filter_host, nu+1 taps, or the default filter),
 *     [full half-fan weights], FDK-weighted BP, then image *= 2/N_bin
 *     (.cpp:253); otherwise [square-root half-fan weights] and plain BP.
 */
typedef struct gtt_ho_op_s *gtt_ho_op;
int gtt_ho_op_create(gtt_ho_op *out);
int gtt_ho_op_destroy(gtt_ho_op op);
int gtt_ho_op_setup(gtt_ho_op op, const cbct_geom *g, const float *angles_deg, const float *offsets,
                    int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                    int use_fbp, int use_offset_correction, const float *filter_host,
                    float samples_per_pixel);
int gtt_ho_op_mult_M(gtt_ho_op op, const float *image_host, float *proj_host, int accumulate);
int gtt_ho_op_mult_MH(gtt_ho_op op, const float *proj_host, float *image_host, int accumulate);
int gtt_ho_op_flags(gtt_ho_op op, int *short_scan, int *use_offset_correction);

/* The computation of CBCT_forwards_projection
 * (xray/CBCT_forwards_projection.cpp:60-254): angles start + i*spacing
 * (float arithmetic, :167-172), zero-filled projections (:197), a
 * hoCuConebeamProjectionOperator set up on the acquisition and binning
 * (nbins <= 0: one bin of all views) with the automatic offset correction
 * (so half-fan offsets put square-root weights on the result), then mult_M
 * of the image (nx, ny, nz[, nbins], host) into projections_out (nu, nv,
 * nviews, host).  angles_out (nullable) receives the angles as generated.
 * The program's file I/O is gtt_write_real / gtt_h5_acquisition_save. */
int gtt_cbct_forwards_projection(const cbct_geom *g, float start_angle, float angular_spacing,
                                 int nviews, const float *offsets, const unsigned *bin_index,
                                 const int *bin_sizes, int nbins, float samples_per_pixel,
                                 const float *image_host, float *projections_out,
                                 float *angles_out);

/* ------------------------------------------------------------------------
 * ".real" arrays (write_nd_array / python/readReal.py:41-76): uint32 ndims,
 * ndims x uint32 sizes (first index fastest), float32 data.  Host memory.
 */
int gtt_write_real(const char *path, int ndims, const size_t *dims, const float *data);
/* dims: room for 16 sizes. */
int gtt_read_real_header(const char *path, int *ndims, size_t *dims);
/* n: the number of elements the header announces. */
int gtt_read_real(const char *path, float *data, size_t n);

/* ------------------------------------------------------------------------
 * Reconstruction server: FDK and OS-SART requests of any size served by
 * worker lanes on one GPU.  No counterpart in the reference, which runs one
 * reconstruction per process (xray/FDK_reconstruct_3d.cpp:24-207,
 * xray/cuCBOS_reconstruct.cpp:126-579); each request runs exactly those
 * programs (gtt_fdk_reconstruct_3d / gtt_os_sart with one bin), on its
 * lane's own HIP stream.  Results do not depend on the number of lanes or on
 * what else is running.
 *
 * Buffers in a request are caller-owned and must stay valid until
 * gtt_server_wait() returns for it.  Admission: the head of the FIFO starts
 * once a lane is free and the estimated device footprints of the running
 * requests plus its own (gtt_request_footprint) fit the budget given at
 * creation (0: three quarters of the free device memory).
 */
enum { GTT_REQ_FDK = 0, GTT_REQ_OS_SART = 1 };

typedef struct {
  int kind;                   /* GTT_REQ_FDK or GTT_REQ_OS_SART */
  cbct_geom geom;
  int nviews;
  const float *angles_deg;    /* nviews */
  const float *offsets;       /* 2*nviews (u, v) or NULL */
  const float *projections;   /* host, nu x nv x nviews (u fastest) */
  const float *filter;        /* FDK: nu+1 taps, or NULL for the default filter */
  float samples_per_pixel;    /* FDK rescale projection; <= 0 selects 1.5 */
  int subsets, iterations;    /* OS-SART */
  float beta, gamma;
  int nonneg;
  float *volume_out;          /* host, nx x ny x nz */
  int storage;                /* CBCT_F32 (0) or CBCT_BF16 (1) working copies */
} gtt_request;

typedef struct {
  int status;                 /* 0, or -1 with message */
  double queue_ms;            /* submit -> start on a lane */
  double run_ms;              /* start -> result in volume_out */
  double latency_ms;          /* submit -> result */
  double scale;               /* FDK least-squares rescale factor */
  char message[256];
} gtt_result;

typedef struct gtt_server_s *gtt_server;
int gtt_server_create(int device, int lanes, size_t hbm_budget_bytes, gtt_server *out);
int gtt_server_destroy(gtt_server s);  /* drains the queue first */
/* OS-SART requests submitted from now on tune the forward projector for
 * their geometry first (cbct_fp_autotune; choices are kept per geometry for
 * the process, so repeated geometries pay once).  Default off. */
int gtt_server_set_autotune(gtt_server s, int on);
/* OS-SART requests submitted from now on replay their iterations 2.. as
 * captured HIP graphs on their lane's stream (gtt_set_graphs; same bits).
 * Default off. */
int gtt_server_set_graphs(gtt_server s, int on);
int gtt_server_submit(gtt_server s, const gtt_request *req, long long *id);
int gtt_server_wait(gtt_server s, long long id, gtt_result *res);
/* gtt_server_wait with a time limit: GTT_WAIT_TIMEOUT (the request is still
 * queued or running and may be waited for again; its buffers stay in use)
 * when it has not finished within timeout_ms milliseconds, else as
 * gtt_server_wait.  For callers that must not block without bound on a
 * request that stops making progress.  No counterpart in the reference. */
enum { GTT_WAIT_TIMEOUT = 2 };
int gtt_server_wait_for(gtt_server s, long long id, long long timeout_ms, gtt_result *res);
const char *gtt_server_last_error(gtt_server s);
size_t gtt_request_footprint(const gtt_request *req);

#ifdef __cplusplus
}
#endif

#endif /* GTT_C_H */
