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
int gtt_server_submit(gtt_server s, const gtt_request *req, long long *id);
int gtt_server_wait(gtt_server s, long long id, gtt_result *res);
const char *gtt_server_last_error(gtt_server s);
size_t gtt_request_footprint(const gtt_request *req);

#ifdef __cplusplus
}
#endif

#endif /* GTT_C_H */
