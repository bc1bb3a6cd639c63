/*
 * libgtt -- C entry points to the C++ operator / solver layer (include/gtt/).
 *
 * These drive the reference's reconstruction programs end to end on host
 * arrays (first index fastest) so that any FFI caller -- and the parity
 * tests -- can run them:
 *   gtt_fdk_reconstruct_3d  FDK_reconstruct_3d (xray/FDK_reconstruct_3d.cpp:24-207)
 *   gtt_os_sart             osSARTSolver over CBSubsetOperator<cuNDArray>
 *                           (solvers/osSARTSolver.h:33-126)
 *   gtt_os_mom              osMOMSolverD + cuTVPrimalDualOperator over
 *                           CBSubsetOperator<cuNDArray> (xray/cuCBOS_reconstruct.cpp:271-523)
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

int gtt_fdk_reconstruct_3d(const cbct_geom *g, const float *angles_deg, const float *offsets,
                           int nviews, const unsigned *bin_index, const int *bin_sizes, int nbins,
                           const float *projections, const float *filter /* nu+1 or NULL */,
                           float samples_per_pixel, float *volume_out, double *scale_out);

int gtt_os_sart(const cbct_geom *g, const float *angles_deg, const float *offsets, int nviews,
                const unsigned *bin_index, const int *bin_sizes, int nbins,
                const float *projections, int nsubsets, int iterations, float beta, float gamma,
                int nonneg, float *volume_out, double *objective_out /* iterations or NULL */);

int gtt_os_mom(const cbct_geom *g, const float *angles_deg, const float *offsets, int nviews,
               const unsigned *bin_index, const int *bin_sizes, int nbins,
               const float *projections, int nsubsets, int iterations, float tv_weight,
               float tau, int reg_steps, float beta, float gamma, int nonneg, int fuse_tv,
               float *volume_out);

typedef struct gtt_subset_op_s *gtt_subset_op;
int gtt_subset_op_create(int nsubsets, gtt_subset_op *out);
int gtt_subset_op_destroy(gtt_subset_op op);
/* Sets up the operator; writes the permuted projections (nullable) and the
 * permutation (nullable, nviews entries). */
int gtt_subset_op_setup(gtt_subset_op op, const cbct_geom *g, const float *angles_deg,
                        const float *offsets, int nviews, const unsigned *bin_index,
                        const int *bin_sizes, int nbins, const float *projections,
                        float *permuted_out, unsigned *permutation_out);
int gtt_subset_op_views(gtt_subset_op op, int subset);
/* subset < 0: the full operator.  Device pointers, default stream. */
int gtt_subset_op_mult_M(gtt_subset_op op, float *vol_dev, float *proj_dev, int subset, int accumulate);
int gtt_subset_op_mult_MH(gtt_subset_op op, float *proj_dev, float *vol_dev, int subset, int accumulate);

#ifdef __cplusplus
}
#endif

#endif /* GTT_C_H */
