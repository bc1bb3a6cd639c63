/*
 * libgtt_h5 -- the reference's HDF5 file formats (optional; built when the
 * HDF5 1.10 headers and libraries are present).
 *
 *   acquisition  xray/CBCT_acquisition.h:177-333 (load/save), geometry part
 *                :76-111 -- /geometry_dataformat_version (2), /angles,
 *                /offsetx, /offsety, /SAD, /SDD, /FOV (or /spacing x
 *                detector size), /projection_dataformat_version (1),
 *                /projections [nviews][nv][nu], optional /airscan.
 *   binning      xray/CBCT_binning.h:106-157 (load); the writer follows
 *                python/l1shroud.py:139-150 -- /binning_dataformat_version
 *                (1), /numBins, /bin_1 .. /bin_N (uint32 view indices).
 *   image        proton/hdf5_utils.h:17-45 -- /image (dimensions reversed)
 *                with attributes input, dimensions, element_size_um,
 *                origin, iterations.
 *
 * Arrays are host arrays, first index fastest (u, then v, then view; x, y,
 * z).  Functions return 0 or -1 with the reference's message in
 * gtt_h5_last_error().
 */
#ifndef GTT_H5_H
#define GTT_H5_H

#ifdef __cplusplus
extern "C" {
#endif

const char *gtt_h5_last_error(void);

/* Sizes of an acquisition file; airscan_dims gets (nu, nv, n) of /airscan
 * or zeros. */
int gtt_h5_acquisition_info(const char *path, int *nu, int *nv, int *nviews, int *airscan_dims);
/* Reads everything; airscan may be NULL.  offsets: 2*nviews (u, v). */
int gtt_h5_acquisition_read(const char *path, float *sad, float *sdd, float *fov /* 2 */,
                            float *angles_deg, float *offsets, float *projections,
                            float *airscan);
int gtt_h5_acquisition_write(const char *path, float sad, float sdd, const float *fov /* 2 */,
                             int nu, int nv, int nviews, const float *angles_deg,
                             const float *offsets /* 2*nviews or NULL */,
                             const float *projections);

int gtt_h5_binning_info(const char *path, int *nbins, int *total);
int gtt_h5_binning_read(const char *path, int *sizes /* nbins */, unsigned *index /* total */);
int gtt_h5_binning_write(const char *path, int nbins, const int *sizes, const unsigned *index);

int gtt_h5_image_write(const char *path, const float *image, int nx, int ny, int nz,
                       const float *dims_mm /* 3 */, const float *origin /* 3 */,
                       const char *arguments, unsigned iterations);

#ifdef __cplusplus
}
#endif

#endif /* GTT_H5_H */
