/* felbm_h5.h -- HDF5 output and restart files in the layout felbm_gpu writes.
 *
 * Host-side C ABI of libfelbm_h5.so (felbm_gpu_amd/csrc/h5io.c), a thin writer
 * over the HDF5 1.10 C library.  The engine library (libfelbm_mi355x.so) does
 * not depend on it: callers download the state or fields from the engine and
 * hand the host arrays to these functions.
 *
 * Every call returns 0 on success or a negative FELBM_H5_E* code, with a
 * message in felbm_h5_last_error().  HDF5's own error stack printing is off.
 *
 * Files (all datasets in the root group, native byte order):
 *   fields      density, concentration, u_x, u_y, u_z, pressure: n values each,
 *               float64 or float32; with deflate > 0 one chunk of n values and
 *               the gzip filter at that level
 *   geometry    coords: n x 3 int32 (x, y, z of each site, same order as the
 *               fields; chunked + gzip when deflate > 0); size: 3 int32, the
 *               largest coordinate + 1 per axis
 *   checkpoint  h, g: Vn float64 each (the post-stream distributions, population
 *               k of site i at k*n + i); step: 1 uint32; target_mass: 1 float64;
 *               Vn: 1 int32
 *   particles   position: m x 3 float64; velocity: m x 3 float64; id: m int32;
 *               optional Dm: m float64, plane_id: m int32, rho_hat: m x 3
 *               float64 and log_rho: m float64
 */
#ifndef FELBM_H5_H
#define FELBM_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FELBM_H5_OK 0
#define FELBM_H5_EARG -1      /* bad argument (null pointer, negative size, bad level) */
#define FELBM_H5_EIO -2       /* the HDF5 library failed (open, create, read, write) */
#define FELBM_H5_EMISMATCH -3 /* the file's Vn or a dataset's shape differs from the caller's */
#define FELBM_H5_EMISSING -4  /* the file has no dataset of that name */

#define FELBM_H5_F64 0
#define FELBM_H5_F32 1
#define FELBM_H5_I32 2

const char *felbm_h5_last_error(void);

/* Replaces write_fields_h5 (src/felbm_gpu_main.cu:47-118), both overloads:
 * the six arrays are float64 when src_f32 == 0 (the std::vector<double>
 * overload) or float32 when src_f32 != 0 (the real_t overload of the float
 * build); out_f32 selects the on-disk type (output_float32), deflate the gzip
 * level 0-9 (output_deflate). */
int felbm_h5_write_fields(const char *path, int64_t n, const void *rho, const void *c, const void *ux,
                          const void *uy, const void *uz, const void *p, int src_f32, int deflate, int out_f32);

/* Replaces write_geometry_h5 (src/felbm_gpu_main.cu:124-143): coords holds
 * x, y, z per site (n x 3, row-major). */
int felbm_h5_write_geometry(const char *path, int64_t n, const int32_t *coords, int deflate);

/* Replaces write_checkpoint (src/felbm_gpu_main.cu:151-166). */
int felbm_h5_write_checkpoint(const char *path, int64_t vn, const double *h, const double *g, uint32_t step,
                              double target_mass);

/* Replaces read_checkpoint (src/felbm_gpu_main.cu:168-189): FELBM_H5_EMISMATCH
 * when the file's Vn differs from vn (the reference's "geometry/precision-set
 * mismatch"), h and g untouched. */
int felbm_h5_read_checkpoint(const char *path, int64_t vn, double *h, double *g, uint32_t *step,
                             double *target_mass);

/* Tracer state file particles_<step>.h5 (ParticleManager::output_state,
 * felbm_gpu_main.cu:712,767-770): position, velocity, id -- the datasets
 * scripts/make_xdmf.py:198-232 reads -- then, when given, Dm, plane_id
 * (README.md:421-436) and rho_hat/log_rho (stretching, README.md:453-455).
 * dm, plane_id, rho_hat and log_rho may be NULL (rho_hat and log_rho together). */
int felbm_h5_write_particles(const char *path, int64_t m, const double *position, const double *velocity,
                             const int32_t *id, const double *dm, const int32_t *plane_id, const double *rho_hat,
                             const double *log_rho);

/* Shape of a dataset (up to the first two extents in dims); FELBM_H5_EMISSING
 * when the file has no such dataset.  With felbm_h5_read_dataset this restores
 * a particle file (ParticleManager::initialize_from_hdf5,
 * felbm_gpu_main.cu:566-575) or reads any file written here. */
int felbm_h5_dataset_shape(const char *path, const char *name, int64_t *dims, int *ndims);

/* Reads a whole dataset of `count` values, converted to `type`
 * (FELBM_H5_F64, FELBM_H5_F32 or FELBM_H5_I32); FELBM_H5_EMISMATCH when the
 * dataset holds another number of values. */
int felbm_h5_read_dataset(const char *path, const char *name, int type, void *buf, int64_t count);

#ifdef __cplusplus
}
#endif

#endif
