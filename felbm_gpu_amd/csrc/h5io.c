/* h5io.c -- HDF5 files in the layout felbm_gpu writes (fields, geometry,
 * checkpoint, particles), over the HDF5 1.10 C library.
 *
 * The call sequence per file is the one the reference's writers make through
 * the C++ wrapper (src/felbm_gpu_main.cu:47-189): create the file with
 * H5F_ACC_TRUNC and default property lists; per dataset, a simple dataspace, a
 * dataset-creation list (chunk = the whole dataset + deflate only when gzip is
 * on), create, write with the native memory type, close.  With the same HDF5
 * library that gives the same bytes on disk.
 */
#include "felbm_h5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *felbm_h5_last_error(void) { return g_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* One dataset of rank 1 or 2; plist < 0 means H5P_DEFAULT. */
static int put(hid_t f, const char *name, int rank, const hsize_t *dims, hid_t ftype, hid_t mtype, const void *buf,
               hid_t plist) {
  hid_t sp = H5Screate_simple(rank, dims, NULL);
  if (sp < 0) return fail(FELBM_H5_EIO, "H5Screate_simple failed for %s", name);
  hid_t ds = H5Dcreate2(f, name, ftype, sp, H5P_DEFAULT, plist < 0 ? H5P_DEFAULT : plist, H5P_DEFAULT);
  if (ds < 0) {
    H5Sclose(sp);
    return fail(FELBM_H5_EIO, "H5Dcreate2 failed for %s", name);
  }
  herr_t w = H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
  H5Dclose(ds);
  H5Sclose(sp);
  if (w < 0) return fail(FELBM_H5_EIO, "H5Dwrite failed for %s", name);
  return FELBM_H5_OK;
}

/* gzip plist chunked over the whole dataset, or H5P_DATASET_CREATE defaults */
static hid_t dcpl(int rank, const hsize_t *dims, int deflate, int64_t n) {
  hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
  if (pl < 0) return pl;
  if (deflate > 0 && n > 0) {
    if (H5Pset_chunk(pl, rank, dims) < 0 || H5Pset_deflate(pl, (unsigned)deflate) < 0) {
      H5Pclose(pl);
      return -1;
    }
  }
  return pl;
}

int felbm_h5_write_fields(const char *path, int64_t n, const void *rho, const void *c, const void *ux,
                          const void *uy, const void *uz, const void *p, int src_f32, int deflate, int out_f32) {
  const void *src[6] = {rho, c, ux, uy, uz, p};
  static const char *names[6] = {"density", "concentration", "u_x", "u_y", "u_z", "pressure"};
  if (!path || n < 0 || deflate < 0 || deflate > 9) return fail(FELBM_H5_EARG, "bad path, n or deflate level");
  for (int i = 0; i < 6; ++i)
    if (!src[i] && n > 0) return fail(FELBM_H5_EARG, "null field %s", names[i]);
  quiet();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot create %s", path);
  hsize_t d[1] = {(hsize_t)n};
  hid_t pl = dcpl(1, d, deflate, n);
  int rc = pl < 0 ? fail(FELBM_H5_EIO, "dataset-creation list") : FELBM_H5_OK;
  /* the on-disk type differs from the source: convert on the host first, as the
     reference does (a plain C cast, round to nearest) */
  void *tmp = NULL;
  if (rc == 0 && n > 0 && (!src_f32) != (!out_f32)) {
    tmp = malloc((size_t)n * (out_f32 ? sizeof(float) : sizeof(double)));
    if (!tmp) rc = fail(FELBM_H5_EARG, "out of host memory");
  }
  hid_t ty = out_f32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
  for (int i = 0; i < 6 && rc == 0; ++i) {
    const void *buf = src[i];
    if (tmp) {
      if (out_f32) {
        const double *s = (const double *)src[i];
        float *t = (float *)tmp;
        for (int64_t j = 0; j < n; ++j) t[j] = (float)s[j];
      } else {
        const float *s = (const float *)src[i];
        double *t = (double *)tmp;
        for (int64_t j = 0; j < n; ++j) t[j] = (double)s[j];
      }
      buf = tmp;
    }
    rc = put(f, names[i], 1, d, ty, ty, buf, pl);
  }
  free(tmp);
  if (pl >= 0) H5Pclose(pl);
  if (H5Fclose(f) < 0 && rc == 0) rc = fail(FELBM_H5_EIO, "closing %s failed", path);
  return rc;
}

int felbm_h5_write_geometry(const char *path, int64_t n, const int32_t *coords, int deflate) {
  if (!path || n < 0 || (n > 0 && !coords) || deflate < 0 || deflate > 9)
    return fail(FELBM_H5_EARG, "bad path, n, coords or deflate level");
  int32_t mx[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      if (coords[3 * i + a] > mx[a]) mx[a] = coords[3 * i + a];
  int32_t dims[3] = {mx[0] + 1, mx[1] + 1, mx[2] + 1};
  quiet();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot create %s", path);
  hsize_t d2[2] = {(hsize_t)n, 3};
  hid_t pl = dcpl(2, d2, deflate, n);
  int rc = pl < 0 ? fail(FELBM_H5_EIO, "dataset-creation list") : FELBM_H5_OK;
  if (rc == 0) rc = put(f, "coords", 2, d2, H5T_NATIVE_INT, H5T_NATIVE_INT, coords, pl);
  if (pl >= 0) H5Pclose(pl);
  hsize_t d1[1] = {3};
  if (rc == 0) rc = put(f, "size", 1, d1, H5T_NATIVE_INT, H5T_NATIVE_INT, dims, -1);
  if (H5Fclose(f) < 0 && rc == 0) rc = fail(FELBM_H5_EIO, "closing %s failed", path);
  return rc;
}

int felbm_h5_write_checkpoint(const char *path, int64_t vn, const double *h, const double *g, uint32_t step,
                              double target_mass) {
  if (!path || vn < 0 || vn > 0x7fffffff || (vn > 0 && (!h || !g)))
    return fail(FELBM_H5_EARG, "bad path, Vn or state pointers");
  quiet();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot create %s", path);
  hsize_t d[1] = {(hsize_t)vn}, d1[1] = {1};
  int32_t v = (int32_t)vn;
  int rc = put(f, "h", 1, d, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, h, -1);
  if (rc == 0) rc = put(f, "g", 1, d, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, g, -1);
  if (rc == 0) rc = put(f, "step", 1, d1, H5T_NATIVE_UINT, H5T_NATIVE_UINT, &step, -1);
  if (rc == 0) rc = put(f, "target_mass", 1, d1, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &target_mass, -1);
  if (rc == 0) rc = put(f, "Vn", 1, d1, H5T_NATIVE_INT, H5T_NATIVE_INT, &v, -1);
  if (H5Fclose(f) < 0 && rc == 0) rc = fail(FELBM_H5_EIO, "closing %s failed", path);
  return rc;
}

/* Reads a whole dataset whose element count must be `count` (rank-agnostic). */
static int get(hid_t f, const char *name, hid_t mtype, void *buf, hsize_t count) {
  hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
  if (ds < 0) return fail(FELBM_H5_EIO, "no dataset %s", name);
  hid_t sp = H5Dget_space(ds);
  hssize_t np = sp < 0 ? -1 : H5Sget_simple_extent_npoints(sp);
  if (sp >= 0) H5Sclose(sp);
  if (np < 0 || (hsize_t)np != count) {
    H5Dclose(ds);
    return fail(FELBM_H5_EMISMATCH, "dataset %s holds %lld values, expected %llu", name, (long long)np,
                (unsigned long long)count);
  }
  herr_t r = H5Dread(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
  H5Dclose(ds);
  return r < 0 ? fail(FELBM_H5_EIO, "reading %s failed", name) : FELBM_H5_OK;
}

int felbm_h5_read_checkpoint(const char *path, int64_t vn, double *h, double *g, uint32_t *step,
                             double *target_mass) {
  if (!path || vn < 0 || (vn > 0 && (!h || !g)) || !step || !target_mass)
    return fail(FELBM_H5_EARG, "bad path, Vn or output pointers");
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot open %s", path);
  int32_t v = 0;
  int rc = get(f, "Vn", H5T_NATIVE_INT, &v, 1);
  if (rc == 0 && (int64_t)v != vn)
    rc = fail(FELBM_H5_EMISMATCH, "restart file Vn=%d != current Vn=%lld (geometry/precision-set mismatch)", v,
              (long long)vn);
  if (rc == 0) rc = get(f, "h", H5T_NATIVE_DOUBLE, h, (hsize_t)vn);
  if (rc == 0) rc = get(f, "g", H5T_NATIVE_DOUBLE, g, (hsize_t)vn);
  if (rc == 0) rc = get(f, "step", H5T_NATIVE_UINT, step, 1);
  if (rc == 0) rc = get(f, "target_mass", H5T_NATIVE_DOUBLE, target_mass, 1);
  H5Fclose(f);
  return rc;
}

int felbm_h5_write_particles(const char *path, int64_t m, const double *position, const double *velocity,
                             const int32_t *id, const double *dm, const int32_t *plane_id, const double *rho_hat,
                             const double *log_rho) {
  if (!path || m < 0 || (m > 0 && (!position || !velocity || !id)) || (!rho_hat) != (!log_rho))
    return fail(FELBM_H5_EARG, "bad path, count or arrays (rho_hat and log_rho go together)");
  quiet();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot create %s", path);
  hsize_t d2[2] = {(hsize_t)m, 3}, d1[1] = {(hsize_t)m};
  int rc = put(f, "position", 2, d2, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, position, -1);
  if (rc == 0) rc = put(f, "velocity", 2, d2, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, velocity, -1);
  if (rc == 0) rc = put(f, "id", 1, d1, H5T_NATIVE_INT, H5T_NATIVE_INT, id, -1);
  if (rc == 0 && dm) rc = put(f, "Dm", 1, d1, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, dm, -1);
  if (rc == 0 && plane_id) rc = put(f, "plane_id", 1, d1, H5T_NATIVE_INT, H5T_NATIVE_INT, plane_id, -1);
  if (rc == 0 && rho_hat) rc = put(f, "rho_hat", 2, d2, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, rho_hat, -1);
  if (rc == 0 && log_rho) rc = put(f, "log_rho", 1, d1, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, log_rho, -1);
  if (H5Fclose(f) < 0 && rc == 0) rc = fail(FELBM_H5_EIO, "closing %s failed", path);
  return rc;
}

int felbm_h5_dataset_shape(const char *path, const char *name, int64_t *dims, int *ndims) {
  if (!path || !name || !dims || !ndims) return fail(FELBM_H5_EARG, "null argument");
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot open %s", path);
  int rc = FELBM_H5_OK;
  *ndims = 0;
  if (H5Lexists(f, name, H5P_DEFAULT) <= 0) {
    rc = fail(FELBM_H5_EMISSING, "no dataset %s in %s", name, path);
  } else {
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    hid_t sp = ds < 0 ? -1 : H5Dget_space(ds);
    int nd = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
    hsize_t d[8];
    if (nd < 0 || nd > 8 || H5Sget_simple_extent_dims(sp, d, NULL) != nd) {
      rc = fail(FELBM_H5_EIO, "cannot read the shape of %s", name);
    } else {
      *ndims = nd;
      for (int i = 0; i < nd && i < 2; ++i) dims[i] = (int64_t)d[i];
    }
    if (sp >= 0) H5Sclose(sp);
    if (ds >= 0) H5Dclose(ds);
  }
  H5Fclose(f);
  return rc;
}

int felbm_h5_read_dataset(const char *path, const char *name, int type, void *buf, int64_t count) {
  if (!path || !name || count < 0 || (count > 0 && !buf) || type < 0 || type > 2)
    return fail(FELBM_H5_EARG, "bad path, name, type or buffer");
  quiet();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail(FELBM_H5_EIO, "cannot open %s", path);
  const hid_t mt = type == FELBM_H5_F64 ? H5T_NATIVE_DOUBLE : type == FELBM_H5_F32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_INT;
  int rc = H5Lexists(f, name, H5P_DEFAULT) > 0 ? get(f, name, mt, buf, (hsize_t)count)
                                               : fail(FELBM_H5_EMISSING, "no dataset %s in %s", name, path);
  H5Fclose(f);
  return rc;
}
