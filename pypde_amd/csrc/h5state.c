/* State files in the reference's HDF5 layout, on the HDF5 C library.
 *
 * The reference writes and reads fields through h5py (pypde/field.py:83-173):
 * one dataset per array, "{group}/v" and "{group}/vhat" per field, "time",
 * "x", "y" and scalar metadata at the root, datasets created on first write
 * and overwritten in place afterwards.  h5py's create_dataset(name, data=...)
 * makes a contiguous little-endian float64 dataset (a scalar dataspace for a
 * 0-d value) and creates missing groups on the way; the functions below make
 * the same objects.  Host memory only: the caller copies device arrays first.
 */
#include <hdf5.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/pypde_amd_h5.h"

#define MAXDIM 8

static char g_err[512];

static int fail(int code, const char* fmt, const char* a) {
  snprintf(g_err, sizeof(g_err), fmt, a ? a : "");
  return code;
}

const char* pypde_h5_last_error(void) { return g_err; }

static void quiet(void) {
  static int done = 0;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    done = 1;
  }
}

/* True if every component of the path exists (H5Lexists fails on a missing
 * intermediate group, so walk the components). */
static int path_exists(hid_t f, const char* name) {
  char buf[1024];
  size_t n = strlen(name);
  if (n == 0 || n >= sizeof(buf)) return 0;
  memcpy(buf, name, n + 1);
  for (size_t i = 1; i <= n; ++i) {
    if (buf[i] == '/' || buf[i] == '\0') {
      char c = buf[i];
      buf[i] = '\0';
      htri_t e = H5Lexists(f, buf, H5P_DEFAULT);
      buf[i] = c;
      if (e <= 0) return 0;
    }
  }
  return 1;
}

int64_t pypde_h5_open(const char* filename, int mode) {
  quiet();
  if (!filename) return fail(PYPDE_H5_EINVAL, "null filename%s", NULL);
  hid_t f;
  if (mode == PYPDE_H5_READ) {
    f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(PYPDE_H5_EIO, "cannot open %s for reading", filename);
    return (int64_t)f;
  }
  if (mode != PYPDE_H5_APPEND) return fail(PYPDE_H5_EINVAL, "unknown mode%s", NULL);
  FILE* probe = fopen(filename, "rb");
  if (probe) {
    fclose(probe);
    f = H5Fopen(filename, H5F_ACC_RDWR, H5P_DEFAULT);
  } else {
    f = H5Fcreate(filename, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  }
  if (f < 0) return fail(PYPDE_H5_EIO, "cannot open or create %s", filename);
  return (int64_t)f;
}

int pypde_h5_close(int64_t h) {
  if (H5Fclose((hid_t)h) < 0) return fail(PYPDE_H5_EIO, "close failed%s", NULL);
  return 0;
}

int pypde_h5_write(int64_t h, const char* name, const double* data, int ndim,
                   const int64_t* dims) {
  if (!name || !data || ndim < 0 || ndim > MAXDIM || (ndim > 0 && !dims))
    return fail(PYPDE_H5_EINVAL, "bad arguments for dataset %s", name);
  hid_t f = (hid_t)h;
  hsize_t d[MAXDIM];
  for (int i = 0; i < ndim; ++i) {
    if (dims[i] < 0) return fail(PYPDE_H5_EINVAL, "negative extent for %s", name);
    d[i] = (hsize_t)dims[i];
  }
  hid_t ds;
  if (path_exists(f, name)) {
    /* write_single_hdf5 (field.py:126-133): overwrite the values in place */
    ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds < 0) return fail(PYPDE_H5_EIO, "%s exists and is not a dataset", name);
    hid_t sp = H5Dget_space(ds);
    int nd = H5Sget_simple_extent_ndims(sp);
    hsize_t cur[MAXDIM];
    int same = (nd == ndim);
    if (same && nd > 0) {
      H5Sget_simple_extent_dims(sp, cur, NULL);
      for (int i = 0; i < nd; ++i) same = same && cur[i] == d[i];
    }
    H5Sclose(sp);
    if (!same) {
      H5Dclose(ds);
      return fail(PYPDE_H5_ESHAPE, "shape of the data differs from dataset %s", name);
    }
  } else {
    hid_t sp = ndim == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(ndim, d, NULL);
    hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    ds = H5Dcreate2(f, name, H5T_IEEE_F64LE, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    H5Pclose(lcpl);
    H5Sclose(sp);
    if (ds < 0) return fail(PYPDE_H5_EIO, "cannot create dataset %s", name);
  }
  herr_t rc = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  if (rc < 0) return fail(PYPDE_H5_EIO, "write failed for %s", name);
  return 0;
}

int pypde_h5_shape(int64_t h, const char* name, int* ndim, int64_t* dims) {
  if (!name || !ndim || !dims) return fail(PYPDE_H5_EINVAL, "bad arguments%s", NULL);
  hid_t f = (hid_t)h;
  if (!path_exists(f, name)) return PYPDE_H5_MISSING;
  hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
  if (ds < 0) return fail(PYPDE_H5_EIO, "%s is not a dataset", name);
  hid_t sp = H5Dget_space(ds);
  int nd = H5Sget_simple_extent_ndims(sp);
  int rc = 0;
  if (nd < 0 || nd > MAXDIM) {
    rc = fail(PYPDE_H5_ESHAPE, "unsupported rank of %s", name);
  } else {
    hsize_t cur[MAXDIM];
    if (nd > 0) H5Sget_simple_extent_dims(sp, cur, NULL);
    *ndim = nd;
    for (int i = 0; i < nd; ++i) dims[i] = (int64_t)cur[i];
  }
  H5Sclose(sp);
  H5Dclose(ds);
  return rc;
}

int pypde_h5_read(int64_t h, const char* name, double* out, int64_t count) {
  if (!name || !out || count < 0) return fail(PYPDE_H5_EINVAL, "bad arguments%s", NULL);
  hid_t f = (hid_t)h;
  if (!path_exists(f, name)) return PYPDE_H5_MISSING;
  hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
  if (ds < 0) return fail(PYPDE_H5_EIO, "%s is not a dataset", name);
  hid_t sp = H5Dget_space(ds);
  hssize_t n = H5Sget_simple_extent_npoints(sp);
  H5Sclose(sp);
  if (n != count) {
    H5Dclose(ds);
    return fail(PYPDE_H5_ESHAPE, "element count of %s differs from the buffer", name);
  }
  /* values of any numeric type are converted to double by the library */
  herr_t rc = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
  H5Dclose(ds);
  if (rc < 0) return fail(PYPDE_H5_EIO, "read failed for %s", name);
  return 0;
}

struct listing {
  char* buf;
  int64_t size, used;
};

static herr_t visit(hid_t g, const char* name, const H5L_info_t* info, void* op) {
  (void)g;
  (void)info;
  struct listing* L = (struct listing*)op;
  int64_t n = (int64_t)strlen(name);
  if (L->buf && L->used + n + 1 < L->size) {
    memcpy(L->buf + L->used, name, (size_t)n);
    L->buf[L->used + n] = '\n';
    L->buf[L->used + n + 1] = '\0';
  }
  L->used += n + 1;
  return 0;
}

int64_t pypde_h5_list(int64_t h, char* buf, int64_t bufsize) {
  struct listing L = {buf, bufsize, 0};
  if (buf && bufsize > 0) buf[0] = '\0';
  if (H5Lvisit((hid_t)h, H5_INDEX_NAME, H5_ITER_INC, visit, &L) < 0)
    return fail(PYPDE_H5_EIO, "cannot list the file%s", NULL);
  return L.used + 1;
}
