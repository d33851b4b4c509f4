/* pypde_amd state files: the reference's HDF5 layout through the HDF5 C library.
 *
 * Library: pypde_amd/libpypde_h5.so (host code only; links libhdf5).  Replaces
 * the h5py calls of the reference's field IO:
 *   - h5py.File(filename, "a" | "r")            pypde/field.py:106,157
 *   - write_single_hdf5 (create or overwrite)    pypde/field.py:125-133
 *   - read_single_hdf5                            pypde/field.py:135-144
 *   - hf.visit(print_grp)                         pypde/field.py:143, 175-176
 * Datasets are contiguous little-endian float64, 0-d values get a scalar
 * dataspace, and missing groups are created on the way, as h5py's
 * create_dataset does.  Arrays are row-major host memory.
 *
 * Return codes: 0 = ok, PYPDE_H5_MISSING = no such dataset (not an error
 * condition: the reference prints the file's keys and reads 0.0), negative =
 * error with pypde_h5_last_error() giving the message.
 */
#ifndef PYPDE_AMD_H5_H
#define PYPDE_AMD_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYPDE_H5_READ 0   /* h5py mode "r" */
#define PYPDE_H5_APPEND 1 /* h5py mode "a": read/write, create if missing */

#define PYPDE_H5_MISSING 1
#define PYPDE_H5_EINVAL (-1)
#define PYPDE_H5_EIO (-2)
#define PYPDE_H5_ESHAPE (-3)

/* Open a file; returns a handle > 0 or a negative code (pypde/field.py:106,157). */
int64_t pypde_h5_open(const char* filename, int mode);
int pypde_h5_close(int64_t handle);

/* Create dataset `name` (e.g. "temp/vhat") with the given extents, or
 * overwrite an existing dataset of the same shape (pypde/field.py:125-133).
 * ndim = 0 writes a scalar. */
int pypde_h5_write(int64_t handle, const char* name, const double* data, int ndim,
                   const int64_t* dims);

/* Rank and extents (dims holds up to 8) of dataset `name`; PYPDE_H5_MISSING
 * if it does not exist. */
int pypde_h5_shape(int64_t handle, const char* name, int* ndim, int64_t* dims);

/* Read dataset `name` (count elements) as float64 (pypde/field.py:135-144). */
int pypde_h5_read(int64_t handle, const char* name, double* out, int64_t count);

/* Every link name in the file, '\n'-terminated, in name order (hf.visit,
 * pypde/field.py:143).  Returns the buffer size needed (including the final
 * NUL); the buffer receives as many whole names as fit. */
int64_t pypde_h5_list(int64_t handle, char* buf, int64_t bufsize);

const char* pypde_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
