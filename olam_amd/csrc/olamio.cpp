// olamio.cpp -- named HDF5 datasets, the storage layer of OLAM's gridfiles and history
// files (gridfile_write omodel/olam_grid.f90:828-1450, history_write
// omodel/history_write.f90:1-190).
//
// OLAM writes every array with the Fortran HDF5 interface, which stores a Fortran array
// of shape (n1, n2, ...) as a dataset of C shape (..., n2, n1): the bytes in the file are
// the array's memory.  The functions below take and return C-order shapes, so a host
// array in the reference's Fortran layout (mza, mwa), held by numpy as (mwa, mza), is
// written and read as it is.  Datasets live in the file's root group under the
// reference's names (e.g. "ZT", "itab_w%iv", "THIL").  Element types: 0 int32,
// 1 float32, 2 float64 (Fortran integer, real, real(r8)).
//
// Every call returns 0 (or a non-negative handle / flag) on success and a negative value
// on failure; HDF5's own error printing is switched off.
#include <hdf5.h>

#include <cstdint>

namespace {

hid_t mem_type(int type) {
  switch (type) {
    case 0: return H5T_NATIVE_INT32;
    case 1: return H5T_NATIVE_FLOAT;
    case 2: return H5T_NATIVE_DOUBLE;
    default: return -1;
  }
}

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

}  // namespace

extern "C" {

// Create (truncate) a file for writing; returns its handle.
int64_t olamio_create(const char *path) {
  quiet();
  return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
}

// Open an existing file, read-only (write = 0) or read-write; returns its handle.
int64_t olamio_open(const char *path, int write) {
  quiet();
  return H5Fopen(path, write ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
}

int olamio_close(int64_t f) { return H5Fclose((hid_t)f) < 0 ? -1 : 0; }

// Write one dataset of the given C-order shape (rank 1..4) from contiguous memory.
int olamio_write(int64_t f, const char *name, int type, int rank, const int64_t *dims, const void *data) {
  const hid_t mt = mem_type(type);
  if (mt < 0 || rank < 1 || rank > 4) return -1;
  hsize_t d[4];
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 1) return -1;
    d[i] = (hsize_t)dims[i];
  }
  const hid_t sp = H5Screate_simple(rank, d, nullptr);
  if (sp < 0) return -1;
  const hid_t ds = H5Dcreate2((hid_t)f, name, mt, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = -1;
  if (ds >= 0) {
    rc = H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0;
    H5Dclose(ds);
  }
  H5Sclose(sp);
  return rc;
}

// 1 if the dataset exists, 0 if not.
int olamio_exists(int64_t f, const char *name) { return H5Lexists((hid_t)f, name, H5P_DEFAULT) > 0 ? 1 : 0; }

// Shape (C order, up to 4 dimensions) and element type of a dataset: type 0 integer,
// 1 4-byte float, 2 8-byte float.  Returns the rank.
int olamio_info(int64_t f, const char *name, int *type, int64_t *dims) {
  const hid_t ds = H5Dopen2((hid_t)f, name, H5P_DEFAULT);
  if (ds < 0) return -1;
  const hid_t sp = H5Dget_space(ds);
  const hid_t ft = H5Dget_type(ds);
  int rank = H5Sget_simple_extent_ndims(sp);
  if (rank >= 0 && rank <= 4) {
    hsize_t d[4] = {0, 0, 0, 0};
    H5Sget_simple_extent_dims(sp, d, nullptr);
    for (int i = 0; i < rank; ++i) dims[i] = (int64_t)d[i];
    const H5T_class_t cls = H5Tget_class(ft);
    if (cls == H5T_INTEGER) *type = 0;
    else if (cls == H5T_FLOAT) *type = H5Tget_size(ft) == 8 ? 2 : 1;
    else rank = -1;
  } else {
    rank = -1;
  }
  H5Tclose(ft);
  H5Sclose(sp);
  H5Dclose(ds);
  return rank;
}

// Read a whole dataset into contiguous memory, converted to the given element type.
int olamio_read(int64_t f, const char *name, int type, void *data) {
  const hid_t mt = mem_type(type);
  if (mt < 0) return -1;
  const hid_t ds = H5Dopen2((hid_t)f, name, H5P_DEFAULT);
  if (ds < 0) return -1;
  const int rc = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0;
  H5Dclose(ds);
  return rc;
}

}  // extern "C"
