/* HDF5 interchange of representatives and vectors: the procedures of
 * chapel/src/MyHDF5.chpl on one process, over libhdf5's C API.
 *
 * Every call opens the file, does its work and closes every handle it opened
 * (the reference's `defer` pattern), so no HDF5 state outlives a call.  The
 * reference halts on an error; here the message goes to a per-thread buffer
 * (ls_amd_hdf5_last_error) and the call returns -1.  libhdf5's own error
 * stack printing is switched off once, at the first call, so expected
 * failures (a missing object) do not print to stderr.
 */
#include "lsamd_hdf5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#define LS_AMD_HDF5_MAX_RANK 32

static __thread char g_error[1024];

static int fail(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
  return -1;
}

const char *ls_amd_hdf5_last_error(void) { return g_error; }

static void quiet(void) {
  static int done = 0;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    done = 1;
  }
  g_error[0] = '\0';
}

static hid_t native_type(int type) {
  switch (type) {
  case LS_AMD_HDF5_U64: return H5T_NATIVE_UINT64;
  case LS_AMD_HDF5_F64: return H5T_NATIVE_DOUBLE;
  default: return H5I_INVALID_HID;
  }
}

static const char *type_name(int type) { return type == LS_AMD_HDF5_U64 ? "uint(64)" : "real(64)"; }

/* A short description of a stored type for the type-mismatch message. */
static void describe_type(hid_t t, char *buf, size_t n) {
  H5T_class_t c = H5Tget_class(t);
  size_t sz = H5Tget_size(t);
  const char *k = "other";
  if (c == H5T_INTEGER) k = H5Tget_sign(t) == H5T_SGN_NONE ? "uint" : "int";
  else if (c == H5T_FLOAT) k = "real";
  else if (c == H5T_STRING) k = "string";
  else if (c == H5T_COMPOUND) k = "compound";
  snprintf(buf, n, "%s(%zu)", k, sz * 8);
}

/* openFile (MyHDF5.chpl:335-344): opens an existing file with `mode`, or
   creates it (H5F_ACC_EXCL) when it does not exist. */
static hid_t open_file(const char *filename, unsigned mode) {
  hid_t f = access(filename, F_OK) == 0 ? H5Fopen(filename, mode, H5P_DEFAULT)
                                        : H5Fcreate(filename, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) fail("HDF5 error: could not open file '%s'", filename);
  return f;
}

static int check_args(const char *filename, const char *name, int rank) {
  if (filename == NULL || name == NULL) return fail("null argument");
  if (rank < 0 || rank > LS_AMD_HDF5_MAX_RANK) return fail("rank %d out of range [0, %d]", rank, LS_AMD_HDF5_MAX_RANK);
  return 0;
}

/* The open dataset with its dataspace, rank- and type-checked as readDataset
   and writeDatasetChunk do (MyHDF5.chpl:82-91, 227-236). */
typedef struct {
  hid_t file, dset, space;
} opened_t;

static void close_opened(opened_t *o) {
  if (o->space >= 0) H5Sclose(o->space);
  if (o->dset >= 0) H5Dclose(o->dset);
  if (o->file >= 0) H5Fclose(o->file);
}

static int open_checked(const char *filename, const char *dataset, int type, int rank, int writable,
                        opened_t *o) {
  o->file = o->dset = o->space = H5I_INVALID_HID;
  if (writable) {
    o->file = open_file(filename, H5F_ACC_RDWR);
    if (o->file < 0) return -1;
  } else {
    o->file = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (o->file < 0) return fail("HDF5 error: could not open file '%s'", filename);
  }
  o->dset = H5Dopen2(o->file, dataset, H5P_DEFAULT);
  if (o->dset < 0) {
    close_opened(o);
    return fail("HDF5 error: could not open dataset '%s'", dataset);
  }
  o->space = H5Dget_space(o->dset);
  int r = o->space < 0 ? -1 : H5Sget_simple_extent_ndims(o->space);
  if (r != rank) {
    close_opened(o);
    return fail("rank mismatch in file: '%s' dataset: '%s'  %d != %d", filename, dataset, rank, r);
  }
  hid_t stored = H5Dget_type(o->dset);
  htri_t eq = stored < 0 ? -1 : H5Tequal(native_type(type), stored);
  if (eq <= 0) {
    char desc[64] = "unknown";
    if (stored >= 0) describe_type(stored, desc, sizeof desc);
    if (stored >= 0) H5Tclose(stored);
    close_opened(o);
    return fail("type mismatch in file: '%s' dataset: '%s'  %s != %s", filename, dataset, type_name(type), desc);
  }
  H5Tclose(stored);
  return 0;
}

/* Selects offset + shape in the file space after checking that it lies
   inside the dataset. */
static int select_slab(opened_t *o, const char *dataset, int rank, const uint64_t *offset, const uint64_t *shape,
                       hid_t *mem) {
  hsize_t dims[LS_AMD_HDF5_MAX_RANK], off[LS_AMD_HDF5_MAX_RANK], cnt[LS_AMD_HDF5_MAX_RANK];
  H5Sget_simple_extent_dims(o->space, dims, NULL);
  for (int i = 0; i < rank; ++i) {
    off[i] = offset[i];
    cnt[i] = shape[i];
    if (offset[i] > dims[i] || shape[i] > dims[i] - offset[i])
      return fail("hyperslab [%llu, %llu) of dimension %d is outside dataset '%s' of extent %llu",
                  (unsigned long long)offset[i], (unsigned long long)(offset[i] + shape[i]), i, dataset,
                  (unsigned long long)dims[i]);
  }
  *mem = H5Screate_simple(rank, cnt, NULL);
  if (*mem < 0) return fail("HDF5 error: could not create a memory dataspace");
  if (rank > 0 && H5Sselect_hyperslab(o->space, H5S_SELECT_SET, off, NULL, cnt, NULL) < 0) {
    H5Sclose(*mem);
    return fail("HDF5 error: could not select the hyperslab in dataset '%s'", dataset);
  }
  return 0;
}

static int has_zero(int rank, const uint64_t *shape) {
  for (int i = 0; i < rank; ++i)
    if (shape[i] == 0) return 1;
  return 0;
}

int ls_amd_hdf5_dataset_rank(const char *filename, const char *dataset) {
  quiet();
  if (check_args(filename, dataset, 0) != 0) return -1;
  hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail("HDF5 error: could not open file '%s'", filename);
  hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
  if (d < 0) {
    H5Fclose(f);
    return fail("HDF5 error: could not open dataset '%s'", dataset);
  }
  hid_t s = H5Dget_space(d);
  int r = s < 0 ? -1 : H5Sget_simple_extent_ndims(s);
  if (s >= 0) H5Sclose(s);
  H5Dclose(d);
  H5Fclose(f);
  if (r < 0) return fail("HDF5 error: could not read the dataspace of '%s'", dataset);
  return r;
}

int ls_amd_hdf5_dataset_shape(const char *filename, const char *dataset, int rank, uint64_t *shape) {
  quiet();
  if (check_args(filename, dataset, rank) != 0) return -1;
  if (shape == NULL && rank > 0) return fail("null argument");
  int r = ls_amd_hdf5_dataset_rank(filename, dataset);
  if (r < 0) return -1;
  if (r != rank) return fail("expected '%s' to be %d-dimensional, but it has rank %d", dataset, rank, r);
  hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
  hid_t d = f < 0 ? H5I_INVALID_HID : H5Dopen2(f, dataset, H5P_DEFAULT);
  hid_t s = d < 0 ? H5I_INVALID_HID : H5Dget_space(d);
  hsize_t dims[LS_AMD_HDF5_MAX_RANK];
  int ok = s >= 0 && H5Sget_simple_extent_dims(s, dims, NULL) == rank;
  if (s >= 0) H5Sclose(s);
  if (d >= 0) H5Dclose(d);
  if (f >= 0) H5Fclose(f);
  if (!ok) return fail("HDF5 error: could not read the shape of '%s'", dataset);
  for (int i = 0; i < rank; ++i) shape[i] = dims[i];
  return 0;
}

int ls_amd_hdf5_read_chunk(const char *filename, const char *dataset, int type, int rank, const uint64_t *offset,
                           const uint64_t *shape, void *out) {
  quiet();
  if (check_args(filename, dataset, rank) != 0) return -1;
  if (native_type(type) < 0) return fail("unknown element type %d", type);
  if (rank > 0 && (offset == NULL || shape == NULL)) return fail("null argument");
  opened_t o;
  if (open_checked(filename, dataset, type, rank, 0, &o) != 0) return -1;
  hid_t mem;
  if (select_slab(&o, dataset, rank, offset, shape, &mem) != 0) {
    close_opened(&o);
    return -1;
  }
  herr_t e = 0;
  if (!has_zero(rank, shape)) {
    if (out == NULL) e = -1;
    else e = H5Dread(o.dset, native_type(type), mem, o.space, H5P_DEFAULT, out);
  }
  H5Sclose(mem);
  close_opened(&o);
  if (e < 0) return fail("HDF5 error: could not read dataset '%s'", dataset);
  return 0;
}

int ls_amd_hdf5_write_chunk(const char *filename, const char *dataset, int type, int rank, const uint64_t *offset,
                            const uint64_t *shape, const void *data) {
  quiet();
  if (check_args(filename, dataset, rank) != 0) return -1;
  if (native_type(type) < 0) return fail("unknown element type %d", type);
  if (rank > 0 && (offset == NULL || shape == NULL)) return fail("null argument");
  opened_t o;
  if (open_checked(filename, dataset, type, rank, 1, &o) != 0) return -1;
  hid_t mem;
  if (select_slab(&o, dataset, rank, offset, shape, &mem) != 0) {
    close_opened(&o);
    return -1;
  }
  herr_t e = 0;
  if (!has_zero(rank, shape)) {
    if (data == NULL) e = -1;
    else e = H5Dwrite(o.dset, native_type(type), mem, o.space, H5P_DEFAULT, data);
  }
  H5Sclose(mem);
  close_opened(&o);
  if (e < 0) return fail("HDF5 error: could not write array to dataset %s", dataset);
  return 0;
}

/* Deletes `name` from an open file if the link exists (writeDataset,
   MyHDF5.chpl:291-292). */
static int delete_if_exists(hid_t f, const char *name) {
  htri_t ex = H5Lexists(f, name, H5P_DEFAULT);
  if (ex < 0) return fail("HDF5 error: could not check whether group '%s' exists", name);
  if (ex > 0 && H5Ldelete(f, name, H5P_DEFAULT) < 0) return fail("HDF5 error: could not delete '%s'", name);
  return 0;
}

static hid_t create_in(hid_t f, const char *dataset, int type, int rank, const uint64_t *shape) {
  hsize_t dims[LS_AMD_HDF5_MAX_RANK];
  for (int i = 0; i < rank; ++i) dims[i] = shape[i];
  hid_t space = H5Screate_simple(rank, dims, NULL);
  if (space < 0) {
    fail("HDF5 error: could not create a dataspace for '%s'", dataset);
    return H5I_INVALID_HID;
  }
  hid_t d = H5Dcreate2(f, dataset, native_type(type), space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5Sclose(space);
  if (d < 0) fail("HDF5 error: could not create dataset '%s'", dataset);
  return d;
}

int ls_amd_hdf5_create_dataset(const char *filename, const char *dataset, int type, int rank,
                               const uint64_t *shape) {
  quiet();
  if (check_args(filename, dataset, rank) != 0) return -1;
  if (native_type(type) < 0) return fail("unknown element type %d", type);
  if (rank > 0 && shape == NULL) return fail("null argument");
  hid_t f = open_file(filename, H5F_ACC_RDWR);
  if (f < 0) return -1;
  int rc = delete_if_exists(f, dataset);
  if (rc == 0) {
    hid_t d = create_in(f, dataset, type, rank, shape);
    if (d < 0) rc = -1;
    else H5Dclose(d);
  }
  if (rc == 0 && H5Fflush(f, H5F_SCOPE_LOCAL) < 0) rc = fail("HDF5 error: could not flush '%s'", filename);
  H5Fclose(f);
  return rc;
}

int ls_amd_hdf5_write_dataset(const char *filename, const char *dataset, int type, int rank, const uint64_t *shape,
                              const void *data) {
  quiet();
  if (check_args(filename, dataset, rank) != 0) return -1;
  if (native_type(type) < 0) return fail("unknown element type %d", type);
  if (rank > 0 && shape == NULL) return fail("null argument");
  if (data == NULL && !has_zero(rank, shape)) return fail("null argument");
  hid_t f = open_file(filename, H5F_ACC_RDWR);
  if (f < 0) return -1;
  int rc = delete_if_exists(f, dataset);
  if (rc == 0) {
    hid_t d = create_in(f, dataset, type, rank, shape);
    if (d < 0) {
      rc = -1;
    } else {
      if (!has_zero(rank, shape) && H5Dwrite(d, native_type(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        rc = fail("HDF5 error: could not write array to dataset %s", dataset);
      H5Dclose(d);
    }
  }
  H5Fclose(f);
  return rc;
}

int ls_amd_hdf5_exists(const char *filename, const char *object) {
  quiet();
  if (check_args(filename, object, 0) != 0) return -1;
  hid_t f = open_file(filename, H5F_ACC_RDONLY);
  if (f < 0) return -1;
  htri_t ex = H5Lexists(f, object, H5P_DEFAULT);
  H5Fclose(f);
  if (ex < 0) return fail("HDF5 error: could not check whether group '%s' exists", object);
  return ex > 0 ? 1 : 0;
}

int ls_amd_hdf5_delete(const char *filename, const char *object) {
  quiet();
  if (check_args(filename, object, 0) != 0) return -1;
  hid_t f = open_file(filename, H5F_ACC_RDWR);
  if (f < 0) return -1;
  herr_t e = H5Ldelete(f, object, H5P_DEFAULT);
  H5Fclose(f);
  if (e < 0) return fail("HDF5 error: could not delete '%s'", object);
  return 0;
}

int ls_amd_hdf5_make_group(const char *filename, const char *group) {
  quiet();
  if (check_args(filename, group, 0) != 0) return -1;
  hid_t f = open_file(filename, H5F_ACC_RDWR);
  if (f < 0) return -1;
  int rc = 0;
  htri_t ex = H5Lexists(f, group, H5P_DEFAULT);
  if (ex < 0) {
    rc = fail("HDF5 error: could not check whether group '%s' exists", group);
  } else if (ex == 0) {
    hid_t g = H5Gcreate2(f, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) rc = fail("HDF5 error: could not create group '%s'", group);
    else H5Gclose(g);
  }
  H5Fclose(f);
  return rc;
}

int ls_amd_hdf5_make_file(const char *filename) {
  quiet();
  if (filename == NULL) return fail("null argument");
  hid_t f = open_file(filename, H5F_ACC_RDWR);
  if (f < 0) return -1;
  H5Fclose(f);
  return 0;
}
