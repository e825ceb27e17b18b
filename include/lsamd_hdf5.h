/* HDF5 interchange of representatives and vectors (liblsamd_hdf5.so).
 *
 * The reference reads and writes its basis representatives (u64, rank 1) and
 * vectors (f64, rank 2: [number of vectors][basis dimension]) through the
 * Chapel module chapel/src/MyHDF5.chpl, on top of libhdf5.  Each entry point
 * below is one of that module's procedures, on one process: the same dataset
 * paths, element types, rank and type checks, and the same create / replace
 * rules, so files written by either side are read by the other.  The
 * reference halts on an error; these return -1 and leave the reference's
 * message in ls_amd_hdf5_last_error().
 *
 * Users of the format in the reference:
 *   chapel/test/TestStatesEnumeration.chpl:22-25   "/representatives", u64, rank 1
 *   chapel/test/TestMatrixVectorProduct.chpl:34,43 "/x", "/y", f64, rank 2
 *   chapel/src/Diagonalize.chpl:227-255,276-279    "basis/representatives",
 *       "hamiltonian/eigenvectors" (rank 2), "hamiltonian/eigenvalues",
 *       "hamiltonian/residuals" (rank 1), groups basis, hamiltonian, observables
 *
 * The library is plain C on libhdf5 (no GPU code); it is built only where the
 * HDF5 headers exist.  Calls are serialised by libhdf5 (thread-safe build).
 */
#ifndef LSAMD_HDF5_H
#define LSAMD_HDF5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types: the native types the reference maps uint(64) and real(64)
   to (HDF5.getHDF5Type); a dataset of any other stored type is a type
   mismatch, as in MyHDF5.chpl:87-91. */
enum { LS_AMD_HDF5_U64 = 0, LS_AMD_HDF5_F64 = 1 };

/* Message of the last failed call on this thread ("" if none). */
const char *ls_amd_hdf5_last_error(void);

/* Rank of a dataset, or -1.  The commented-out datasetRank of
   MyHDF5.chpl:12-15. */
int ls_amd_hdf5_dataset_rank(const char *filename, const char *dataset);

/* datasetShape (MyHDF5.chpl:26-45): shape[rank] of the dataset; fails with
   "expected '<dataset>' to be <rank>-dimensional, but it has rank <r>". */
int ls_amd_hdf5_dataset_shape(const char *filename, const char *dataset, int rank, uint64_t *shape);

/* readDatasetChunk (MyHDF5.chpl:105-144, 157-162): the hyperslab offset[rank] +
   shape[rank] of the dataset into out (row-major, shape[0] * ... *
   shape[rank-1] elements of `type`).  readDataset (MyHDF5.chpl:71-103) is
   this call with offset 0 and the dataset's own shape.  Rank and type are
   checked as there; a hyperslab outside the dataset is an error (the
   reference does not check H5Dread's status). */
int ls_amd_hdf5_read_chunk(const char *filename, const char *dataset, int type, int rank,
                           const uint64_t *offset, const uint64_t *shape, void *out);

/* writeDatasetChunk (MyHDF5.chpl:215-252): writes data (row-major, shape)
   into the existing dataset at offset, opening the file read-write (created
   if absent, as openFile at MyHDF5.chpl:335-344 does). */
int ls_amd_hdf5_write_chunk(const char *filename, const char *dataset, int type, int rank,
                            const uint64_t *offset, const uint64_t *shape, const void *data);

/* The first half of writeDatasetAsBlocks (MyHDF5.chpl:309-333): deletes the
   object if it exists and creates a contiguous dataset of `type` and shape;
   the blocks are then written with ls_amd_hdf5_write_chunk. */
int ls_amd_hdf5_create_dataset(const char *filename, const char *dataset, int type, int rank,
                               const uint64_t *shape);

/* writeDataset (MyHDF5.chpl:288-308): deletes the object if it exists,
   creates the dataset and writes all of data. */
int ls_amd_hdf5_write_dataset(const char *filename, const char *dataset, int type, int rank,
                              const uint64_t *shape, const void *data);

/* doesObjectExist (MyHDF5.chpl:363-374): 1 if the link exists, 0 if not,
   -1 on error (an intermediate group that does not exist is an error, as
   H5Lexists reports it). */
int ls_amd_hdf5_exists(const char *filename, const char *object);

/* deleteObject (MyHDF5.chpl:352-361). */
int ls_amd_hdf5_delete(const char *filename, const char *object);

/* makeGroup (MyHDF5.chpl:376-391): creates the group unless it exists. */
int ls_amd_hdf5_make_group(const char *filename, const char *group);

/* makeFile (MyHDF5.chpl:393-396): creates the file unless it exists. */
int ls_amd_hdf5_make_file(const char *filename);

#ifdef __cplusplus
}
#endif

#endif /* LSAMD_HDF5_H */
