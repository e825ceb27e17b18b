"""HDF5 interchange of representatives and vectors (liblsamd_hdf5.so,
include/lsamd_hdf5.h).

One function per procedure of the reference's chapel/src/MyHDF5.chpl, with
the same arguments, dataset paths, element types (uint(64) -> numpy.uint64,
real(64) -> numpy.float64), rank and type checks and create / replace rules,
so files written by the reference are read here and the other way round.
Where the reference halts, these raise HDF5Error with the reference's message.

The *_as_blocks procedures distribute a dataset over Chapel locales; on one
GPU there is one block, so they read or write the whole dataset, streamed in
slabs (`slab_bytes`) so that a vector in device memory never needs a second
full copy in host memory.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from ._abi import PACKAGE_ROOT, EngineError, load as _load_engine

LIBRARY_PATH = os.path.join(PACKAGE_ROOT, "liblsamd_hdf5.so")
U64, F64 = 0, 1
_TYPES = {np.dtype(np.uint64): U64, np.dtype(np.float64): F64}
_u64p = C.POINTER(C.c_uint64)

_lib = None


class HDF5Error(EngineError):
    pass


def load():
    """Loads liblsamd_hdf5.so; fails loudly when it was not built (the HDF5
    headers were absent at build time) or libhdf5 cannot be loaded."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise HDF5Error(f"{LIBRARY_PATH} does not exist: build it with `make -C lattice-symmetries_amd` "
                        "where the HDF5 headers are installed (HDF5_PREFIX)")
    try:
        lib = C.CDLL(LIBRARY_PATH)
    except OSError as e:
        raise HDF5Error(f"cannot load {LIBRARY_PATH}: {e}") from e
    s, i = C.c_char_p, C.c_int
    sig = {
        "ls_amd_hdf5_last_error": (C.c_char_p, []),
        "ls_amd_hdf5_dataset_rank": (i, [s, s]),
        "ls_amd_hdf5_dataset_shape": (i, [s, s, i, _u64p]),
        "ls_amd_hdf5_read_chunk": (i, [s, s, i, i, _u64p, _u64p, C.c_void_p]),
        "ls_amd_hdf5_write_chunk": (i, [s, s, i, i, _u64p, _u64p, C.c_void_p]),
        "ls_amd_hdf5_create_dataset": (i, [s, s, i, i, _u64p]),
        "ls_amd_hdf5_write_dataset": (i, [s, s, i, i, _u64p, C.c_void_p]),
        "ls_amd_hdf5_exists": (i, [s, s]),
        "ls_amd_hdf5_delete": (i, [s, s]),
        "ls_amd_hdf5_make_group": (i, [s, s]),
        "ls_amd_hdf5_make_file": (i, [s]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def available() -> bool:
    try:
        load()
        return True
    except HDF5Error:
        return False


def _b(x: str) -> bytes:
    return os.fsencode(x)


def _ok(rc: int) -> int:
    if rc < 0:
        raise HDF5Error((_lib.ls_amd_hdf5_last_error() or b"").decode("utf-8", "replace"))
    return rc


def _type_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt not in _TYPES:
        raise HDF5Error(f"unsupported element type {dt}: the files hold uint64 or float64")
    return _TYPES[dt]


def _dims(xs: Sequence[int]):
    a = np.ascontiguousarray(np.asarray(list(xs), dtype=np.uint64).reshape(-1))
    return a, a.ctypes.data_as(_u64p)


def dataset_rank(filename: str, dataset: str) -> int:
    """The commented-out datasetRank (MyHDF5.chpl:12-15)."""
    return _ok(load().ls_amd_hdf5_dataset_rank(_b(filename), _b(dataset)))


def dataset_shape(filename: str, dataset: str, rank: int) -> tuple:
    """datasetShape (MyHDF5.chpl:26-45)."""
    shape = np.zeros(max(rank, 1), dtype=np.uint64)
    _ok(load().ls_amd_hdf5_dataset_shape(_b(filename), _b(dataset), rank, shape.ctypes.data_as(_u64p)))
    return tuple(int(v) for v in shape[:rank])


def read_dataset_chunk(filename: str, dataset: str, offset: Sequence[int], out: Optional[np.ndarray] = None,
                       dtype=None, shape: Optional[Sequence[int]] = None) -> np.ndarray:
    """readDatasetChunk: into `out` (MyHDF5.chpl:105-144), or a new array of
    `dtype` and `shape` (MyHDF5.chpl:157-162)."""
    if out is None:
        if dtype is None or shape is None:
            raise ValueError("give out, or dtype and shape")
        out = np.empty(tuple(shape), dtype=dtype)
    if not out.flags.c_contiguous:
        raise ValueError("out must be C-contiguous")
    if len(offset) != out.ndim:
        raise ValueError("offset and array ranks differ")
    off, offp = _dims(offset)
    shp, shpp = _dims(out.shape)
    _ok(load().ls_amd_hdf5_read_chunk(_b(filename), _b(dataset), _type_code(out.dtype), out.ndim, offp, shpp,
                                      out.ctypes.data))
    return out


def read_dataset(filename: str, dataset: str, dtype, rank: int) -> np.ndarray:
    """readDataset (MyHDF5.chpl:71-103)."""
    r = dataset_rank(filename, dataset)
    if r != rank:
        raise HDF5Error(f"rank mismatch in file: '{filename}' dataset: '{dataset}'  {rank} != {r}")
    shape = dataset_shape(filename, dataset, rank)
    return read_dataset_chunk(filename, dataset, (0,) * rank, dtype=dtype, shape=shape)


def read_dataset_as_blocks(filename: str, dataset: str, rank: int = 2, dtype=np.float64) -> np.ndarray:
    """readDatasetAsBlocks (MyHDF5.chpl:272-286) with one block."""
    return read_dataset(filename, dataset, dtype, rank)


def write_dataset_chunk(filename: str, dataset: str, offset: Sequence[int], arr: np.ndarray) -> None:
    """writeDatasetChunk (MyHDF5.chpl:215-252): into an existing dataset."""
    arr = np.ascontiguousarray(arr)
    if len(offset) != arr.ndim:
        raise ValueError("offset and array ranks differ")
    off, offp = _dims(offset)
    shp, shpp = _dims(arr.shape)
    _ok(load().ls_amd_hdf5_write_chunk(_b(filename), _b(dataset), _type_code(arr.dtype), arr.ndim, offp, shpp,
                                       arr.ctypes.data))


def create_dataset(filename: str, dataset: str, dtype, shape: Sequence[int]) -> None:
    """The create step of writeDatasetAsBlocks (MyHDF5.chpl:309-324): replaces
    an existing object by a new contiguous dataset."""
    shp, shpp = _dims(shape)
    _ok(load().ls_amd_hdf5_create_dataset(_b(filename), _b(dataset), _type_code(dtype), len(shape), shpp))


def write_dataset(filename: str, dataset: str, arr: np.ndarray) -> None:
    """writeDataset (MyHDF5.chpl:288-308)."""
    arr = np.ascontiguousarray(arr)
    shp, shpp = _dims(arr.shape)
    _ok(load().ls_amd_hdf5_write_dataset(_b(filename), _b(dataset), _type_code(arr.dtype), arr.ndim, shpp,
                                         arr.ctypes.data if arr.size else None))


def _slabs(shape: Sequence[int], itemsize: int, slab_bytes: int):
    """(leading index, start, count) slabs along the last axis of a row-major
    array, each at most slab_bytes."""
    shape = tuple(int(v) for v in shape)
    last = shape[-1] if shape else 1
    step = max(1, slab_bytes // itemsize)
    lead = list(np.ndindex(*shape[:-1])) if len(shape) > 1 else [()]
    for idx in lead:
        for a in range(0, last, step):
            yield idx, a, min(step, last - a)


def write_dataset_as_blocks(filename: str, dataset: str, arr: np.ndarray, slab_bytes: int = 1 << 28) -> None:
    """writeDatasetAsBlocks (MyHDF5.chpl:309-333) with one block: create, then
    write in slabs along the last axis."""
    arr = np.ascontiguousarray(arr)
    create_dataset(filename, dataset, arr.dtype, arr.shape)
    if arr.size == 0:
        return
    for idx, a, n in _slabs(arr.shape, arr.itemsize, slab_bytes):
        write_dataset_chunk(filename, dataset, idx + (a,), arr[idx + (slice(a, a + n),)].reshape((1,) * len(idx) + (n,)))


def write_device_dataset(filename: str, dataset: str, dev, dtype, shape: Sequence[int],
                         slab_bytes: int = 1 << 28) -> None:
    """writeDatasetAsBlocks of a row-major array in device memory (a
    DeviceArray): each slab is copied to host memory and written, so host
    memory holds one slab at a time."""
    dt = np.dtype(dtype)
    shape = tuple(int(v) for v in shape)
    if not shape:
        raise ValueError("device arrays are written as datasets of rank 1 or more")
    total = int(np.prod(shape)) * dt.itemsize
    if total > dev.nbytes:
        raise ValueError("shape exceeds the device allocation")
    create_dataset(filename, dataset, dt, shape)
    if total == 0:
        return
    lib = _load_engine()
    strides = np.cumprod((1,) + shape[::-1])[::-1][1:]
    buf = np.empty(max(1, min(slab_bytes // dt.itemsize, shape[-1])), dtype=dt)
    for idx, a, n in _slabs(shape, dt.itemsize, slab_bytes):
        elem = int(sum(int(i) * int(s) for i, s in zip(idx, strides[:-1]))) + a
        if lib.ls_amd_memcpy_d2h(buf.ctypes.data, dev.ptr + elem * dt.itemsize, n * dt.itemsize) != 0:
            raise EngineError("device to host copy failed")
        write_dataset_chunk(filename, dataset, idx + (a,), buf[:n].reshape((1,) * len(idx) + (n,)))


def read_device_dataset(filename: str, dataset: str, dtype, rank: int, slab_bytes: int = 1 << 28):
    """readDatasetAsBlocks into device memory: returns (DeviceArray, shape),
    read and uploaded slab by slab."""
    from . import DeviceArray
    dt = np.dtype(dtype)
    if rank < 1:
        raise ValueError("device arrays are read from datasets of rank 1 or more")
    shape = dataset_shape(filename, dataset, rank)
    total = int(np.prod(shape)) * dt.itemsize
    dev = DeviceArray(total)
    if total == 0:
        return dev, shape
    lib = _load_engine()
    strides = np.cumprod((1,) + shape[::-1])[::-1][1:]
    buf = np.empty(max(1, min(slab_bytes // dt.itemsize, shape[-1])), dtype=dt)
    for idx, a, n in _slabs(shape, dt.itemsize, slab_bytes):
        read_dataset_chunk(filename, dataset, idx + (a,), out=buf[:n].reshape((1,) * len(idx) + (n,)))
        elem = int(sum(int(i) * int(s) for i, s in zip(idx, strides[:-1]))) + a
        if lib.ls_amd_memcpy_h2d(dev.ptr + elem * dt.itemsize, buf.ctypes.data, n * dt.itemsize) != 0:
            raise EngineError("host to device copy failed")
    return dev, shape


def does_object_exist(filename: str, obj: str) -> bool:
    """doesObjectExist (MyHDF5.chpl:363-374)."""
    return _ok(load().ls_amd_hdf5_exists(_b(filename), _b(obj))) == 1


def delete_object(filename: str, obj: str) -> None:
    """deleteObject (MyHDF5.chpl:352-361)."""
    _ok(load().ls_amd_hdf5_delete(_b(filename), _b(obj)))


def make_group(filename: str, group: str) -> None:
    """makeGroup (MyHDF5.chpl:376-391)."""
    _ok(load().ls_amd_hdf5_make_group(_b(filename), _b(group)))


def make_file(filename: str) -> None:
    """makeFile (MyHDF5.chpl:393-396)."""
    _ok(load().ls_amd_hdf5_make_file(_b(filename)))
