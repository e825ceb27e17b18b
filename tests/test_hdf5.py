"""HDF5 interchange on the CPU (liblsamd_hdf5.so, include/lsamd_hdf5.h),
against the procedures of the reference's chapel/src/MyHDF5.chpl.

- The library exports every entry point the header declares.
- The fixtures under golden/hdf5 (written by h5py, in the layout of the
  reference's golden files) are read bit for bit, and their contents are
  pinned by closed forms: the representatives of an unprojected sector are
  all states of the weight, those of a spin-inversion sector are the states
  below their flip, the symmetric sector has its Burnside dimension, and
  y = H x follows from a Heisenberg matrix built here from the model file's
  bond list alone.
- Files written here are read back by h5py (an independent HDF5 reader)
  with the element types, shapes and layout the reference's files have.
- The reference's error behaviour: rank mismatch, type mismatch, missing
  file, dataset or intermediate group, a hyperslab outside the dataset.
"""
import os
import re
import subprocess

import numpy as np
import pytest
import yaml

from lsamd import hdf5 as H

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "hdf5")
MODELS = os.path.join(HERE, "golden", "models")
HEADER = os.path.join(os.path.dirname(HERE), "include", "lsamd_hdf5.h")
PYTHON39 = os.environ.get("PYTHON39", "/opt/conda/bin/python3.9")
NAMES = ["heisenberg_chain_10", "heisenberg_kagome_12", "heisenberg_kagome_12_symm"]

pytestmark = pytest.mark.skipif(not H.available(), reason="liblsamd_hdf5.so not built (no HDF5 headers)")


def _have_h5py() -> bool:
    if not os.path.exists(PYTHON39):
        return False
    return subprocess.run([PYTHON39, "-c", "import h5py"], capture_output=True).returncode == 0


needs_h5py = pytest.mark.skipif(not _have_h5py(), reason="no h5py interpreter to cross-check with")


def h5py_run(code: str, *args: str) -> None:
    subprocess.run([PYTHON39, "-c", code, *args], check=True)


def h5py_dump(path: str, out_npz: str) -> dict:
    """Every dataset of a file as h5py reads it, with dtype / chunk info."""
    h5py_run(r"""
import sys, json, numpy as np, h5py
arrays, meta = {}, {}
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        arrays[name] = obj[()]
        meta[name] = [obj.dtype.str, list(obj.shape), obj.chunks is None]
    else:
        meta[name] = "group"
with h5py.File(sys.argv[1], "r") as f:
    f.visititems(visit)
np.savez(sys.argv[2], **{k.replace("/", "|"): v for k, v in arrays.items()})
open(sys.argv[2] + ".json", "w").write(json.dumps(meta))
""", path, out_npz)
    import json
    with np.load(out_npz) as d:
        arrays = {k.replace("|", "/"): d[k] for k in d.files}
    with open(out_npz + ".json") as f:
        meta = json.load(f)
    return {"arrays": arrays, "meta": meta}


# ---------------------------------------------------------------- library


def test_library_exports_every_declared_symbol():
    with open(HEADER) as f:
        text = f.read()
    names = sorted(set(re.findall(r"\b(ls_amd_hdf5_\w+)\s*\(", text)))
    assert len(names) == 11
    lib = H.load()
    for n in names:
        assert hasattr(lib, n), n


# ---------------------------------------------------------------- fixtures


def popcount(xs):
    return np.array([bin(int(v)).count("1") for v in xs])


def heisenberg_matrix(name, states):
    """H on the listed states from the bond list of the model file: per bond
    c (sz sz + 2 (s+ s- + s- s+)), c = 1 for Pauli matrices, 1/4 for spin-1/2
    operators (expression written with S)."""
    with open(os.path.join(MODELS, name + ".yaml"), encoding="utf-8") as f:
        m = yaml.safe_load(f)
    terms = m["hamiltonian"]["terms"]
    c = 0.25 if terms[0]["expression"].startswith("S") else 1.0
    bonds = terms[0]["sites"]
    assert all(t["sites"] == bonds for t in terms) and len(terms) == 3
    index = {int(s): k for k, s in enumerate(states)}
    M = np.zeros((len(states), len(states)))
    for k, s in enumerate(states):
        s = int(s)
        for i, j in bonds:
            bi, bj = (s >> i) & 1, (s >> j) & 1
            M[k, k] += c * (1.0 if bi == bj else -1.0)
            if bi != bj:
                M[index[s ^ (1 << i) ^ (1 << j)], k] += 2.0 * c
    return M


@pytest.mark.parametrize("name", NAMES)
def test_fixture_layout_and_read(name):
    reps = H.read_dataset(os.path.join(GOLDEN, "construction", name + ".h5"), "/representatives", np.uint64, 1)
    assert reps.dtype == np.uint64 and reps.ndim == 1
    f = os.path.join(GOLDEN, "matvec", name + ".h5")
    assert H.dataset_shape(f, "/x", 2) == (2, reps.size)
    x = H.read_dataset_as_blocks(f, "/x")
    y = H.read_dataset_as_blocks(f, "/y")
    assert x.shape == y.shape == (2, reps.size)
    # a hyperslab equals the same slice of the whole
    part = H.read_dataset_chunk(f, "/y", (1, 5), dtype=np.float64, shape=(1, 37))
    assert np.array_equal(part, y[1:2, 5:42])


def test_fixture_contents_pinned_by_closed_forms():
    from helpers import burnside_dimension, weight_states
    # unprojected kagome 12: every state of weight 6, and y = H x
    reps = H.read_dataset(os.path.join(GOLDEN, "construction", "heisenberg_kagome_12.h5"), "/representatives",
                          np.uint64, 1)
    assert np.array_equal(reps, weight_states(12, 6))
    f = os.path.join(GOLDEN, "matvec", "heisenberg_kagome_12.h5")
    M = heisenberg_matrix("heisenberg_kagome_12", reps)
    x, y = H.read_dataset(f, "/x", np.float64, 2), H.read_dataset(f, "/y", np.float64, 2)
    assert np.allclose(y, x @ M.T, rtol=1e-12, atol=1e-12)
    # chain 10, spin inversion -1: the states of weight 5 below their flip;
    # sector matrix M[i, j] = H[r_i, r_j] - H[r_i, ~r_j]
    reps = H.read_dataset(os.path.join(GOLDEN, "construction", "heisenberg_chain_10.h5"), "/representatives",
                          np.uint64, 1)
    full = weight_states(10, 5)
    assert np.array_equal(reps, full[full < (full ^ np.uint64(1023))])
    Hf = heisenberg_matrix("heisenberg_chain_10", full)
    pos = {int(s): k for k, s in enumerate(full)}
    ri = [pos[int(r)] for r in reps]
    fi = [pos[int(r) ^ 1023] for r in reps]
    Ms = Hf[np.ix_(ri, ri)] - Hf[np.ix_(ri, fi)]
    f = os.path.join(GOLDEN, "matvec", "heisenberg_chain_10.h5")
    x, y = H.read_dataset(f, "/x", np.float64, 2), H.read_dataset(f, "/y", np.float64, 2)
    assert np.allclose(y, x @ Ms.T, rtol=1e-12, atol=1e-12)
    # kagome 12 with symmetries: Burnside dimension, weight 6, ascending
    with open(os.path.join(MODELS, "heisenberg_kagome_12_symm.yaml"), encoding="utf-8") as fh:
        m = yaml.safe_load(fh)
    import lsamd
    S = lsamd.Symmetries([lsamd.Symmetry(s["permutation"], s["sector"]) for s in m["basis"]["symmetries"]])
    chars = S.characters
    dim = burnside_dimension(12, 6, [list(p) for p in S.elements()], list(chars), 0)
    reps = H.read_dataset(os.path.join(GOLDEN, "construction", "heisenberg_kagome_12_symm.h5"),
                          "/representatives", np.uint64, 1)
    assert reps.size == round(dim.real) and np.all(popcount(reps) == 6) and np.all(np.diff(reps.astype(float)) > 0)


@needs_h5py
@pytest.mark.parametrize("name", NAMES)
def test_fixture_reads_match_h5py(tmp_path, name):
    for kind, dsets in (("construction", [("/representatives", np.uint64, 1)]),
                        ("matvec", [("/x", np.float64, 2), ("/y", np.float64, 2)])):
        f = os.path.join(GOLDEN, kind, name + ".h5")
        ref = h5py_dump(f, str(tmp_path / f"{kind}.npz"))["arrays"]
        for d, dt, r in dsets:
            mine = H.read_dataset(f, d, dt, r)
            want = ref[d.lstrip("/")]
            assert mine.dtype == want.dtype and mine.shape == want.shape
            assert mine.tobytes() == want.tobytes()


# ---------------------------------------------------------------- writing


@needs_h5py
def test_written_files_read_back_by_h5py(tmp_path):
    f = str(tmp_path / "out.h5")
    rng = np.random.default_rng(5)
    reps = np.sort(rng.choice(1 << 40, 1000, replace=False)).astype(np.uint64)
    vecs = rng.standard_normal((3, 1000))
    evals = rng.standard_normal(3)
    H.make_file(f)
    for g in ("basis", "hamiltonian", "observables"):
        H.make_group(f, g)
    H.make_group(f, "basis")  # exists: no-op
    H.write_dataset_as_blocks(f, "basis/representatives", reps, slab_bytes=8 * 77)
    H.write_dataset_as_blocks(f, "hamiltonian/eigenvectors", vecs, slab_bytes=8 * 300)
    H.write_dataset(f, "hamiltonian/eigenvalues", evals)
    H.write_dataset(f, "hamiltonian/residuals", np.zeros(3))
    # replace with a different shape and type, as writeDataset deletes first
    H.write_dataset(f, "hamiltonian/residuals", np.arange(4, dtype=np.uint64))
    H.write_dataset(f, "hamiltonian/residuals", np.full(3, 1e-9))
    # chunk writes into an existing dataset
    H.create_dataset(f, "observables/m", np.float64, (2, 10))
    H.write_dataset_chunk(f, "observables/m", (0, 0), np.ones((2, 10)))
    H.write_dataset_chunk(f, "observables/m", (1, 3), np.full((1, 4), 7.0))
    d = h5py_dump(f, str(tmp_path / "dump.npz"))
    meta, arr = d["meta"], d["arrays"]
    assert meta["basis"] == meta["hamiltonian"] == meta["observables"] == "group"
    assert meta["basis/representatives"] == ["<u8", [1000], True]
    assert meta["hamiltonian/eigenvectors"] == ["<f8", [3, 1000], True]
    assert meta["hamiltonian/residuals"] == ["<f8", [3], True]
    assert arr["basis/representatives"].tobytes() == reps.tobytes()
    assert arr["hamiltonian/eigenvectors"].tobytes() == vecs.tobytes()
    assert arr["hamiltonian/eigenvalues"].tobytes() == evals.tobytes()
    want = np.ones((2, 10))
    want[1, 3:7] = 7.0
    assert np.array_equal(arr["observables/m"], want)


def test_object_management(tmp_path):
    f = str(tmp_path / "m.h5")
    assert not os.path.exists(f)
    # doesObjectExist opens through openFile, which creates a missing file
    assert H.does_object_exist(f, "basis") is False
    assert os.path.exists(f)
    H.make_group(f, "basis")
    assert H.does_object_exist(f, "basis")
    H.write_dataset(f, "basis/representatives", np.arange(5, dtype=np.uint64))
    assert H.does_object_exist(f, "basis/representatives")
    H.delete_object(f, "basis/representatives")
    assert not H.does_object_exist(f, "basis/representatives")
    with pytest.raises(H.HDF5Error, match="could not delete 'basis/representatives'"):
        H.delete_object(f, "basis/representatives")


def test_empty_datasets(tmp_path):
    f = str(tmp_path / "e.h5")
    H.write_dataset(f, "/representatives", np.zeros(0, dtype=np.uint64))
    assert H.dataset_shape(f, "/representatives", 1) == (0,)
    assert H.read_dataset(f, "/representatives", np.uint64, 1).size == 0
    H.write_dataset_as_blocks(f, "/x", np.zeros((2, 0)))
    assert H.read_dataset(f, "/x", np.float64, 2).shape == (2, 0)


# ---------------------------------------------------------------- errors


def test_reference_error_messages(tmp_path):
    f = str(tmp_path / "err.h5")
    H.write_dataset(f, "/representatives", np.arange(4, dtype=np.uint64))
    H.write_dataset(f, "/x", np.zeros((1, 4)))
    with pytest.raises(H.HDF5Error, match=r"type mismatch in file: '.*' dataset: '/representatives'"):
        H.read_dataset(f, "/representatives", np.float64, 1)
    with pytest.raises(H.HDF5Error, match=r"rank mismatch in file: '.*' dataset: '/x'  1 != 2"):
        H.read_dataset(f, "/x", np.float64, 1)
    with pytest.raises(H.HDF5Error, match=r"expected '/x' to be 1-dimensional, but it has rank 2"):
        H.dataset_shape(f, "/x", 1)
    with pytest.raises(H.HDF5Error, match="could not open dataset '/y'"):
        H.read_dataset(f, "/y", np.float64, 2)
    with pytest.raises(H.HDF5Error, match="could not open file"):
        H.read_dataset(str(tmp_path / "absent.h5"), "/x", np.float64, 2)
    with pytest.raises(H.HDF5Error, match="outside dataset '/x'"):
        H.read_dataset_chunk(f, "/x", (0, 2), dtype=np.float64, shape=(1, 3))
    with pytest.raises(H.HDF5Error, match="could not open dataset '/z'"):
        H.write_dataset_chunk(f, "/z", (0,), np.zeros(2))
    with pytest.raises(H.HDF5Error, match=r"type mismatch.*'/x'"):
        H.write_dataset_chunk(f, "/x", (0, 0), np.zeros((1, 2), dtype=np.uint64))
    # a missing intermediate group fails H5Lexists, as in doesObjectExist
    with pytest.raises(H.HDF5Error, match="could not check whether group 'basis/representatives' exists"):
        H.does_object_exist(f, "basis/representatives")
    with pytest.raises(H.HDF5Error):
        H.write_dataset(f, "nogroup/x", np.zeros(3))
    with pytest.raises(H.HDF5Error, match="unsupported element type"):
        H.write_dataset(f, "/i", np.zeros(3, dtype=np.int64))


@needs_h5py
def test_stored_types_other_than_u64_f64_are_type_mismatches(tmp_path):
    f = str(tmp_path / "types.h5")
    h5py_run(r"""
import sys, numpy as np, h5py
with h5py.File(sys.argv[1], "w") as f:
    f.create_dataset("i64", data=np.arange(4, dtype="<i8"))
    f.create_dataset("u64be", data=np.arange(4, dtype=">u8"))
    f.create_dataset("f32", data=np.zeros((1, 4), dtype="<f4"))
    f.create_dataset("u64", data=np.arange(4, dtype="<u8"))
""", f)
    for d, dt, r in (("i64", np.uint64, 1), ("u64be", np.uint64, 1), ("f32", np.float64, 2)):
        with pytest.raises(H.HDF5Error, match="type mismatch"):
            H.read_dataset(f, d, dt, r)
    assert np.array_equal(H.read_dataset(f, "u64", np.uint64, 1), np.arange(4, dtype=np.uint64))
