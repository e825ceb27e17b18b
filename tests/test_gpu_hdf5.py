"""HDF5 interchange on the GPU: the reference's two HDF5 tests and its
Diagonalize application, run through the engine.

- TestStatesEnumeration (chapel/test/TestStatesEnumeration.chpl:9-38): the
  basis enumerated on the device equals "/representatives" of the file.
- TestMatrixVectorProduct (chapel/test/TestMatrixVectorProduct.chpl:24-60):
  y = H x for the rows of "/x" equals "/y" within atol 1e-13, rtol 1e-11
  (its approxEqual at :15-23), on-the-fly and through the CSR; once more
  with x read into and y written from device memory.
- Diagonalize (chapel/src/Diagonalize.chpl:226-330): eigenvalues against a
  dense diagonalisation of the oracle's matrix, eigenvectors checked by their
  residuals and orthonormality, the file layout read back, and a second run
  on the same file reusing the stored representatives.
"""
import os

import numpy as np
import pytest

import lsamd
from lsamd import hdf5 as H
from lsamd.diagonalize import diagonalize
from oracle import oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not H.available(), reason="liblsamd_hdf5.so not built (no HDF5 headers)")]

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "hdf5")
MODELS = os.path.join(HERE, "golden", "models")
NAMES = ["heisenberg_chain_10", "heisenberg_kagome_12", "heisenberg_kagome_12_symm"]
ATOL, RTOL = 1e-13, 1e-11


def approx_equal(a, b):
    return np.abs(a - b) <= np.maximum(ATOL, RTOL * np.maximum(np.abs(a), np.abs(b)))


class Row:
    """Row k of a [rows][n] f64 device array, as the products take it."""

    def __init__(self, base, k, n):
        self.base = base
        self.ptr = base.ptr + 8 * k * n
        self.nbytes = 8 * n


def model(name):
    return lsamd.Model.from_file(os.path.join(MODELS, name + ".yaml"))


@pytest.mark.parametrize("name", NAMES)
def test_states_enumeration_against_file(engine, name):
    print("hdf5 enumeration", name, flush=True)
    b = model(name).device_basis()
    b.build()
    want = H.read_dataset_as_blocks(os.path.join(GOLDEN, "construction", name + ".h5"), "/representatives",
                                    rank=1, dtype=np.uint64)
    assert np.array_equal(b.states, want)


@pytest.mark.parametrize("name", NAMES)
def test_matrix_vector_product_against_file(engine, name):
    print("hdf5 matvec", name, flush=True)
    m = model(name)
    b = m.device_basis()
    b.build()
    op = m.device_operator(b)
    f = os.path.join(GOLDEN, "matvec", name + ".h5")
    x = H.read_dataset_as_blocks(f, "/x")
    y = H.read_dataset_as_blocks(f, "/y")
    op.build_csr(64)
    for k in range(x.shape[0]):
        for mode in ("otf", "csr"):
            z = op.matvec(x[k], mode)
            ok = approx_equal(z, y[k])
            assert ok.all(), (mode, k, np.flatnonzero(~ok)[:10], np.abs(z - y[k]).max())
    # device-resident: x read slab by slab into HBM, y streamed back out
    dx, shape = H.read_device_dataset(f, "/x", np.float64, 2, slab_bytes=8 * 100)
    n = shape[1]
    dy = lsamd.DeviceArray(8 * x.size)
    for k in range(shape[0]):
        op.matvec_otf(Row(dx, k, n), Row(dy, k, n))
    out = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"lsamd_y_{name}_{os.getpid()}.h5")
    try:
        H.write_device_dataset(out, "/y", dy, np.float64, x.shape, slab_bytes=8 * 100)
        z = H.read_dataset(out, "/y", np.float64, 2)
    finally:
        if os.path.exists(out):
            os.remove(out)
    assert approx_equal(z, y).all()


def oracle_matrix(m, reps):
    n, si = m.number_sites, m.spin_inversion or 0
    if m.symmetries:
        S = lsamd.Symmetries([lsamd.Symmetry(p, s) for p, s in m.symmetries])
        og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    else:
        og = O.Group(n, np.arange(n, dtype=np.int64)[None, :], np.ones(1), np.zeros(1), si)
    M = O.matrix_dense(og, n, si, reps, O.Terms(m.terms(-1, True)), O.Terms(m.terms(-1, False)))
    return M.real


@pytest.mark.parametrize("name,nev", [("heisenberg_kagome_12_symm", 2), ("heisenberg_chain_10", 1),
                                      ("heisenberg_kagome_12", 3)])
def test_diagonalize_writes_the_reference_layout(engine, tmp_path, name, nev):
    print("diagonalize", name, nev, flush=True)
    out = str(tmp_path / "ed.h5")
    r = diagonalize(os.path.join(MODELS, name + ".yaml"), out, num_evals=nev, eps=1e-10, max_block_size=nev)
    assert r["converged"] and not r["representatives_reused"]
    reps = H.read_dataset(out, "basis/representatives", np.uint64, 1)
    assert np.array_equal(reps, H.read_dataset(os.path.join(GOLDEN, "construction", name + ".h5"),
                                               "/representatives", np.uint64, 1))
    for g in ("basis", "hamiltonian", "observables"):
        assert H.does_object_exist(out, g)
    evals = H.read_dataset(out, "hamiltonian/eigenvalues", np.float64, 1)
    res = H.read_dataset(out, "hamiltonian/residuals", np.float64, 1)
    V = H.read_dataset(out, "hamiltonian/eigenvectors", np.float64, 2)
    assert evals.shape == res.shape == (nev,) and V.shape == (nev, reps.size)
    M = oracle_matrix(model(name), reps)
    w = np.linalg.eigvalsh(M)[:nev]
    assert np.allclose(evals, w, rtol=1e-9, atol=1e-9), (evals, w)
    assert np.allclose(V @ V.T, np.eye(nev), atol=1e-8)
    R = V @ M.T - evals[:, None] * V
    true_res = np.linalg.norm(R, axis=1)
    assert np.all(true_res <= 1e-8 * np.maximum(1, np.abs(evals)))
    assert np.all(res <= 1e-8 * np.maximum(1, np.abs(evals)))

    # a second run resumes on the stored representatives
    r2 = diagonalize(os.path.join(MODELS, name + ".yaml"), out, num_evals=nev, eps=1e-10, max_block_size=nev)
    assert r2["representatives_reused"]
    assert np.allclose(r2["eigenvalues"], evals, rtol=1e-9, atol=1e-9)


def test_diagonalize_rejects_a_stored_list_of_another_basis(engine, tmp_path):
    out = str(tmp_path / "bad.h5")
    H.make_file(out)
    H.make_group(out, "basis")
    reps = H.read_dataset(os.path.join(GOLDEN, "construction", "heisenberg_kagome_12.h5"), "/representatives",
                          np.uint64, 1)
    # every state of weight 6, but the symmetric sector holds only orbit minima
    H.write_dataset(out, "basis/representatives", reps)
    with pytest.raises(lsamd.EngineError, match="is not a representative of this basis"):
        diagonalize(os.path.join(MODELS, "heisenberg_kagome_12_symm.yaml"), out)
    H.write_dataset(out, "basis/representatives", reps[::-1].copy())
    with pytest.raises(lsamd.EngineError, match="strictly increasing"):
        diagonalize(os.path.join(MODELS, "heisenberg_kagome_12_symm.yaml"), out)
    with pytest.raises(ValueError, match="invalid numEvals: 0"):
        diagonalize(os.path.join(MODELS, "heisenberg_kagome_12_symm.yaml"), out, num_evals=0)
