"""GPU parity of the push-form product on operators that are not Hermitian,
and of the three coefficient paths of the reference: full projection,
spin inversion only, and no projection (fixed Hamming weight, and the
identity index of the whole space).

The reference's product is y[index(rep beta)] += Re(c) x_alpha over the
pairs every alpha emits (chapel/src/DistributedMatrixVector.chpl:91-143,
chapel/src/BatchedOperator.chpl:139-253) for any Operator(basis, expr),
python/lattice_symmetries/__init__.py:626-631, 715-729.  Every product form
of the engine (reference entry point with the CSR cache and on the fly,
device on-the-fly, CSR, FP64 and FP32 blocks, complex vectors) must match
the oracle's push form to 1e-12 (FP64) or 2e-5 (FP32)."""
import os
import sys

import numpy as np
import pytest

import lsamd
from lsamd import models
from lsamd.expr import sigma_minus, sigma_plus, sigma_z
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12
RTOL32 = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def non_hermitian(n):
    """sigma+_0 sigma-_3 + 0.3 sigma^z_1 + (0.4 + 0.2i) sigma+_(n-2) sigma-_0.
    The last term also applies to orbit minima of a translation group (bit 0
    up), so the projected matrices are not Hermitian either."""
    return (sigma_plus(0) * sigma_minus(3) + sigma_z(1) * 0.3
            + sigma_plus(n - 2) * sigma_minus(0) * (0.4 + 0.2j))


# name, n, Hamming weight, spin inversion, generators, sectors
BASES = [
    ("projected_k0_p", 12, 6, 1, [models.chain_translation(12), models.chain_parity(12)], [0, 0]),
    ("projected_k3_complex", 12, 6, 0, [models.chain_translation(12)], [3]),
    ("inversion_plus", 12, 6, 1, None, None),
    ("inversion_minus", 12, 6, -1, None, None),
    ("unprojected_weight", 12, 6, 0, None, None),
    ("unprojected_identity", 10, None, 0, None, None),
]


def make_bases(n, h, si, gens, sectors):
    S = None
    og = None
    if gens is not None:
        S = lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])
        og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    sb = lsamd.SpinBasis(n, h, si if si else None, S)
    sb.build()
    db = lsamd.DeviceBasis(n, h, si if si else None, S)
    db.build()
    return og, sb, db


def push_reference(og, n, si, reps, expr, x):
    diag, off = expr.split()
    return O.matvec_push(og, n, si, reps, O.Terms(diag), O.Terms(off), x)


@pytest.mark.parametrize("name,n,h,si,gens,sectors", BASES, ids=[b[0] for b in BASES])
@pytest.mark.parametrize("which", ["non_hermitian", "heisenberg"])
def test_every_product_form_matches_the_push_form(engine, name, n, h, si, gens, sectors, which):
    say("push form", name, which)
    expr = non_hermitian(n) if which == "non_hermitian" else models.heisenberg_chain(n, 1.0, 0.25)
    og, sb, db = make_bases(n, h, si, gens, sectors)
    reps = sb.states.copy()
    assert np.array_equal(reps, db.states)
    if gens is None:
        want_states = O.enumerate_unprojected(n, h, si)
        assert np.array_equal(reps, want_states)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(reps.size)
    want = push_reference(og, n, si, reps, expr, x)
    scale = np.abs(want).max()
    assert scale > 0
    diag, off = expr.split()
    M = O.matrix_dense(og, n, si, reps, O.Terms(diag), O.Terms(off))
    if which == "non_hermitian":
        # the case is what it claims: Re(M) is not symmetric
        assert np.abs(M.real - M.real.T).max() > 0.1
    assert np.abs(M.real @ x - want).max() <= RTOL64 * scale
    # reference entry point: CSR cache (default), then on the fly
    y = lsamd.Operator(sb, expr).matvec(x)
    assert np.abs(y - want).max() <= RTOL64 * scale
    os.environ["LSAMD_MATVEC"] = "otf"
    try:
        y = lsamd.Operator(sb, expr).matvec(x)
    finally:
        del os.environ["LSAMD_MATVEC"]
    assert np.abs(y - want).max() <= RTOL64 * scale
    # device API
    dop = lsamd.DeviceOperator(db, expr)
    assert np.abs(dop.matvec(x, "otf") - want).max() <= RTOL64 * scale
    dop.build_csr(96)
    assert np.abs(dop.matvec(x, "csr") - want).max() <= RTOL64 * scale
    X = rng.standard_normal((reps.size, 8))
    ref8 = M.real @ X
    assert np.abs(dop.matvec_block(X, 64) - ref8).max() <= RTOL64 * np.abs(ref8).max()
    assert np.abs(dop.matvec_block(X, 32) - ref8).max() <= RTOL32 * np.abs(ref8).max()
    # complex vectors: y = M x with the complex matrix
    xc = x + 1j * rng.standard_normal(reps.size)
    wc = M @ xc
    assert np.abs(dop.matvec_c(xc) - wc).max() <= RTOL64 * np.abs(wc).max()
    # the unpacked CSR stores the same matrix
    plain = lsamd.DeviceOperator(db, expr)
    plain.set_csr_packing(False)
    plain.build_csr(96)
    assert np.abs(plain.matvec(x, "csr") - want).max() <= RTOL64 * scale
    # Hermiticity as the eigensolvers see it
    herm = which == "heisenberg"
    for mode in (0, 1, 2, 3):
        assert dop.is_hermitian(mode) == herm, mode


def test_csr_rows_are_the_push_rows_in_ascending_source_order(engine):
    """Row j of the cached CSR lists the sources i of M[j, i] in ascending i,
    with the values of the oracle's dense push matrix."""
    n, h = 12, 6
    og, sb, db = make_bases(n, h, 0, [models.chain_translation(n)], [3])
    expr = non_hermitian(n)
    dop = lsamd.DeviceOperator(db, expr)
    dop.set_csr_packing(False)
    dop.build_csr(128)
    rp, cols, vals, diag = dop.csr_arrays()
    reps = db.states
    d, o = expr.split()
    M = O.matrix_dense(og, n, 0, reps, O.Terms(d), O.Terms(o))
    R = np.zeros_like(M)
    for j in range(reps.size):
        cs = cols[rp[j]:rp[j + 1]]
        assert np.all(np.diff(cs.astype(np.int64)) >= 0)
        np.add.at(R, (j, cs), vals[rp[j]:rp[j + 1]])
    R[np.arange(reps.size), np.arange(reps.size)] += diag
    assert np.abs(R - M).max() <= 1e-13 * np.abs(M).max()


def test_eigensolvers_refuse_an_operator_that_is_not_hermitian(engine):
    n, h = 12, 6
    og, sb, db = make_bases(n, h, 0, None, None)
    dop = lsamd.DeviceOperator(db, non_hermitian(n))
    dop.build_csr(96)
    with pytest.raises(lsamd.EngineError, match="not symmetric"):
        dop.lanczos("fp64", max_steps=20)
    with pytest.raises(lsamd.EngineError, match="not symmetric"):
        dop.block_eigensolve("fp32", block_size=4, nev=1, max_blocks=6)
    with pytest.raises(lsamd.EngineError, match="not Hermitian"):
        dop.lanczos("otf_complex", max_steps=20)


def test_non_hermitian_sector_batch(engine):
    """A sector batch of non-Hermitian sector operators: the batched complex
    product is M x per sector, and the batched eigensolver refuses it."""
    n, h = 10, 5
    expr = non_hermitian(n)
    d, o = expr.split()
    ops, mats, keep = [], [], []
    for k in (0, 1, 5):
        S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), k)])
        og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 0)
        db = lsamd.DeviceBasis(n, h, None, S)
        db.build()
        dop = lsamd.DeviceOperator(db, expr)
        dop.build_csr(128)
        ops.append(dop)
        keep.append((db, S))
        mats.append(O.matrix_dense(og, n, 0, db.states, O.Terms(d), O.Terms(o)))
    batch = lsamd.DeviceSectorBatch(ops)
    rng = np.random.default_rng(3)
    dims = [m.shape[0] for m in mats]
    X = rng.standard_normal((sum(dims), 2)) + 1j * rng.standard_normal((sum(dims), 2))
    Y = batch.matvec(X, 64)
    off = 0
    for m, dim in zip(mats, dims):
        want = m @ X[off:off + dim]
        assert np.abs(Y[off:off + dim] - want).max() <= RTOL64 * max(np.abs(want).max(), 1e-300)
        off += dim
    with pytest.raises(lsamd.EngineError, match="not Hermitian"):
        batch.eigensolve(nev=1, block_size=1, max_blocks=4, keep_blocks=2)
