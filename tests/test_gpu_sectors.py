"""Many sectors in one product and the batched block Lanczos, against the
oracle's dense sector matrices and exact results.

Every momentum x spin-inversion sector of a J1-J2 chain is built on the
device with the norm tolerance 1e-9 (complex characters), its complex CSR
cached, and all of them concatenated into one block-diagonal operator.
"""
import sys

import helpers as H
import numpy as np
import pytest

import lsamd
from lsamd import models, sectors as LS
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_sector_matrices(n, h, expr, sector_list):
    return H.chain_sector_matrices(n, h, expr, sector_list, LS.DEFAULT_NORM_TOLERANCE)


def test_sector_product_matches_oracle(engine):
    n, h = 12, 6
    expr = models.j1j2_chain(n, 1.0, 0.5)
    say("sector product", n)
    batch, bases, ops = LS.chain_sector_batch(n, h, expr, keep_operators=True)
    ref = oracle_sector_matrices(n, h, expr, batch.sectors)
    assert batch.count == 2 * n
    for b, (reps, M) in zip(bases, ref):
        assert np.array_equal(b.states, reps)
        assert np.abs(M - M.conj().T).max() < 1e-12
    assert list(batch.offsets) == list(np.cumsum([0] + [r.size for r, _ in ref]))
    rng = np.random.default_rng(7)
    for k in (1, 2, 4, 8):
        X = rng.standard_normal((batch.number_states, k)) + 1j * rng.standard_normal((batch.number_states, k))
        want = np.concatenate([M @ X[batch.offsets[s]:batch.offsets[s + 1]] for s, (_, M) in enumerate(ref)])
        scale = np.abs(want).max()
        y64 = batch.matvec(X, 64)
        assert np.abs(y64 - want).max() <= RTOL64 * scale
        y32 = batch.matvec(X, 32)
        assert np.abs(y32 - want).max() <= 2e-5 * scale
    # every schedule computes the same product
    X = rng.standard_normal((batch.number_states, 8)) + 1j * rng.standard_normal((batch.number_states, 8))
    want = batch.matvec(X, 64)
    for lanes in (1, 2, 4, 8, 16):
        for blocks, die in ((8, 1), (8, 0), (64, 1), (64, 0), (1024, 1), (1000, 0)):
            batch.set_schedule(lanes, blocks, die)
            assert np.abs(batch.matvec(X, 64) - want).max() <= RTOL64 * np.abs(want).max()
    sched = batch.autotune(8)
    assert sched["lanes"] in (1, 2, 4, 8, 16)
    assert np.abs(batch.matvec(X, 64) - want).max() <= RTOL64 * np.abs(want).max()
    # the graph is replayed for repeated calls on the same buffers
    E = batch.to_embedded(X, np.float32)
    dx = lsamd.DeviceArray.from_numpy(E)
    dy = lsamd.DeviceArray(E.nbytes)
    first = None
    for _ in range(3):
        batch.matvec_device(8, dx, dy, 32)
        y = batch.from_embedded(dy.to_numpy(np.float32, E.shape))
        if first is None:
            first = y
        assert np.array_equal(y, first)


def test_sector_product_rejects_bad_arguments(engine):
    n, h = 8, 4
    expr = models.j1j2_chain(n, 1.0, 0.5)
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), 0)])
    b = lsamd.DeviceBasis(n, h, None, S)
    b.build()
    op = lsamd.DeviceOperator(b, expr)
    op.build_csr(96)
    with pytest.raises(lsamd.EngineError, match="complex CSR"):
        lsamd.DeviceSectorBatch([op])
    op.build_csr(128)
    with pytest.raises(lsamd.EngineError, match="complex values"):
        op.csr_matvec(lsamd.DeviceArray(8 * b.number_states), lsamd.DeviceArray(8 * b.number_states))
    batch = lsamd.DeviceSectorBatch([op])
    with pytest.raises(lsamd.EngineError):
        batch.set_schedule(3, 64, True)
    with pytest.raises(lsamd.EngineError):
        batch.set_schedule(4, 12, True)


def test_sector_eigensolve_matches_dense(engine):
    """Lowest 4 eigenvalues of all 32 sectors of the 16-site J1-J2 chain at
    the Majumdar-Ghosh point, FP32 product then FP64, against dense
    diagonalisation of the oracle's matrices; the two dimer ground states
    (E0 = -3N/2 with sigma.sigma couplings) sit at k = 0 and k = pi, both
    with spin inversion +1."""
    n, h = 16, 8
    expr = models.j1j2_chain(n, 1.0, 0.5)
    say("sector eigensolve", n)
    batch = LS.chain_sector_batch(n, h, expr)
    ref = oracle_sector_matrices(n, h, expr, batch.sectors)
    r = batch.eigensolve(nev=4, block_size=4, max_blocks=10, keep_blocks=2, max_cycles=40, tol=1e-8)
    say("cycles", r["cycles"], "fp64", r["fp64_cycles"], "products", r["products"], "resid", r["max_residual"])
    assert r["converged"]
    assert r["fp64_cycles"] >= 1
    assert r["max_residual"] < 1e-6
    for s, (_, M) in enumerate(ref):
        exact = np.linalg.eigvalsh(M)[:4]
        got = r["eigenvalues"][s]
        assert np.abs(got - exact).max() <= 1e-9 * max(1.0, np.abs(exact).max()), (batch.sectors[s], got, exact)
    e0 = {sec: r["eigenvalues"][s][0] for s, sec in enumerate(batch.sectors)}
    assert abs(e0[(0, 1)] + 1.5 * n) < 1e-10 * 1.5 * n
    assert abs(e0[(n // 2, 1)] + 1.5 * n) < 1e-10 * 1.5 * n
    assert min(e0.values()) > -1.5 * n - 1e-9


def test_complex_lanczos_mode_matches_dense(engine):
    n, h = 16, 8
    expr = models.j1j2_chain(n, 1.0, 0.5)
    say("complex lanczos", n)
    (reps, M), = oracle_sector_matrices(n, h, expr, [(3, -1)])
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), 3)])
    b = lsamd.DeviceBasis(n, h, -1, S)
    b.set_norm_tolerance(LS.DEFAULT_NORM_TOLERANCE)
    b.build()
    assert np.array_equal(b.states, reps)
    op = lsamd.DeviceOperator(b, expr)
    res = op.lanczos("otf_complex", max_steps=300, k=3, tol=1e-12)
    exact = np.linalg.eigvalsh(M)
    assert np.abs(np.array(res["eigenvalues"][:3]) - exact[:3]).max() <= 1e-10 * np.abs(exact[0])
    assert res["residual_norm"] < 1e-6


def test_fermionic_sector_batch_matches_oracle(engine):
    """Momentum sectors of a 6-site Hubbard ring (3 up + 3 down, fermionic
    permutation signs, complex characters) as one batch: product and lowest
    eigenvalues against the oracle's dense sector matrices."""
    n, up, down = 6, 3, 3
    expr = models.hubbard(n, models.chain_bonds(n), t=1.0, U=4.0)
    diag, off = expr.split()
    perm = models.chain_translation(n)
    gen = np.concatenate([perm, perm + n]).astype(np.int32)
    ops, mats = [], []
    for k in range(n):
        S = lsamd.Symmetries([lsamd.Symmetry(gen, k)])
        b = lsamd.DeviceBasis.fermions(n, up, down, S)
        b.set_norm_tolerance(LS.DEFAULT_NORM_TOLERANCE)
        b.build()
        op = lsamd.DeviceOperator(b, expr)
        op.build_csr(128)
        ops.append(op)
        og = O.Group(2 * n, S.elements(), S.characters.real, S.characters.imag, 0)
        O.lib().ora_group_set_fermionic.argtypes = [O.C.c_void_p, O.C.c_int]
        O.lib().ora_group_set_norm_eps.argtypes = [O.C.c_void_p, O.C.c_double]
        assert O.lib().ora_group_set_fermionic(og.h, 2) == 0
        O.lib().ora_group_set_norm_eps(og.h, LS.DEFAULT_NORM_TOLERANCE)
        reps = b.states.copy()
        M = O.matrix_dense(og, 2 * n, 0, reps, O.Terms(diag), O.Terms(off))
        assert np.abs(M - M.conj().T).max() < 1e-12
        mats.append(M)
    batch = lsamd.DeviceSectorBatch(ops)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((batch.number_states, 4)) + 1j * rng.standard_normal((batch.number_states, 4))
    want = np.concatenate([M @ X[batch.offsets[s]:batch.offsets[s + 1]] for s, M in enumerate(mats)])
    assert np.abs(batch.matvec(X, 64) - want).max() <= RTOL64 * np.abs(want).max()
    assert np.abs(batch.matvec(X, 32) - want).max() <= 2e-5 * np.abs(want).max()
    r = batch.eigensolve(nev=2, block_size=4, max_blocks=8, keep_blocks=2, max_cycles=60, tol=1e-9)
    assert r["converged"]
    for s, M in enumerate(mats):
        exact = np.linalg.eigvalsh(M)[:2]
        assert np.abs(r["eigenvalues"][s] - exact).max() <= 1e-9 * max(1.0, np.abs(exact).max())


def test_chain28_sector_batch_full_size(engine):
    """All 56 sectors of the 28-site J1-J2 chain (40 116 600 states) in one
    batch: the exact Majumdar-Ghosh ground states at k = 0 and k = pi, and two
    complex sectors against single-sector FP64 Lanczos."""
    n, h = 28, 14
    expr = models.j1j2_chain(n, 1.0, 0.5)
    say("chain28 sectors")
    batch, bases, ops = LS.chain_sector_batch(n, h, expr, keep_operators=True)
    assert batch.count == 56 and batch.number_states == 40116600
    sizes = np.diff(batch.offsets)
    assert sizes.min() == 716014 and sizes.max() == 716728
    # every sector's dimension is its Burnside count (cycle formula, no enumeration)
    from test_dimensions import dim
    want = [dim(n, h, [models.chain_translation(n)], [k], si) for k, si in batch.sectors]
    assert [int(d) for d in sizes] == want
    names = [f"{k}{'+' if si == 1 else '-'}" for k, si in batch.sectors]
    refs = {nm: ops[names.index(nm)] for nm in ("5+", "11-")}
    del bases, ops
    batch.autotune(8)
    r = batch.eigensolve(nev=4, block_size=4, max_blocks=16, keep_blocks=4, max_cycles=40, tol=1e-7)
    say("cycles", r["cycles"], "fp64", r["fp64_cycles"], "seconds", r["total_seconds"])
    assert r["converged"] and r["max_residual"] < 1e-6
    ev = r["eigenvalues"]
    for nm in ("0+", "14+"):
        assert abs(ev[names.index(nm)][0] + 1.5 * n) <= 1e-10 * 1.5 * n
    assert ev[:, 0].min() >= -1.5 * n - 1e-9
    for nm, op in refs.items():
        ref = op.lanczos("otf_complex", max_steps=400, k=4, tol=1e-12)
        assert np.abs(ev[names.index(nm)] - np.array(ref["eigenvalues"][:4])).max() <= 1e-10 * 1.5 * n, nm
