"""Edge cases of the device path: sectors with no states, sectors smaller
than the Krylov space (diagonalised densely), batches mixing them with large
sectors, operators with only diagonal terms, one-state and empty bases.

Products are checked against the oracle's sector matrices; spectra against
numpy's projection of the full Hamiltonian onto symmetrised vectors."""
import sys

import helpers as H
import numpy as np
import pytest

import lsamd
from lsamd import models, sectors as LS
from lsamd.expr import sigma_z, sum_exprs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_batch_product(batch, mats, rng, ks=(1, 4)):
    for k in ks:
        X = rng.standard_normal((batch.number_states, k)) + 1j * rng.standard_normal((batch.number_states, k))
        want = np.concatenate([M @ X[batch.offsets[s]:batch.offsets[s + 1]] for s, M in enumerate(mats)])
        scale = max(np.abs(want).max(), 1.0)
        assert np.abs(batch.matvec(X, 64) - want).max() <= RTOL64 * scale
        assert np.abs(batch.matvec(X, 32) - want).max() <= 2e-5 * scale


def check_spectra(r, nev, spectra):
    for s, w in enumerate(spectra):
        got = r["eigenvalues"][s]
        m = min(nev, w.size)
        assert np.abs(got[:m] - w[:m]).max(initial=0.0) <= 1e-9 * max(1.0, np.abs(w).max(initial=0.0)), (s, got, w)
        assert np.isnan(got[m:]).all(), (s, got)
        assert (r["residuals"][s][:m] < 1e-8).all()


def test_batch_of_tiny_and_empty_sectors(engine):
    """All 8 sectors of the 4-site chain (three of them empty): product,
    and every sector diagonalised densely (none fits a Krylov space)."""
    n, h = 4, 2
    expr = models.j1j2_chain(n, 1.0, 0.5)
    sl = LS.chain_sectors(n)
    say("tiny sectors", n)
    bases, ops = LS.chain_sector_operators(n, h, expr, sl)
    ref = H.chain_sector_matrices(n, h, expr, sl, LS.DEFAULT_NORM_TOLERANCE)
    sizes = [r.size for r, _ in ref]
    assert sizes.count(0) == 3
    for b, (reps, _) in zip(bases, ref):
        assert np.array_equal(b.states, reps)
    batch = lsamd.DeviceSectorBatch(ops)
    assert list(batch.offsets) == list(np.cumsum([0] + sizes))
    rng = np.random.default_rng(11)
    check_batch_product(batch, [M for _, M in ref], rng)
    r = batch.eigensolve(nev=2, block_size=4, max_blocks=8, keep_blocks=2, max_cycles=10, tol=1e-9)
    assert r["converged"] and r["products"] == 0 and r["cycles"] == 0
    check_spectra(r, 2, [H.chain_sector_spectrum(n, h, expr, k, si) for k, si in sl])
    # the options are checked even when no sector needs the Krylov process
    with pytest.raises(lsamd.EngineError, match="block size"):
        batch.eigensolve(nev=1, block_size=3)


def test_batch_mixing_large_small_and_empty_sectors(engine):
    """Three 16-site sectors (Krylov), all 10-site sectors (dense) and the
    4-site sectors (some empty), interleaved in one batch."""
    expr16 = models.j1j2_chain(16, 1.0, 0.5)
    expr10 = models.j1j2_chain(10, 1.0, 0.5)
    expr4 = models.j1j2_chain(4, 1.0, 0.5)
    parts = [(16, 8, expr16, [(0, 1)]), (4, 2, expr4, LS.chain_sectors(4)[:4]), (10, 5, expr10, LS.chain_sectors(10)),
             (16, 8, expr16, [(8, 1), (3, -1)]), (4, 2, expr4, LS.chain_sectors(4)[4:])]
    say("mixed batch")
    ops, mats, spectra, keep, names = [], [], [], [], []
    for n, h, expr, sl in parts:
        names += [(n, k, si) for k, si in sl]
        bases, o = LS.chain_sector_operators(n, h, expr, sl)
        keep.append(bases)
        ops += o
        ref = H.chain_sector_matrices(n, h, expr, sl, LS.DEFAULT_NORM_TOLERANCE)
        mats += [M for _, M in ref]
        spectra += [np.linalg.eigvalsh(M) if M.size else np.zeros(0) for _, M in ref]
    sizes = [M.shape[0] for M in mats]
    assert min(sizes) == 0 and max(sizes) > 300
    batch = lsamd.DeviceSectorBatch(ops)
    rng = np.random.default_rng(5)
    check_batch_product(batch, mats, rng, ks=(2, 8))
    r = batch.eigensolve(nev=4, block_size=4, max_blocks=10, keep_blocks=2, max_cycles=40, tol=1e-8)
    say("cycles", r["cycles"], "fp64", r["fp64_cycles"], "resid", r["max_residual"])
    assert r["converged"] and r["products"] > 0
    check_spectra(r, 4, spectra)
    # the Majumdar-Ghosh ground states of the 16-site chain (k = 0 and pi, +)
    for name in ((16, 0, 1), (16, 8, 1)):
        assert abs(r["eigenvalues"][names.index(name)][0] + 24.0) < 1e-9


def zz_chain(n):
    return sum_exprs([sigma_z(i) * sigma_z((i + 1) % n) for i in range(n)])


def test_operator_without_off_diagonal_terms(engine):
    """sigma^z sigma^z ring: the CSR has no entries and every product is the
    diagonal."""
    n, h = 12, 6
    expr = zz_chain(n)
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), 0)])
    b = lsamd.DeviceBasis(n, h, None, S)
    b.build()
    g, reps, diag, off = H.sector_setup(n, h, [models.chain_translation(n)], [0], 0, expr)
    assert np.array_equal(b.states, reps)
    M = O.matrix_dense(g, n, 0, reps, diag, off).real
    assert np.count_nonzero(M - np.diag(np.diag(M))) == 0
    op = lsamd.DeviceOperator(b, expr)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(reps.size)
    want = M @ x
    assert np.abs(op.matvec(x, "otf") - want).max() <= RTOL64 * np.abs(want).max()
    assert op.build_csr(96) == 0
    assert np.abs(op.matvec(x, "csr") - want).max() <= RTOL64 * np.abs(want).max()
    X = rng.standard_normal((reps.size, 4))
    W = M @ X
    assert np.abs(op.matvec_block(X, 64) - W).max() <= RTOL64 * np.abs(W).max()
    assert np.abs(op.matvec_block(X, 32) - W).max() <= 2e-6 * np.abs(W).max()
    # the complex CSR and a sector batch of it
    assert op.build_csr(128) == 0
    batch = lsamd.DeviceSectorBatch([op])
    check_batch_product(batch, [M.astype(complex)], rng)


def test_one_state_and_empty_bases(engine):
    """Hamming weight 0 (one state), weight 1 with translations (one state per
    momentum) and an empty sector: products, CSR and errors."""
    n = 8
    expr = models.j1j2_chain(n, 1.0, 0.5)
    for h, k in ((0, 0), (1, 3)):
        S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), k)])
        b = lsamd.DeviceBasis(n, h, None, S)
        b.build()
        g, reps, diag, off = H.sector_setup(n, h, [models.chain_translation(n)], [k], 0, expr)
        assert reps.size == 1 and np.array_equal(b.states, reps)
        M = O.matrix_dense(g, n, 0, reps, diag, off)
        op = lsamd.DeviceOperator(b, expr)
        x = np.array([0.75 - 0.5j])
        assert np.abs(op.matvec_c(x) - M @ x).max() <= RTOL64 * max(1.0, np.abs(M).max())
        op.build_csr(128)
        batch = lsamd.DeviceSectorBatch([op])
        y = batch.matvec(x.reshape(1, 1), 64)
        assert np.abs(y[0, 0] - (M @ x)[0]).max() <= RTOL64 * max(1.0, np.abs(M).max())
        assert abs(H.chain_sector_spectrum(n, h, expr, k, None)[0] - M[0, 0].real) < 1e-12
    # k = 0 with spin inversion -1 on 4 sites has no states
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(4), 0)])
    b = lsamd.DeviceBasis(4, 2, -1, S)
    b.set_norm_tolerance(LS.DEFAULT_NORM_TOLERANCE)
    b.build()
    assert b.number_states == 0 and b.states.size == 0
    op = lsamd.DeviceOperator(b, models.j1j2_chain(4, 1.0, 0.5))
    assert op.matvec(np.zeros(0), "otf").size == 0
    assert op.build_csr(96) == 0
    assert op.matvec(np.zeros(0), "csr").size == 0
    with pytest.raises(lsamd.EngineError, match="empty basis"):
        op.lanczos("fp64", max_steps=10)


def test_block_solvers_survive_an_exhausted_krylov_space(engine):
    """sigma^z sigma^z ring: a diagonal operator with few distinct values, so
    the block Krylov space of any start block is exhausted after a few steps
    (A V_j falls inside the stored blocks exactly).  The solvers refill such
    a block with fresh random directions and still return the lowest
    eigenvalues: -16 twice (the two Neel states) on the 16-site ring."""
    n, h = 16, 8
    expr = zz_chain(n)
    b = lsamd.DeviceBasis(n, h, None, None)
    b.build()
    op = lsamd.DeviceOperator(b, expr)
    op.build_csr(96)
    for mode in ("fp64", "fp32"):
        say("exhausted krylov", mode)
        r = op.block_eigensolve(mode, block_size=8, nev=2, max_blocks=12, keep_blocks=2, max_cycles=10, tol=1e-8)
        assert r["converged"], r
        assert np.abs(r["eigenvalues"] - np.array([-16.0, -16.0])).max() < 1e-10
    sl = [(0, None), (3, None), (8, None)]
    bases, ops = LS.chain_sector_operators(n, h, expr, sl)
    ref = H.chain_sector_matrices(n, h, expr, sl, LS.DEFAULT_NORM_TOLERANCE)
    batch = lsamd.DeviceSectorBatch(ops)
    r = batch.eigensolve(nev=2, block_size=4, max_blocks=12, keep_blocks=3, max_cycles=10, tol=1e-8)
    assert r["converged"]
    for s, (_, M) in enumerate(ref):
        assert np.abs(r["eigenvalues"][s] - np.linalg.eigvalsh(M)[:2]).max() < 1e-10, (sl[s], r["eigenvalues"][s])


def test_batched_solver_keeps_its_basis_orthonormal_when_a_direction_vanishes(engine):
    """The same exhausted sectors from twelve start blocks.  A Krylov block can
    lose one direction (a combination of its columns) while every column keeps
    part of its norm; the step must then orthogonalise twice, and a direction
    left at rounding-noise size must be replaced, not normalised, or the basis
    stops being orthonormal and Ritz values leave the spectrum."""
    n, h = 16, 8
    expr = zz_chain(n)
    sl = [(0, None), (3, None), (8, None)]
    bases, ops = LS.chain_sector_operators(n, h, expr, sl)
    ref = [np.linalg.eigvalsh(M)[:2] for _, M in H.chain_sector_matrices(n, h, expr, sl, LS.DEFAULT_NORM_TOLERANCE)]
    batch = lsamd.DeviceSectorBatch(ops)
    for seed in range(1, 13):
        r = batch.eigensolve(nev=2, block_size=4, max_blocks=12, keep_blocks=3, max_cycles=10, tol=1e-8, seed=seed)
        say("exhausted sectors, seed", seed, r["cycles"], r["max_residual"])
        assert r["converged"] and r["max_residual"] <= 1e-7, (seed, r)
        for s in range(len(sl)):
            assert np.abs(r["eigenvalues"][s] - ref[s]).max() < 1e-10, (seed, sl[s], r["eigenvalues"][s])
    # a tolerance below the FP64 refill limit of 1e-8 ||W||: the limit follows it
    for seed in (1, 2, 3):
        r = batch.eigensolve(nev=2, block_size=4, max_blocks=12, keep_blocks=3, max_cycles=20, tol=1e-10, seed=seed)
        say("exhausted sectors, tol 1e-10, seed", seed, r["cycles"], r["max_residual"])
        assert r["converged"] and r["max_residual"] <= 1e-9, (seed, r)
