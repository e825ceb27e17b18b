"""The operator cache behind ls_chpl_matrix_vector_product: entries follow
the data they were built from, are dropped with it, and stay within their
device-memory cap, so a basis and operator rebuilt at recycled addresses
never meet a stale matrix (the reference recomputes every product,
chapel/src/DistributedMatrixVector.chpl:1090-1105)."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import lsamd
from lsamd import _abi, models
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N, H = 12, 6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sector(k):
    # momenta 0, 3 and 6 of the 12-site chain: characters 1, -i, -1 keep every
    # stabiliser sum exact (other momenta meet the reference's NaN norms)
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(N), k)])
    sb = lsamd.SpinBasis(N, H, None, S)
    sb.build()
    og = O.Group(N, S.elements(), S.characters.real, S.characters.imag, 0)
    return S, sb, og


def want(og, reps, expr, x):
    d, o = expr.split()
    return O.matvec_push(og, N, 0, reps, O.Terms(d), O.Terms(o), x)


def entries():
    return int(lsamd.load().ls_amd_operator_cache_entries())


def test_destroying_a_basis_drops_its_entries(engine):
    expr = models.j1j2_chain(N, 1.0, 0.5)
    S, sb, og = sector(0)
    op = lsamd.Operator(sb, expr)
    x = np.random.default_rng(1).standard_normal(sb.number_states)
    before = entries()
    y = op.matvec(x)
    assert np.abs(y - want(og, sb.states, expr, x)).max() <= 1e-12 * np.abs(y).max()
    assert entries() == before + 1
    # the basis's group and index data go away while the operator lives on
    sb.__del__()
    assert entries() == before
    del op


def test_recycled_basis_struct_gets_the_new_matrix(engine):
    """The same ls_hs_basis and ls_hs_operator structs, now describing another
    sector of equal dimension: the product follows the new data."""
    expr = models.j1j2_chain(N, 1.0, 0.5)
    Sa, sa, oga = sector(0)
    Sb, sb, ogb = sector(6)
    assert sa.number_states == sb.number_states
    op = lsamd.Operator(sa, expr)
    x = np.random.default_rng(2).standard_normal(sa.number_states)
    ya = op.matvec(x)
    assert np.abs(ya - want(oga, sa.states, expr, x)).max() <= 1e-12 * np.abs(ya).max()
    saved = C.create_string_buffer(C.sizeof(_abi.ls_hs_basis))
    C.memmove(saved, C.addressof(sa._basis), C.sizeof(_abi.ls_hs_basis))
    try:
        C.memmove(C.addressof(sa._basis), C.addressof(sb._basis), C.sizeof(_abi.ls_hs_basis))
        yb = op.matvec(x)
        wb = want(ogb, sb.states, expr, x)
        assert np.abs(yb - wb).max() <= 1e-12 * np.abs(wb).max()
        assert np.abs(ya - wb).max() > 1e-3
    finally:
        C.memmove(C.addressof(sa._basis), saved, C.sizeof(_abi.ls_hs_basis))
    yb2 = op.matvec(x)
    assert np.abs(yb2 - ya).max() <= 1e-12 * np.abs(ya).max()
    del op


def test_build_multiply_destroy_loop_stays_bounded(engine):
    expr = models.heisenberg_chain(N, 1.0, 0.25)
    base = entries()
    rng = np.random.default_rng(3)
    for it in range(6):
        k = (0, 6)[it % 2]
        say("loop", it, "k", k)
        S, sb, og = sector(k)
        op = lsamd.Operator(sb, expr)
        x = rng.standard_normal(sb.number_states)
        y = op.matvec(x)
        w = want(og, sb.states, expr, x)
        assert np.abs(y - w).max() <= 1e-12 * np.abs(w).max()
        del op, sb, S
        assert entries() <= base + 1


def test_device_memory_cap_evicts_least_recently_used(engine):
    expr = models.heisenberg_chain(N, 1.0, 0.25)
    keep = []
    os.environ["LSAMD_CACHE_BYTES"] = "1"
    try:
        for k in (0, 3, 6):
            S, sb, og = sector(k)
            op = lsamd.Operator(sb, expr)
            x = np.random.default_rng(k).standard_normal(sb.number_states)
            y = op.matvec(x)
            w = want(og, sb.states, expr, x)
            assert np.abs(y - w).max() <= 1e-12 * np.abs(w).max()
            keep.append((S, sb, op))
            assert entries() == 1
    finally:
        del os.environ["LSAMD_CACHE_BYTES"]
    # evicted entries are rebuilt on their next use
    S, sb, op = keep[0]
    x = np.random.default_rng(9).standard_normal(sb.number_states)
    og = O.Group(N, S.elements(), S.characters.real, S.characters.imag, 0)
    w = want(og, sb.states, expr, x)
    assert np.abs(op.matvec(x) - w).max() <= 1e-12 * np.abs(w).max()
