"""Pins of the oracle's helper products and counters against its sequential
push-form product and dense matrix (themselves pinned against independent
numpy projectors, the reference's golden vectors and HPhi energies):

- ora_matvec_rows is row i of Re(M)^T x (the pairs alpha_i emits), which
  equals the push form Re(M) x exactly when Re(M) is symmetric;
- ora_matvec_push_parallel (the CPU baseline) is the push form restricted to
  a strided set of source rows;
- ora_count_projected_parallel counts what ora_enumerate_projected lists.
"""
import numpy as np
import pytest

import helpers as H
import test_fermions as F
from lsamd import models
from lsamd.expr import sigma_minus, sigma_plus, sigma_z
from oracle import oracle as O


def spin_sector(n, h, gens, sectors, si, norm_eps=0.0):
    perms, num, den, re, im = O.group_closure(n, gens, sectors)
    g = O.Group(n, perms, re, im, si)
    if norm_eps:
        O.lib().ora_group_set_norm_eps.argtypes = [O.C.c_void_p, O.C.c_double]
        O.lib().ora_group_set_norm_eps(g.h, norm_eps)
    lo = (1 << h) - 1
    hi = lo << (n - h) if not si else lo << (n - h - 1)
    return g, g.enumerate(lo, hi), lo, hi


def trivial_group(n, si):
    """The identity alone plus spin inversion: the kind-2 path of a basis with
    spin inversion only."""
    return O.Group(n, np.arange(n, dtype=np.int32)[None, :], np.ones(1), np.zeros(1), si)


def non_hermitian(n):
    return (sigma_plus(0) * sigma_minus(3) + sigma_z(1) * 0.3
            + sigma_plus(n - 2) * sigma_minus(0) * (0.4 + 0.2j))


CASES = [
    ("chain16_k0_p_z", 16, 8, [models.chain_translation(16), models.chain_parity(16)], [0, 0], 1,
     models.heisenberg_chain(16, 1.0, 0.25), 0.0),
    ("chain12_k3_complex", 12, 6, [models.chain_translation(12)], [3], 0, models.j1j2_chain(12, 1.0, 0.5), 0.0),
    ("chain12_k1_tolerance", 12, 6, [models.chain_translation(12)], [1], 0, models.j1j2_chain(12, 1.0, 0.5), 1e-9),
    ("square4_j1j2_z-", 16, 8, models.square_generators(4), [0, 0, 0, 0, 0], -1, models.j1j2_square(4, 1.0, 0.5),
     0.0),
]


def dense(g, n, si, reps, expr):
    d, o = expr.split()
    return O.matrix_dense(g, n, si, reps, O.Terms(d), O.Terms(o))


@pytest.mark.parametrize("name,n,h,gens,sectors,si,expr,eps", CASES, ids=[c[0] for c in CASES])
def test_matvec_rows_is_the_transposed_product(name, n, h, gens, sectors, si, expr, eps):
    g, reps, lo, hi = spin_sector(n, h, gens, sectors, si, eps)
    d, o = expr.split()
    M = dense(g, n, si, reps, expr)
    x = np.random.default_rng(1).standard_normal(reps.size)
    push = O.matvec_push(g, n, si, reps, O.Terms(d), O.Terms(o), x)
    rows = np.arange(reps.size)
    got = O.matvec_rows(g, reps, O.Terms(d), O.Terms(o), x, rows)
    scale = np.abs(push).max()
    assert np.abs(got - M.real.T @ x).max() <= 1e-12 * scale
    # Hermitian operators respecting the group: Re(M) symmetric, rows = push
    assert np.abs(M.real - M.real.T).max() <= 1e-13 * np.abs(M).max()
    assert np.abs(got - push).max() <= 1e-12 * scale
    # a strided subset of rows gives the same values
    sub = rows[1::7]
    assert np.array_equal(O.matvec_rows(g, reps, O.Terms(d), O.Terms(o), x, sub), got[sub])


def test_matvec_rows_on_an_operator_that_is_not_hermitian():
    """Where Re(M) is not symmetric the rows are Re(M)^T x, not the push form."""
    n, h = 12, 6
    g, reps, lo, hi = spin_sector(n, h, [models.chain_translation(n)], [0], 0)
    expr = non_hermitian(n)
    d, o = expr.split()
    M = dense(g, n, 0, reps, expr)
    x = np.random.default_rng(2).standard_normal(reps.size)
    push = O.matvec_push(g, n, 0, reps, O.Terms(d), O.Terms(o), x)
    got = O.matvec_rows(g, reps, O.Terms(d), O.Terms(o), x, np.arange(reps.size))
    assert np.abs(push - M.real @ x).max() <= 1e-12 * np.abs(push).max()
    assert np.abs(got - M.real.T @ x).max() <= 1e-12 * np.abs(push).max()
    assert np.abs(got - push).max() > 0.1


@pytest.mark.parametrize("si", [1, -1])
def test_inversion_only_equals_the_trivial_group_with_inversion(si):
    """BatchedOperator.chpl:166-205 (inversion only) against the full
    projection path on the group {identity} with spin inversion."""
    n, h = 12, 6
    reps = O.enumerate_unprojected(n, h, si)
    g = trivial_group(n, si)
    lo, hi = (1 << h) - 1, ((1 << h) - 1) << (n - h - 1)
    assert np.array_equal(g.enumerate(lo, hi), reps)
    expr = models.j1j2_chain(n, 1.0, 0.5) + non_hermitian(n)
    d, o = expr.split()
    x = np.random.default_rng(3).standard_normal(reps.size)
    inv = O.matvec_push(None, n, si, reps, O.Terms(d), O.Terms(o), x)
    full = O.matvec_push(g, n, si, reps, O.Terms(d), O.Terms(o), x)
    assert np.abs(inv - full).max() <= 1e-12 * np.abs(inv).max()
    M1 = dense(None, n, si, reps, expr)
    M2 = dense(g, n, si, reps, expr)
    assert np.abs(M1 - M2).max() <= 1e-13 * np.abs(M1).max()


def test_unprojected_push_form_is_the_term_matrix():
    """No projection: M is the operator's matrix on the weight-h states, built
    from the terms by brute force (helpers.term_matrix), not symmetric here."""
    n, h = 10, 5
    reps = O.enumerate_unprojected(n, h)
    assert np.array_equal(reps, H.weight_states(n, h))
    expr = non_hermitian(n) + models.heisenberg_chain(n, 1.0, 0.25)
    d, o = expr.split()
    full = H.term_matrix(n, d + o)
    want = full[np.ix_(reps.astype(np.int64), reps.astype(np.int64))]
    M = dense(None, n, 0, reps, expr)
    assert np.abs(M - want).max() <= 1e-14
    x = np.random.default_rng(4).standard_normal(reps.size)
    y = O.matvec_push(None, n, 0, reps, O.Terms(d), O.Terms(o), x)
    assert np.abs(y - want.real @ x).max() <= 1e-12 * np.abs(y).max()


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("name,n,h,gens,sectors,si,expr,eps", CASES[:2] + CASES[3:],
                         ids=[c[0] for c in CASES[:2] + CASES[3:]])
def test_matvec_push_parallel(name, n, h, gens, sectors, si, expr, eps, threads):
    g, reps, lo, hi = spin_sector(n, h, gens, sectors, si, eps)
    d, o = expr.split()
    x = np.random.default_rng(5).standard_normal(reps.size)
    want = O.matvec_push(g, n, si, reps, O.Terms(d), O.Terms(o), x)
    y = np.zeros(reps.size)
    O.matvec_push_parallel(g, reps, O.Terms(d), O.Terms(o), x, y, 0, reps.size, threads)
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    # strided source rows: the columns of M at those rows
    M = dense(g, n, si, reps, expr)
    sel = np.arange(1, reps.size, 3)
    y = np.zeros(reps.size)
    nnz = O.matvec_push_parallel(g, reps, O.Terms(d), O.Terms(o), x, y, 1, reps.size, threads, stride=3)
    w = M.real[:, sel] @ x[sel]
    assert np.abs(y - w).max() <= 1e-12 * np.abs(w).max()
    if np.abs(M.imag).max() == 0:
        # one per source row plus every nonzero off-diagonal pair it emits
        # (the oracle's pair list without diagonal terms)
        cap = reps.size * (len(o) + 1) + 1
        rows = np.zeros(cap, dtype=np.int64)
        cols = np.zeros(cap, dtype=np.int64)
        re = np.zeros(cap)
        im = np.zeros(cap)
        k = O.lib().ora_matrix_coo(g.h, 8, 2, n, si, reps.size, O._p(reps, O._u64p), *O.Terms([]).args(with_x=False),
                                   *O.Terms(o).args(), cap, O._p(rows, O._i64p), O._p(cols, O._i64p),
                                   O._p(re, O._f64p), O._p(im, O._f64p))
        assert nnz == sel.size + np.count_nonzero(np.isin(cols[:k], sel) & (re[:k] != 0))


def test_count_projected_parallel_counts_the_enumeration():
    n, h = 16, 8
    g, reps, lo, hi = spin_sector(n, h, [models.chain_translation(n), models.chain_parity(n)], [0, 0], 1)
    r0, r1 = O.state_to_index(lo), O.state_to_index(hi)
    for threads in (1, 3, 8):
        assert g.count_parallel(lo, r1 - r0 + 1, threads) == reps.size
    part = 5000
    upper = O.index_to_state(r0 + part - 1, h)
    assert g.count_parallel(lo, part, 4) == g.enumerate(lo, upper).size


def test_fermionic_sector_rows_and_push():
    """3x3 Hubbard, 2 up + 2 down, T x C4v A1 (fermionic signs)."""
    n = 9
    expr = models.hubbard(n, models.square_nn_bonds(3), t=1.0, U=4.0)
    d, o = expr.split()
    g, perms, chars = F.fermion_group(n, models.square_generators(3), [0, 0, 0, 0, 0])
    states = F.spinful_states(n, 2, 2)
    f, _ = g.is_representative(states)
    reps = states[f == 1]
    M = O.matrix_dense(g, 2 * n, 0, reps, O.Terms(d), O.Terms(o))
    x = np.random.default_rng(6).standard_normal(reps.size)
    push = O.matvec_push(g, 2 * n, 0, reps, O.Terms(d), O.Terms(o), x)
    rows = O.matvec_rows(g, reps, O.Terms(d), O.Terms(o), x, np.arange(reps.size))
    scale = np.abs(push).max()
    assert np.abs(push - M.real @ x).max() <= 1e-12 * scale
    assert np.abs(rows - push).max() <= 1e-12 * scale
    y = np.zeros(reps.size)
    O.matvec_push_parallel(g, reps, O.Terms(d), O.Terms(o), x, y, 0, reps.size, 4)
    assert np.abs(y - push).max() <= 1e-12 * scale
