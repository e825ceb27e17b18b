"""GPU parity for fermionic bases with permutation symmetries (workload 3:
Hubbard model on the 4x4 lattice, T x C4v, A1, with exact fermionic signs)."""
import sys

import numpy as np
import pytest

import helpers as H
import lsamd
from lsamd import models
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def spinful_generators(gens, n):
    return [np.concatenate([g, g + n]).astype(np.int32) for g in gens]


def spinful_states(n, up, down):
    a = H.weight_states(n, up)
    b = H.weight_states(n, down)
    return np.sort((b[:, None] << np.uint64(n) | a[None, :]).ravel())


def setup(L, up, down, sectors, norm_eps=None):
    """Fermionic basis on the L x L torus: the generators of the model files
    (translations, reflections, rotation), or the two translations alone when
    `sectors` has two entries."""
    n = L * L
    gens = spinful_generators(models.square_generators(L)[:len(sectors)], n)
    S = lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])
    og = O.Group(2 * n, S.elements(), S.characters.real, S.characters.imag, 0)
    O.lib().ora_group_set_fermionic.argtypes = [O.C.c_void_p, O.C.c_int]
    assert O.lib().ora_group_set_fermionic(og.h, 2) == 0
    db = lsamd.DeviceBasis.fermions(n, up, down, S)
    if norm_eps is not None:
        O.lib().ora_group_set_norm_eps.argtypes = [O.C.c_void_p, O.C.c_double]
        O.lib().ora_group_set_norm_eps(og.h, norm_eps)
        db.set_norm_tolerance(norm_eps)
    return n, S, og, db


# A1 and B1 of T x C4v at Gamma (real characters), and momentum (2 pi / 3, 0)
# of the translations alone (complex characters, numerically-zero stabiliser
# sums counted as zero with the tolerance 1e-9)
@pytest.mark.parametrize("sectors,norm_eps", [([0, 0, 0, 0, 0], None), ([0, 0, 1, 1, 0], None), ([1, 0], 1e-9)])
def test_hubbard_3x3_bit_exact(engine, sectors, norm_eps):
    say("hubbard 3x3", sectors)
    n, S, og, db = setup(3, 2, 2, sectors, norm_eps)
    if norm_eps is None:
        assert np.abs(S.characters.imag).max() < 1e-15  # +-1 up to the rounding of sin(pi)
    else:
        assert np.abs(S.characters.imag).max() > 0.5
    states = spinful_states(n, 2, 2)
    r, c, nrm = db.state_info(states)
    orr, oc, onrm = og.state_info(states)
    assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(nrm, onrm, equal_nan=True)
    f, fn = db.is_representative(states)
    of, ofn = og.is_representative(states)
    assert np.array_equal(f, of) and np.array_equal(fn, ofn, equal_nan=True)
    count = db.build()
    want = states[of == 1]
    assert count == want.size and np.array_equal(db.states, want)
    expr = models.hubbard(n, models.square_nn_bonds(3), t=1.0, U=4.0)
    diag, off = expr.split()
    M = O.matrix_dense(og, 2 * n, 0, want, O.Terms(diag), O.Terms(off))
    dop = lsamd.DeviceOperator(db, expr)
    rng = np.random.default_rng(7)
    xc = rng.standard_normal(count) + 1j * rng.standard_normal(count)
    yc = dop.matvec_c(xc)
    assert np.abs(yc - M @ xc).max() <= 1e-12 * np.abs(M @ xc).max()
    x = rng.standard_normal(count)
    dop.build_csr(64)
    assert np.abs(dop.matvec(x, "csr") - M.real @ x).max() <= 1e-12 * np.abs(M.real @ x).max()


def test_hubbard_4x4_a1_full_size(engine):
    say("hubbard 4x4 A1")
    n, S, og, db = setup(4, 7, 7, [0, 0, 0, 0, 0])
    count = db.build()
    assert count == 1022734  # fermionic Burnside count (tests/test_fermions.py)
    reps = db.states
    assert np.all(reps[1:] > reps[:-1])
    # every entry passes the device's is_representative (a kernel separate
    # from the enumeration): with the count and the strict order, the list is
    # exactly the sorted list of all representatives
    f_all, n_all = db.is_representative(reps)
    assert np.all(f_all == 1) and np.all(n_all > 0)
    rng = np.random.default_rng(5)
    sample = reps[rng.choice(reps.size, 5000, replace=False)]
    f, _ = og.is_representative(sample)
    assert np.all(f == 1)
    expr = models.hubbard(n, models.square_nn_bonds(4), t=1.0, U=4.0)
    diag, off = expr.split()
    dop = lsamd.DeviceOperator(db, expr)
    x = rng.standard_normal(count)
    xi = rng.standard_normal(count)
    y_otf = dop.matvec(x, "otf")
    dop.build_csr(64)
    y = dop.matvec(x, "csr")
    assert np.abs(y - y_otf).max() <= 1e-12 * np.abs(y).max()
    rows = np.sort(rng.choice(count, 2000, replace=False))
    want = O.matvec_rows(og, reps, O.Terms(diag), O.Terms(off), x, rows)
    assert np.abs(y[rows] - want).max() <= 1e-12 * np.abs(y).max()
    # complex x with the seeded N(0,1) real and imaginary parts: the A1
    # characters are real, so real and imaginary parts do not mix
    yc = dop.matvec_c(x + 1j * xi)
    assert np.abs(yc.real - y).max() <= 1e-12 * np.abs(y).max()
    yi = dop.matvec(xi, "csr")
    assert np.abs(yc.imag - yi).max() <= 1e-12 * np.abs(yi).max()
    # symmetry of the projected operator
    assert abs(np.dot(xi, y) - np.dot(yi, x)) <= 1e-12 * np.abs(y).sum()
