"""Full-size checks on the 6x6 J1-J2 workload: T x C4v x Z2 (576 elements),
Hamming weight 18, 15 804 956 states, the Hamiltonian the benchmark times.

The representative list is pinned exactly without enumerating all 4.5 G
candidates on the CPU: it has the Burnside count of representatives
(tests/test_dimensions.py), it is strictly ascending, and every entry is its
own orbit minimum with a nonzero norm (the device's state_info, a kernel
separate from the enumeration, on all of them; the oracle on a sample).  A
strictly ascending list of that many representatives can only be the sorted
list of all of them.  Three stretches of the candidate range are also
enumerated by the oracle and compared entry by entry.

The product is checked through size-independent properties (symmetry,
linearity, CSR = on-the-fly), sampled rows against the oracle, and the
block forms against the single-vector FP64 product."""
import sys

import numpy as np
import pytest

import lsamd
from lsamd import models
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_SITES, WEIGHT, DIM = 36, 18, 15804956


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def square6(engine):
    gens = models.square_generators(6)
    S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in gens])
    say("square6: build")
    b = lsamd.DeviceBasis(N_SITES, WEIGHT, 1, S)
    count = b.build()
    og = O.Group(N_SITES, S.elements(), S.characters.real, S.characters.imag, 1)
    return S, b, count, og


def test_square6_basis_is_the_sorted_list_of_all_representatives(square6):
    S, b, count, og = square6
    assert len(S) == 288 and count == DIM
    reps = b.states
    assert reps.size == DIM and np.all(reps[1:] > reps[:-1])
    assert all(bin(int(v)).count("1") == WEIGHT for v in reps[:: DIM // 997])
    # every entry is its own orbit minimum with a nonzero norm (device state_info)
    say("square6: state_info on every state")
    r, _, nrm = b.state_info(reps)
    assert np.array_equal(r, reps) and np.all(nrm > 0)
    del r, nrm
    # a sample through the oracle, bit for bit against the device kernels
    rng = np.random.default_rng(6)
    sample = reps[np.sort(rng.choice(DIM, size=4000, replace=False))]
    r_o, c_o, n_o = og.state_info(sample)
    r_d, c_d, n_d = b.state_info(sample)
    assert np.array_equal(r_o, sample) and np.array_equal(r_d, r_o)
    assert np.array_equal(n_d, n_o) and np.array_equal(c_d, c_o)
    f_o, w_o = og.is_representative(sample)
    f_d, w_d = b.is_representative(sample)
    assert np.all(f_o == 1) and np.array_equal(f_d, f_o) and np.array_equal(w_d, w_o)
    # random candidates of the range (mostly not representatives), flags and index
    ranks = rng.integers(0, O.state_to_index(b.max_state_estimate) + 1, 20000)
    cand = np.unique(np.array([O.index_to_state(int(i), WEIGHT) for i in ranks], dtype=np.uint64))
    f_o, w_o = og.is_representative(cand)
    f_d, w_d = b.is_representative(cand)
    assert np.array_equal(f_d, f_o) and np.array_equal(w_d, w_o)
    assert 0 < f_o.sum() < cand.size
    idx = b.index(cand)
    assert np.array_equal(idx, O.state_index(reps, cand))
    assert np.array_equal(idx >= 0, f_o == 1)
    # the whole list through the device index
    assert np.array_equal(b.index(reps), np.arange(DIM))
    # three stretches of a million candidates, exhaustively through the oracle
    for start in (b.min_state_estimate, int(reps[DIM // 2]), int(reps[DIM - 50000])):
        lo_rank = O.state_to_index(start)
        hi = O.index_to_state(lo_rank + 1_000_000, WEIGHT)
        say("square6: oracle enumeration from", start)
        want = og.enumerate(start, hi)
        got = reps[np.searchsorted(reps, start):np.searchsorted(reps, hi, side="right")]
        assert want.size > 0 and np.array_equal(got, want)


def test_square6_product_properties(square6):
    S, b, count, og = square6
    expr = models.j1j2_square(6, 1.0, 0.5)
    op = lsamd.DeviceOperator(b, expr)
    say("square6: csr")
    nnz = op.build_csr(96)
    assert nnz > 70 * DIM
    assert op.csr_info["packed"]
    op.autotune()
    rng = np.random.default_rng(36)
    x1 = rng.standard_normal(DIM)
    x2 = rng.standard_normal(DIM)
    y1 = op.matvec(x1, "csr")
    y2 = op.matvec(x2, "csr")
    scale = np.abs(y1).max()
    # symmetry <x2, H x1> = <H x2, x1>, linearity
    assert abs(np.dot(x2, y1) - np.dot(y2, x1)) <= 1e-12 * np.abs(y1).sum()
    y12 = op.matvec(2.0 * x1 - 0.5 * x2, "csr")
    assert np.abs(y12 - (2.0 * y1 - 0.5 * y2)).max() <= 1e-12 * np.abs(y12).max()
    # sampled rows against the oracle
    reps = b.states
    rows = np.sort(rng.choice(DIM, size=1500, replace=False))
    diag, off = expr.split()
    want = O.matvec_rows(og, reps, O.Terms(diag), O.Terms(off), x1, rows)
    assert np.abs(y1[rows] - want).max() <= 1e-12 * scale
    del reps
    # on-the-fly product = CSR product
    say("square6: otf")
    assert np.abs(op.matvec(x1, "otf") - y1).max() <= 1e-12 * scale
    # the block forms: 8 columns, column 0 = x1
    X = rng.standard_normal((DIM, 8))
    X[:, 0] = x1
    say("square6: block products")
    W64 = op.matvec_block(X, 64)
    assert np.abs(W64[:, 0] - y1).max() <= 1e-12 * scale
    W32 = op.matvec_block(X, 32)
    assert np.abs(W32 - W64).max() <= 2e-6 * np.abs(W64).max()
    del X, W64, W32


def test_square6_through_the_reference_entry_points(square6):
    """The drop-in path at full size: the reference-named basis build
    (ls_hs_build_representatives -> ls_chpl_enumerate_representatives) and
    product (ls_chpl_matrix_vector_product, host vectors) give the device
    basis and the device product."""
    S, b, count, og = square6
    expr = models.j1j2_square(6, 1.0, 0.5)
    say("square6: reference entry points")
    sb = lsamd.SpinBasis(N_SITES, WEIGHT, 1, S)
    sb.build()
    assert sb.number_states == DIM and np.array_equal(sb.states, b.states)
    hop = lsamd.Operator(sb, expr)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(DIM)
    y = hop.matvec(x)
    op = lsamd.DeviceOperator(b, expr)
    op.build_csr(64)
    want = op.matvec(x, "csr")
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    # a second call reuses the cached operator
    assert np.array_equal(hop.matvec(x), y)
