"""Full-size checks on the 32-site chain (workload 2: k = 0, P+, Z+,
4 707 969 states) through size-independent properties and sampled rows."""
import sys

import numpy as np
import pytest

import lsamd
from lsamd import models
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def chain32(engine):
    n, h = 32, 16
    gens = [models.chain_translation(n), models.chain_parity(n)]
    S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in gens])
    say("chain32: build")
    b = lsamd.DeviceBasis(n, h, 1, S)
    count = b.build()
    return n, h, S, b, count


def test_chain32_basis(chain32):
    n, h, S, b, count = chain32
    assert count == 4707969  # Burnside (tests/test_dimensions.py)
    reps = b.states
    assert np.all(reps[1:] > reps[:-1])
    # every entry is its own orbit minimum with a nonzero norm (the device's
    # state_info, a kernel separate from the enumeration): with the Burnside
    # count and the strict order, the list is exactly the sorted list of all
    # representatives
    r_all, _, n_all = b.state_info(reps)
    assert np.array_equal(r_all, reps) and np.all(n_all > 0)
    rng = np.random.default_rng(3)
    sample = reps[rng.choice(reps.size, size=20000, replace=False)]
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 1)
    r, c, nrm = og.state_info(sample)
    assert np.array_equal(r, sample) and np.all(nrm > 0)
    f, _ = og.is_representative(sample)
    assert np.all(f == 1)
    # the first states of the enumeration range, exhaustively
    lo = b.min_state_estimate
    first = og.enumerate(lo, O.index_to_state(O.state_to_index(lo) + 3_000_000, h))
    assert np.array_equal(reps[:first.size], first)
    # index round trip on the device
    idx = b.index(reps)
    assert np.array_equal(idx, np.arange(reps.size))
    # absent states
    flips = reps[:1000] ^ np.uint64(1)
    got = b.index(flips)
    want = O.state_index(reps, flips)
    assert np.array_equal(got, want)


def test_chain32_product_properties(chain32):
    n, h, S, b, count = chain32
    expr = models.heisenberg_chain(n)
    op = lsamd.DeviceOperator(b, expr)
    say("chain32: csr")
    nnz = op.build_csr(96)
    assert nnz > 0
    rng = np.random.default_rng(42)
    x1 = rng.standard_normal(count)
    x2 = rng.standard_normal(count)
    y1 = op.matvec(x1, "csr")
    y2 = op.matvec(x2, "csr")
    # symmetry <x2, H x1> = <H x2, x1>, linearity
    assert abs(np.dot(x2, y1) - np.dot(y2, x1)) <= 1e-12 * np.abs(y1).sum()
    y12 = op.matvec(2.0 * x1 - 0.5 * x2, "csr")
    assert np.abs(y12 - (2.0 * y1 - 0.5 * y2)).max() <= 1e-12 * np.abs(y12).max()
    # sampled rows against the oracle
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 1)
    reps = b.states
    rows = np.sort(rng.choice(count, size=3000, replace=False))
    diag, off = expr.split()
    want = O.matvec_rows(og, reps, O.Terms(diag), O.Terms(off), x1, rows)
    assert np.abs(y1[rows] - want).max() <= 1e-12 * np.abs(y1).max()
    # on-the-fly product = CSR product
    say("chain32: otf")
    y_otf = op.matvec(x1, "otf")
    assert np.abs(y_otf - y1).max() <= 1e-12 * np.abs(y1).max()
