"""Host vectors through the reference's product entry point
(ls_chpl_matrix_vector_product) at sizes that take the pinned staging copies
(vectors above 4 MiB, moved in 32 MiB chunks, the last one partial):

- an operator with off-diagonal terms only, whose product accumulates into
  the caller's y, so y travels to the device as well as back;
- two host threads calling at once: one holds the staging buffers and the
  cached vectors, the other takes plain copies and vectors of its own.

Both agree with the device API's product on device-resident vectors.
"""
import threading

import numpy as np
import pytest

import lsamd
from lsamd import models
from lsamd.expr import sigma_minus, sigma_plus, sum_exprs

pytestmark = pytest.mark.gpu

N, H = 26, 13  # 10 400 600 states: 83.2 MB per vector


def xy_chain(n):
    """sum over bonds of s+_i s-_j + s-_i s+_j: no diagonal terms."""
    return sum_exprs(sigma_plus(i) * sigma_minus(j) + sigma_minus(i) * sigma_plus(j)
                     for i, j in models.chain_bonds(n))


@pytest.fixture(scope="module")
def setup(engine):
    print("staging: chain", N, flush=True)
    expr = xy_chain(N)
    sb = lsamd.SpinBasis(N, H)
    sb.build()
    hop = lsamd.Operator(sb, expr)
    db = lsamd.DeviceBasis(N, H)
    db.build()
    dop = lsamd.DeviceOperator(db, expr)
    assert sb.number_states == db.number_states == 10400600
    return hop, dop, sb.number_states


def reference_product(dop, x, y0):
    dx = lsamd.DeviceArray.from_numpy(x)
    dy = lsamd.DeviceArray.from_numpy(y0)
    dop.matvec_otf(dx, dy, accumulate=True)
    return dy.to_numpy(np.float64, x.shape)


def test_accumulating_product_through_staging(setup):
    hop, dop, n = setup
    rng = np.random.default_rng(11)
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(n)
    want = reference_product(dop, x, y0)
    y = y0.copy()
    hop.matvec(x, out=y)
    scale = np.abs(want).max()
    assert np.abs(y - want).max() <= 1e-12 * scale
    # the cached vectors are reused; a second call from a fresh y0 gives the same
    y2 = y0.copy()
    hop.matvec(x, out=y2)
    assert np.array_equal(y2, y)


def test_two_threads_at_once(setup):
    hop, dop, n = setup
    rng = np.random.default_rng(12)
    xs = [rng.standard_normal(n) for _ in range(2)]
    ys = [np.zeros(n) for _ in range(2)]
    errors = []

    def run(k):
        try:
            for _ in range(3):
                ys[k][:] = 0.0
                hop.matvec(xs[k], out=ys[k])
        except Exception as e:  # noqa: BLE001  (reported below)
            errors.append(e)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(2):
        want = reference_product(dop, xs[k], np.zeros(n))
        assert np.abs(ys[k] - want).max() <= 1e-12 * np.abs(want).max()


def test_chunked_product_has_the_bits_of_one_launch(setup):
    """The reference entry point runs its CSR product in row chunks, each
    chunk of y copied to the host while the next is computed; every row gets
    the bits of the single-launch product under the same schedule."""
    hop, dop, n = setup
    csr = lsamd.DeviceOperator(dop.basis, xy_chain(N))
    csr.build_csr(64)  # the same default schedule as the cached operator
    x = np.random.default_rng(13).standard_normal(n)
    y = np.zeros(n)
    hop.matvec(x, out=y)
    assert np.array_equal(y, csr.matvec(x, "csr"))
