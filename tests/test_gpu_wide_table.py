"""The wide value table: a real CSR with more than 256 but at most 4096
distinct stored values keeps the 32-bit column and a 16-bit index into a
table staged in LDS.  Random couplings on every bond of the 4x4 lattice, on
the translation-symmetric basis, give several hundred distinct values (each
coupling times the sector's norm ratios).

The wide format must store the same matrix as the unpacked one (the
exported CSR is equal), its FP64 single-vector and block products and its
FP32 block product must equal the unpacked format's bit for bit (the same
values in the same order), and the product must match the oracle's push
form under every schedule.
"""
import numpy as np
import pytest

import lsamd
from lsamd import models
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L, H = 4, 8
RTOL64 = 1e-12


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup(engine):
    n = L * L
    S = lsamd.Symmetries([lsamd.Symmetry(models.square_translation_x(L), 0),
                          lsamd.Symmetry(models.square_translation_y(L), 0)])
    db = lsamd.DeviceBasis(n, H, None, S)
    db.build()
    expr = models.random_bond_square(L, seed=3)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 0)
    return n, db, expr, og


def test_wide_table_is_chosen_and_stores_the_same_matrix(setup):
    n, db, expr, og = setup
    plain = lsamd.DeviceOperator(db, expr)
    plain.set_csr_packing(False)
    plain.build_csr(96)
    wide = lsamd.DeviceOperator(db, expr)
    wide.build_csr(96)
    info = wide.csr_info
    say("wide csr info", info)
    assert info["wide"] and not info["packed"] and 256 < info["table_size"] <= 4096
    assert not plain.csr_info["wide"] and not plain.csr_info["packed"]
    for a, b in zip(wide.csr_arrays(), plain.csr_arrays()):
        assert np.array_equal(a, b)
    reps = db.states
    rng = np.random.default_rng(5)
    x = rng.standard_normal(reps.size)
    X = rng.standard_normal((reps.size, 8))
    diag, off = expr.split()
    want = O.matvec_push(og, n, 0, reps, O.Terms(diag), O.Terms(off), x)
    # the same bits as the unpacked format, in every product form
    assert np.array_equal(wide.matvec(x, "csr"), plain.matvec(x, "csr"))
    assert np.array_equal(wide.matvec_block(X, 32), plain.matvec_block(X, 32))
    assert np.array_equal(wide.matvec_block(X, 64), plain.matvec_block(X, 64))
    for flags in (0, 1, 2, 3, 6, 7):
        for lanes in (1, 8, 64):
            say("wide schedule", lanes, flags)
            wide.set_schedule(lanes, flags, lanes, flags)
            y = wide.matvec(x, "csr")
            assert np.abs(y - want).max() <= RTOL64 * np.abs(want).max()
    wide.autotune()
    assert np.abs(wide.matvec(x, "csr") - want).max() <= RTOL64 * np.abs(want).max()


def test_wide_table_serves_more_than_2_to_24_rows(engine):
    """A matrix with few distinct values but 2^24 rows or more cannot pack
    its index into the column word: the unprojected 28-site chain (40 116 600
    rows, a handful of values) takes the wide table, and its CSR product
    matches the on-the-fly product (an independent path, no CSR)."""
    n, h = 28, 14
    db = lsamd.DeviceBasis(n, h)
    assert db.build() == 40116600
    op = lsamd.DeviceOperator(db, models.heisenberg_chain(n, 1.0, 0.5))
    op.build_csr(64)
    info = op.csr_info
    say("28-site chain csr info", info)
    assert info["wide"] and info["table_size"] <= 256
    x = np.random.default_rng(9).standard_normal(40116600)
    y = op.matvec(x, "csr")
    want = op.matvec(x, "otf")
    assert np.abs(y - want).max() <= RTOL64 * np.abs(want).max()
