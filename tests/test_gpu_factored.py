"""The translation-factorised orbit search of the fused kernels against the
element-by-element search (the reference's order, kernels/generator.cpp:10-166):
for groups that are torus translations times point elements with real +-1
characters the stored CSR must be bit-identical, and the on-the-fly product
must agree to rounding (its atomic adds have no fixed order)."""
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


CASES = [
    # name, n, h, generators, sectors, spin inversion, expected factorised
    ("chain16_k0_p_z", 16, 8, [models.chain_translation(16), models.chain_parity(16)], [0, 0], 1, True),
    ("chain20_kpi_m_zm", 20, 10, [models.chain_translation(20), models.chain_parity(20)], [10, 1], -1, True),
    ("chain18_k0_zm", 18, 9, [models.chain_translation(18)], [0], -1, True),
    ("square4_a1", 16, 8, models.square_generators(4), [0, 0, 0, 0, 0], 1, True),
    ("square4_b1_zm", 16, 8, models.square_generators(4), [0, 0, 1, 1, 0], -1, True),
    ("square4_kpi", 16, 8, models.square_generators(4), [2, 2, 0, 0, 0], 1, True),
    ("chain16_k4_complex", 16, 8, [models.chain_translation(16)], [4], 0, False),
]


def device_basis(n, h, gens, sectors, si):
    S = lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])
    b = lsamd.DeviceBasis(n, h, si if si else None, S)
    b.build()
    return S, b


@pytest.mark.parametrize("name,n,h,gens,sectors,si,factored", CASES, ids=[c[0] for c in CASES])
def test_factored_csr_is_bit_identical(engine, name, n, h, gens, sectors, si, factored):
    say("factored", name)
    try:
        S, b = device_basis(n, h, gens, sectors, si)
    except ValueError:
        pytest.skip("inconsistent sector")
    expr = models.j1j2_square(4, 1.0, 0.5) if name.startswith("square") else models.j1j2_chain(n, 1.0, 0.5)
    fast = lsamd.DeviceOperator(b, expr)
    slow = lsamd.DeviceOperator(b, expr)
    slow.set_factored(False)
    assert fast.uses_factored == factored
    assert not slow.uses_factored
    for prec in (128, 96):
        for op in (fast, slow):
            op.set_csr_packing(False)
            op.build_csr(prec)
        for a, c in zip(fast.csr_arrays(), slow.csr_arrays()):
            assert a.dtype == c.dtype and a.shape == c.shape
            assert np.array_equal(a.view(np.uint8), c.view(np.uint8))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(b.number_states)
    y_fast, y_slow = fast.matvec(x, "otf"), slow.matvec(x, "otf")
    assert np.abs(y_fast - y_slow).max() <= 1e-13 * np.abs(y_slow).max()
    # and both are the reference's product
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    d, o = expr.split()
    want = O.matvec_push(og, n, si, b.states, O.Terms(d), O.Terms(o), x)
    assert np.abs(fast.matvec(x, "csr") - want).max() <= 1e-12 * np.abs(want).max()


def test_square6_factored_csr_gives_the_same_product(engine):
    """Full 6x6 workload: the CSR built through the factorised search and the
    one built element by element give bit-identical products."""
    n, h = 36, 18
    S, b = device_basis(n, h, models.square_generators(6), [0, 0, 0, 0, 0], 1)
    expr = models.j1j2_square(6, 1.0, 0.5)
    x = lsamd.DeviceArray.from_numpy(np.random.default_rng(2).standard_normal(b.number_states))
    ys = []
    for factored in (True, False):
        op = lsamd.DeviceOperator(b, expr)
        op.set_factored(factored)
        assert op.uses_factored == factored
        nnz = op.build_csr(64)
        say("square6 factored", factored, nnz)
        y = lsamd.DeviceArray(8 * b.number_states)
        op.csr_matvec(x, y)
        ys.append(y.to_numpy(np.float64, (b.number_states,)))
        del op
    assert np.array_equal(ys[0].view(np.uint64), ys[1].view(np.uint64))


def test_fermionic_groups_keep_the_element_search(engine):
    n = 9
    gens = [np.concatenate([g, g + n]).astype(np.int32) for g in models.square_generators(3)]
    S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in gens])
    b = lsamd.DeviceBasis.fermions(n, 2, 2, S)
    b.build()
    op = lsamd.DeviceOperator(b, models.hubbard(n, models.square_nn_bonds(3), t=1.0, U=4.0))
    assert not op.uses_factored
