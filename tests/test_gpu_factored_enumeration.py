"""The translation-factorised basis filter against the element-by-element
one (the reference's is_representative with norm > 0,
kernels/generator.cpp:168-272, chapel/src/StatesEnumeration.chpl:242-268):
for groups of torus translations times point elements with characters +-1
both must list exactly the same representatives, and the oracle's
enumeration too.  The 6x6 benchmark basis is compared at full size."""
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


@pytest.fixture
def toggle(engine):
    lib = lsamd.load()
    yield lambda on: lib.ls_amd_set_factored_enumeration(1 if on else 0)
    lib.ls_amd_set_factored_enumeration(1)


# name, n, Hamming weight, spin inversion, generators, sectors
CASES = [
    ("chain16_k0_p_z", 16, 8, 1, [models.chain_translation(16), models.chain_parity(16)], [0, 0]),
    ("chain16_kpi_zm", 16, 8, -1, [models.chain_translation(16)], [8]),
    ("chain18_k0", 18, 9, 0, [models.chain_translation(18)], [0]),
    ("chain20_kpi_m", 20, 10, 0, [models.chain_translation(20), models.chain_parity(20)], [10, 1]),
    ("chain24_k0_p_z", 24, 12, 1, [models.chain_translation(24), models.chain_parity(24)], [0, 0]),
    ("square4_a1", 16, 8, 1, models.square_generators(4), [0, 0, 0, 0, 0]),
    ("square4_b1_zm", 16, 8, -1, models.square_generators(4), [0, 0, 1, 1, 0]),
    ("square4_kpi", 16, 8, 0, models.square_generators(4), [2, 2, 0, 0, 0]),
    ("square4_weight6", 16, 6, 0, models.square_generators(4), [0, 0, 0, 0, 0]),
]


def build(n, h, si, S):
    b = lsamd.DeviceBasis(n, h, si if si else None, S)
    b.build()
    return b


@pytest.mark.parametrize("name,n,h,si,gens,sectors", CASES, ids=[c[0] for c in CASES])
def test_factored_filter_lists_the_same_states(toggle, name, n, h, si, gens, sectors):
    say("enumeration", name)
    try:
        S = lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])
    except ValueError:
        pytest.skip("inconsistent sector")
    toggle(True)
    fast = build(n, h, si, S).states
    toggle(False)
    slow = build(n, h, si, S).states
    assert np.array_equal(fast, slow)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    lo = (1 << h) - 1
    hi = lo << (n - h) if not si else lo << (n - h - 1)
    if n <= 20:
        assert np.array_equal(fast, og.enumerate(lo, hi))


def test_square6_basis_both_filters(toggle):
    """15 804 956 representatives of the 6x6 benchmark sector, both ways."""
    S = lsamd.Symmetries([lsamd.Symmetry(g, 0) for g in models.square_generators(6)])
    toggle(True)
    b = build(36, 18, 1, S)
    fast = b.states
    toggle(False)
    slow = build(36, 18, 1, S).states
    say("square6", fast.size)
    assert fast.size == 15804956
    assert np.array_equal(fast, slow)
