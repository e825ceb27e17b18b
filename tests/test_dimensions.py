"""Sector dimensions of the benchmark workloads by Burnside's lemma from the
cycle structure of every group element (no enumeration, no orbit code)."""
import numpy as np
import pytest

import helpers as H
from lsamd import models
from oracle import oracle as O


def dim(n, h, gens, sectors, si):
    perms, num, den, re, im = O.group_closure(n, gens, sectors)
    d = H.burnside_by_cycles(n, h, perms, re + 1j * im, si)
    assert abs(d.imag) < 1e-6
    return int(round(d.real))


def test_cycle_formula_matches_enumeration_on_small_sectors():
    for n, h, gens, sectors, si in [
        (12, 6, [models.chain_translation(12), models.chain_parity(12)], [0, 0], 1),
        (12, 6, [models.chain_translation(12), models.chain_parity(12)], [6, 1], -1),
        (16, 8, models.square_generators(4), [0, 0, 0, 0, 0], 1),
        (14, 7, [models.chain_translation(14)], [3], 1),
    ]:
        perms, num, den, re, im = O.group_closure(n, gens, sectors)
        assert dim(n, h, gens, sectors, si) == int(round(H.burnside_dimension(n, h, perms, re + 1j * im, si).real))


def test_workload_dimensions():
    # chain 16, S.S, k=0 P+ Z+ (workload 1)
    assert dim(16, 8, [models.chain_translation(16), models.chain_parity(16)], [0, 0], 1) == 257
    # chain 32, k=0 P+ Z+ (workload 2; chapel/data/heisenberg_chain_32_symm.yaml)
    assert dim(32, 16, [models.chain_translation(32), models.chain_parity(32)], [0, 0], 1) == 4707969
    # 6x6, T x C4v x Z2 trivial sector (workload 4; heisenberg_square_6x6.yaml)
    assert dim(36, 18, models.square_generators(6), [0] * 5, 1) == 15804956


def test_chain28_sectors_sum_to_full_space():
    # workload 5: 56 sectors (k = 0..27, spin inversion +-1) of the N=28 chain
    n, h = 28, 14
    total = 0
    sizes = []
    for k in range(n):
        for si in (1, -1):
            d = dim(n, h, [models.chain_translation(n)], [k], si)
            sizes.append(d)
            total += d
    from math import comb
    assert total == comb(28, 14)
    assert min(sizes) == 716014 and max(sizes) == 716728
