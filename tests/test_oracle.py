"""Pins the CPU oracle (oracle/oracle.c) before it is used as the checker.

Every check compares against something the oracle does not compute itself:
golden vectors from the reference's own C kernels, the reference's known
answers, closed forms, or brute force over explicit permutations.
"""
import numpy as np
import pytest

import helpers as H
from lsamd import models
from lsamd.expr import (fermion_annihilate, fermion_create, fermion_number, sigma_minus, sigma_plus,
                        sigma_x, sigma_y, sigma_z, sum_exprs)
from oracle import oracle as O


# --- golden vectors of kernels/reference.c and kernels/indexing.c ------------

def _golden_terms(golden, name, key):
    v = golden[f"{name}_{key}_v"]
    terms = [O.Terms([])]
    from lsamd.expr import Term
    ts = [Term(complex(v[i]), int(golden[f"{name}_{key}_m"][i]), int(golden[f"{name}_{key}_l"][i]),
               int(golden[f"{name}_{key}_r"][i]), int(golden[f"{name}_{key}_x"][i]),
               int(golden[f"{name}_{key}_s"][i])) for i in range(v.size)]
    del terms
    return ts


@pytest.mark.parametrize("name", ["spin", "fermion"])
def test_apply_diag_matches_reference_c(golden, name):
    terms = O.Terms(_golden_terms(golden, name, "diag"))
    a, xs = golden[f"{name}_alphas"], golden[f"{name}_xs"]
    assert np.array_equal(O.apply_diag(terms, a, xs), golden[f"{name}_diag_ys"])
    assert np.array_equal(O.apply_diag(terms, a, None), golden[f"{name}_diag_ys_nox"])


@pytest.mark.parametrize("name", ["spin", "fermion"])
def test_apply_off_diag_matches_reference_c(golden, name):
    terms = O.Terms(_golden_terms(golden, name, "off"))
    b, c, o = O.apply_off_diag(terms, golden[f"{name}_alphas"], golden[f"{name}_xs"])
    assert np.array_equal(o, golden[f"{name}_off_offsets"])
    assert np.array_equal(b, golden[f"{name}_off_betas"])
    assert np.array_equal(c, golden[f"{name}_off_coeffs"])


@pytest.mark.parametrize("case", ["idx_a", "idx_b", "idx_c"])
def test_state_index_matches_reference_c(golden, case):
    got = O.state_index(golden[f"{case}_reps"], golden[f"{case}_queries"])
    assert np.array_equal(got, golden[f"{case}_indices"])


# --- groups, characters, Benes ------------------------------------------------

@pytest.mark.parametrize("n,gens,order", [
    (32, [models.chain_translation(32), models.chain_parity(32)], 64),   # heisenberg_chain_32_symm.yaml
    (16, models.square_generators(4), 128),                               # heisenberg_square_4x4.yaml
    (36, models.square_generators(6), 288),                               # heisenberg_square_6x6.yaml
])
def test_group_orders_of_reference_models(n, gens, order):
    perms, num, den, re, im = O.group_closure(n, gens, [0] * len(gens))
    assert len(perms) == order
    # ascending order, identity first, closed under composition
    assert all(tuple(perms[i]) < tuple(perms[i + 1]) for i in range(len(perms) - 1))
    assert np.array_equal(perms[0], np.arange(n))
    s = {tuple(p) for p in perms}
    for a in perms[:: max(1, order // 16)]:
        for b in perms:
            assert tuple(b[a]) in s
    assert np.all(re == 1.0) and np.all(im == 0.0)


@pytest.mark.parametrize("n,k", [(8, 0), (8, 1), (8, 3), (12, 5), (10, 5)])
def test_characters_closed_form(n, k):
    perms, num, den, re, im = O.group_closure(n, [models.chain_translation(n)], [k])
    assert len(perms) == n
    for p, cr, ci in zip(perms, re, im):
        j = int(p[0])  # p = T^j: bit i of T^j(x) is bit i + j of x
        assert abs(cr + 1j * ci - np.exp(-2j * np.pi * k * j / n)) < 1e-14


def test_inconsistent_sectors_rejected():
    # a reflection with sector 1 combined with translations at k = 1 is fine,
    # but the same permutation listed twice with different sectors is not
    n = 6
    with pytest.raises(ValueError):
        O.group_closure(n, [models.chain_translation(n), models.chain_translation(n)], [0, 1])


def test_benes_fixed_vectors():
    # haskell/test/LatticeSymmetries/BenesSpec.hs:34-42
    assert O.permute_bits([0, 1, 2], 0b100) == 0b100
    assert O.permute_bits([1, 2, 0], 0b100) == 0b010
    assert O.permute_bits([1, 2, 0], 0b101) == 0b110
    assert O.permute_bits([1, 2, 3, 0, 5, 6, 7, 4, 9, 10, 11, 8], 0b100111010010) == 0b110011100001


# --- combinadics and enumeration ----------------------------------------------

def test_gosper_and_combinadics():
    n, h = 14, 5
    states = H.weight_states(n, h)
    v = int(states[0])
    for i, s in enumerate(states):
        assert v == int(s)
        assert O.state_to_index(v) == i
        assert O.index_to_state(i, h) == v
        if i + 1 < states.size:
            v = O.next_state(v)


SECTORS = [
    # n, h, generators, sectors, spin inversion
    (12, 6, [models.chain_translation(12), models.chain_parity(12)], [0, 0], 1),
    (12, 6, [models.chain_translation(12), models.chain_parity(12)], [6, 1], -1),
    (12, 5, [models.chain_translation(12)], [3], 0),          # complex characters
    (12, 6, [models.chain_translation(12)], [4], 1),
    (10, 5, [models.chain_translation(10), models.chain_parity(10)], [5, 1], -1),
    (16, 8, models.square_generators(4), [0, 0, 0, 0, 0], 1),
    (16, 8, models.square_generators(4), [0, 0, 1, 1, 0], 0),
    (9, 4, [models.chain_translation(9)], [2], 0),
]


@pytest.mark.parametrize("n,h,gens,sectors,si", SECTORS)
def test_enumeration_count_equals_burnside(n, h, gens, sectors, si):
    g = H.oracle_group(n, gens, sectors, si)
    lo = (1 << h) - 1
    hi = lo << (n - h) if not si else lo << (n - h - 1)
    reps = g.enumerate(lo, hi)
    perms, num, den, re, im = O.group_closure(n, gens, sectors)
    dim = H.burnside_dimension(n, h, perms, re + 1j * im, si)
    assert abs(dim.imag) < 1e-9
    assert len(reps) == int(round(dim.real))


@pytest.mark.parametrize("n,h,gens,sectors,si", SECTORS)
def test_state_info_matches_brute_force(n, h, gens, sectors, si):
    g = H.oracle_group(n, gens, sectors, si)
    perms, num, den, re, im = O.group_closure(n, gens, sectors)
    chars = re + 1j * im
    xs = H.weight_states(n, h)
    rep, c, norm = g.state_info(xs)
    brep, stab = H.brute_orbit_data(n, perms, chars, si, xs)
    assert np.array_equal(rep, brep)
    order = len(perms) * (2 if si else 1)
    # Mathematically zero stabiliser sums can come out as +-1e-16 in the
    # restated Halide order (characters like cos(2 pi / 3) are inexact); the
    # reference then yields sqrt of a tiny negative (NaN) or a tiny norm.
    # Those states are excluded here: only the restated order decides them.
    clear = np.abs(stab) > 1e-9
    assert np.allclose(norm[clear] ** 2 * order, stab.real[clear], atol=1e-12)
    assert np.all(np.isnan(norm[~clear]) | (norm[~clear] < 1e-7))
    # for states with a nonzero norm every minimiser carries the same
    # character: chi must be that of any element mapping x to its rep
    flip = np.uint64((1 << n) - 1)
    nz = np.flatnonzero(norm > 1e-8)
    for i in nz[:200]:
        x = xs[i]
        want = None
        for p, ch in zip(perms, chars):
            y = H.permute_array(p, np.array([x]))[0]
            if y == rep[i]:
                want = ch
                break
            if si and (y ^ flip) == rep[i]:
                want = si * ch
                break
        if rep[i] == x:
            want = 1.0
        assert abs(c[i] - want) < 1e-14
    # representatives: flag set exactly for rep == x with norm > 0
    f, nrm = g.is_representative(xs)
    assert np.array_equal(f.astype(bool)[clear], ((brep == xs) & (stab.real > 1e-9))[clear])


# --- known answers and spectra ------------------------------------------------

def test_getting_started_energy():
    # python/example/getting_started.py:8-51
    n = 10
    expr = sum_exprs([(sigma_plus(i) * sigma_minus(j) + sigma_plus(j) * sigma_minus(i)) * 2 +
                      sigma_z(i) * sigma_z(j) for i, j in models.chain_bonds(n)])
    g, reps, d, o = H.sector_setup(n, 5, [models.chain_translation(n), models.chain_parity(n)], [5, 1], -1, expr)
    M = O.matrix_dense(g, n, -1, reps, d, o)
    assert np.abs(M - M.conj().T).max() < 1e-12
    assert abs(np.linalg.eigvalsh(M)[0] - (-18.06178542)) < 5e-9


def test_sector_spectra_reassemble_full_spectrum():
    n, h = 12, 6
    expr = models.j1j2_chain(n, 1.0, 0.4)
    diag, off = expr.split()
    full_states = H.weight_states(n, h)
    Mfull = O.matrix_dense(None, n, 0, full_states, O.Terms(diag), O.Terms(off))
    want = np.sort(np.linalg.eigvalsh(Mfull))
    # independent projector construction (numpy) reassembles the spectrum
    got = []
    spectra = {}
    for k in range(n):
        perms, num, den, re, im = O.group_closure(n, [models.chain_translation(n)], [k])
        Ms = H.sector_matrix_numpy(n, perms, re + 1j * im, 0, Mfull, full_states)
        spectra[k] = np.linalg.eigvalsh(Ms)
        got.extend(spectra[k])
    assert len(got) == len(want)
    assert np.allclose(np.sort(got), want, atol=1e-10)
    # the reference's formula (oracle) reproduces the sector spectra where
    # its characters are exact (k = 0, 3, 6, 9: phases multiples of 1/4)
    for k in (0, 3, 6, 9):
        g, reps, d, o = H.sector_setup(n, h, [models.chain_translation(n)], [k], 0, expr)
        M = O.matrix_dense(g, n, 0, reps, d, o)
        assert np.abs(M - M.conj().T).max() < 1e-12
        assert np.allclose(np.linalg.eigvalsh(M), spectra[k], atol=1e-10)


def test_numerically_zero_norm_halts_like_the_reference():
    # k = 4 on 12 sites: the period-4 states have a stabiliser character sum
    # 1 + cos(-4 pi / 3) + cos(-8 pi / 3) that is +-1e-16 in floating point;
    # the reference's coefficient becomes NaN (sqrt of a negative) or huge,
    # passes its `c != 0` test and the index lookup halts
    # (DistributedMatrixVector.chpl:127-135).  The oracle does the same.
    n, h = 12, 6
    expr = models.j1j2_chain(n, 1.0, 0.4)
    g, reps, d, o = H.sector_setup(n, h, [models.chain_translation(n)], [4], 0, expr)
    with pytest.raises(ValueError):
        O.matvec_push(g, n, 0, reps, d, o, np.ones(reps.size))


def test_norm_tolerance_recovers_every_momentum_sector():
    """With the opt-in tolerance (stabiliser sums <= 1e-9 count as zero) the
    reference formula reproduces the numpy-projector spectra of all 12
    momenta of the 12-site J1-J2 chain, including the complex ones."""
    n, h = 12, 6
    expr = models.j1j2_chain(n, 1.0, 0.4)
    diag, off = expr.split()
    full_states = H.weight_states(n, h)
    Mfull = O.matrix_dense(None, n, 0, full_states, O.Terms(diag), O.Terms(off))
    O.lib().ora_group_set_norm_eps.argtypes = [O.C.c_void_p, O.C.c_double]
    for k in range(n):
        perms, num, den, re, im = O.group_closure(n, [models.chain_translation(n)], [k])
        g = O.Group(n, perms, re, im, 0)
        O.lib().ora_group_set_norm_eps(g.h, 1e-9)
        lo = (1 << h) - 1
        reps = g.enumerate(lo, lo << (n - h))
        M = O.matrix_dense(g, n, 0, reps, O.Terms(diag), O.Terms(off))
        Ms = H.sector_matrix_numpy(n, perms, re + 1j * im, 0, Mfull, full_states)
        assert reps.size == Ms.shape[0]
        assert np.abs(M - M.conj().T).max() < 1e-12
        assert np.allclose(np.linalg.eigvalsh(M), np.linalg.eigvalsh(Ms), atol=1e-10)


def test_push_matvec_is_real_part_of_projected_matrix():
    n, h = 12, 6
    expr = models.j1j2_chain(n, 1.0, 0.5)
    g, reps, d, o = H.sector_setup(n, h, [models.chain_translation(n), models.chain_parity(n)], [0, 0], 1, expr)
    M = O.matrix_dense(g, n, 1, reps, d, o)
    x = np.random.default_rng(1).standard_normal(reps.size)
    y = O.matvec_push(g, n, 1, reps, d, o, x)
    assert np.allclose(y, M.real @ x, rtol=1e-13, atol=1e-12)


def _hphi_spin(n, bonds_groups, extra):
    terms = []
    for J, bonds in bonds_groups:
        for i, j in bonds:
            terms.append((sigma_x(i) * sigma_x(j) + sigma_y(i) * sigma_y(j) + sigma_z(i) * sigma_z(j)) * (J / 4))
    return sum_exprs(terms + extra)


def test_hphi_kagome_energy():
    # test/01_spin_kagome: 9 sites, Hamming weight 5, E0 from HPhi
    bonds = [(1.0, [(1, 2), (1, 5), (4, 5), (4, 8), (7, 8), (7, 2)]),
             (0.5, [(0, 2), (2, 3), (3, 5), (5, 6), (6, 8), (8, 0)]),
             (0.5, [(0, 1), (1, 6), (3, 4), (4, 0), (6, 7), (7, 3)])]
    expr = _hphi_spin(9, bonds, [])
    diag, off = expr.split()
    states = H.weight_states(9, 5)
    M = O.matrix_dense(None, 9, 0, states, O.Terms(diag), O.Terms(off))
    assert abs(np.linalg.eigvalsh(M)[0] - (-3.0170209179017471)) < 1e-12


def test_hphi_ladder_field_energy():
    # test/02_spin_ladder_DM: 6 sites, no symmetries, 0.5 S^x field
    bonds = [(1.0, [(0, 2), (0, 4), (0, 1), (1, 3), (1, 5), (2, 4), (2, 3), (3, 5), (4, 5)])]
    expr = _hphi_spin(6, bonds, [sigma_x(i) * 0.25 for i in range(6)])
    diag, off = expr.split()
    states = np.arange(1 << 6, dtype=np.uint64)
    M = O.matrix_dense(None, 6, 0, states, O.Terms(diag), O.Terms(off))
    assert abs(np.linalg.eigvalsh(M)[0] - (-3.0527756377319952)) < 1e-12


def test_hphi_hubbard_energy():
    # test/04_hubbard_square: 8 sites, 4 up + 4 down, t = 1 (doubly counted
    # boundary hoppings as in the model file), U = 4
    n = 8
    bonds = [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (4, 7), (5, 6), (6, 7),
             (0, 4), (1, 5), (2, 6), (3, 7)]
    terms = []
    for i, j in bonds:
        for off in (0, n):
            terms.append(fermion_create(j + off) * fermion_annihilate(i + off) * (-1))
            terms.append(fermion_create(i + off) * fermion_annihilate(j + off) * (-1))
    terms += [fermion_number(i) * fermion_number(i + n) * 4.0 for i in range(n)]
    expr = sum_exprs(terms)
    up = H.weight_states(n, 4)
    states = np.sort((up[None, :] << np.uint64(n) | up[:, None]).ravel())
    diag, off = expr.split()
    M = O.matrix_dense(None, 2 * n, 0, states, O.Terms(diag), O.Terms(off))
    assert abs(np.linalg.eigvalsh(M)[0] - (-10.2529529552635879)) < 1e-11


def test_lanczos_cfg1_matches_dense():
    # Heisenberg chain N=16, S.S, h=8, k=0, P+, Z+ (257 states); 100-step Lanczos
    n = 16
    expr = models.heisenberg_chain(n, 1.0, scale=0.25)
    g, reps, d, o = H.sector_setup(n, 8, [models.chain_translation(n), models.chain_parity(n)], [0, 0], 1, expr)
    assert reps.size == 257
    M = O.matrix_dense(g, n, 1, reps, d, o).real
    e0 = np.linalg.eigvalsh(M)[0]
    e_l = O.lanczos_ground_state(lambda v: O.matvec_push(g, n, 1, reps, d, o, v), reps.size, steps=100)
    assert abs(e_l - e0) < 1e-10
    # the N=16 Heisenberg ring ground state (S.S): -7.1422964...
    assert abs(e0 - (-7.142296361)) < 1e-8
