"""Fermionic permutation symmetries (the oracle's convention, pinned by brute
force: the reference has no fermionic symmetries, Basis.hs:941-948)."""
import numpy as np
import pytest

import helpers as H
from lsamd import models
from lsamd.expr import fermion_create
from oracle import oracle as O


def spinful_generators(gens, n):
    """Site permutations acting on both species: bits 0..n-1 up, n..2n-1 down."""
    return [np.concatenate([g, g + n]).astype(np.int32) for g in gens]


def apply_relabelled(perm, x, nbits):
    """U(g)|x> by moving every creation operator: |x> = c+_{b1} ... c+_{bk}|0>
    with b1 < ... < bk; U c+_b U^dagger = c+_{pi(b)} where pi(b) is the i with
    perm[i] = b.  Evaluated with the term algebra (Jordan-Wigner masks)."""
    inv = {int(perm[i]): i for i in range(nbits)}
    occ = [b for b in range(nbits) if (x >> b) & 1]
    state, coeff = 0, 1.0
    for b in reversed(occ):  # rightmost operator acts first
        (t,) = sum(fermion_create(inv[b]).split(), [])
        r = t.apply(state)
        assert r is not None
        c, state = r
        coeff *= c.real
    return coeff, state


@pytest.mark.parametrize("n", [6, 9])
def test_fermion_sign_matches_operator_algebra(n):
    rng = np.random.default_rng(n)
    perms = [rng.permutation(n).astype(np.int32) for _ in range(6)]
    g = O.Group(n, np.stack(perms), np.ones(6), np.zeros(6), 0)
    O.lib().ora_group_set_fermionic.argtypes = [O.C.c_void_p, O.C.c_int]
    O.lib().ora_fermion_sign.argtypes = [O.C.c_void_p, O.C.c_int, O.C.c_uint64]
    assert O.lib().ora_group_set_fermionic(g.h, 1) == 0
    for _ in range(200):
        x = int(rng.integers(0, 1 << n))
        e = int(rng.integers(0, 6))
        c, y = apply_relabelled(perms[e], x, n)
        assert y == O.permute_bits(perms[e], x)
        assert O.lib().ora_fermion_sign(g.h, e, x) == int(c)


def fermion_dim(n_sites, up, down, gens, sectors):
    perms, num, den, re, im = O.group_closure(n_sites, gens, sectors)

    def coeff(cyc, N):
        c = [0] * (N + 1)
        c[0] = 1
        for L in cyc:
            s = (-1) ** (L - 1)
            for w in range(N, L - 1, -1):
                c[w] += s * c[w - L]
        return c[N]

    tot = 0j
    for p, ch in zip(perms, re + 1j * im):
        cyc = H._cycles(p)
        tot += np.conj(ch) * coeff(cyc, up) * coeff(cyc, down)
    return tot / len(perms)


def test_hubbard_4x4_a1_dimension():
    # workload 3: 4x4, 7 up + 7 down, T x C4v, A1 at Gamma
    d = fermion_dim(16, 7, 7, models.square_generators(4), [0] * 5)
    assert abs(d - 1022734) < 1e-6


def spinful_states(n, up, down):
    a = H.weight_states(n, up)
    b = H.weight_states(n, down)
    return np.sort((b[:, None] << np.uint64(n) | a[None, :]).ravel())


def fermion_group(n, gens, sectors):
    g2 = spinful_generators(gens, n)
    perms, num, den, re, im = O.group_closure(2 * n, g2, sectors)
    g = O.Group(2 * n, perms, re, im, 0)
    O.lib().ora_group_set_fermionic.argtypes = [O.C.c_void_p, O.C.c_int]
    assert O.lib().ora_group_set_fermionic(g.h, 2) == 0
    return g, perms, re + 1j * im


def numpy_fermion_sector(nbits, perms, chars, Hfull, states):
    index = {int(s): i for i, s in enumerate(states)}
    seen, vecs = set(), []
    for s in states:
        s = int(s)
        if s in seen:
            continue
        v = np.zeros(len(states), dtype=complex)
        for p, c in zip(perms, chars):
            sign, y = apply_relabelled(p, s, nbits)
            v[index[y]] += np.conj(c) * sign
            seen.add(y)
        nv = np.linalg.norm(v)
        if nv > 1e-8:
            vecs.append(v / nv)
    P = np.stack(vecs, axis=1)
    return P.conj().T @ Hfull @ P


def test_hubbard_3x3_sectors():
    """3x3 Hubbard, 2 up + 2 down, T x C4v: projected spectra from the
    oracle's formula (with fermionic signs) equal those of explicit
    projectors, and the A1 count matches the fermionic Burnside count."""
    L, n = 3, 9
    expr = models.hubbard(n, models.square_nn_bonds(L), t=1.0, U=4.0)
    diag, off = expr.split()
    states = spinful_states(n, 2, 2)
    Hfull = O.matrix_dense(None, 2 * n, 0, states, O.Terms(diag), O.Terms(off))
    assert np.abs(Hfull - Hfull.conj().T).max() < 1e-12
    gens = models.square_generators(L)
    for sectors in ([0, 0, 0, 0, 0], [0, 0, 1, 1, 0], [0, 0, 1, 1, 1]):
        try:
            g, perms, chars = fermion_group(n, gens, sectors)
        except ValueError:
            continue  # inconsistent sector
        f, nrm = g.is_representative(states)
        reps = states[f == 1]
        want_dim = fermion_dim(n, 2, 2, gens, sectors)
        assert reps.size == int(round(want_dim.real))
        M = O.matrix_dense(g, 2 * n, 0, reps, O.Terms(diag), O.Terms(off))
        assert np.abs(M - M.conj().T).max() < 1e-12
        Ms = numpy_fermion_sector(2 * n, perms, chars, Hfull, states)
        assert np.allclose(np.linalg.eigvalsh(M), np.linalg.eigvalsh(Ms), atol=1e-10)


def test_hubbard_ring_momentum_spectra_reassemble():
    """6-site Hubbard ring, 3 up + 3 down: explicit fermionic projectors of
    all momenta reassemble the full spectrum; the oracle's formula agrees on
    the momenta with exact characters (k = 0, 3)."""
    n = 6
    expr = models.hubbard(n, models.chain_bonds(n), t=1.0, U=4.0)
    diag, off = expr.split()
    states = spinful_states(n, 3, 3)
    Hfull = O.matrix_dense(None, 2 * n, 0, states, O.Terms(diag), O.Terms(off))
    want = np.sort(np.linalg.eigvalsh(Hfull))
    got = []
    for k in range(n):
        g, perms, chars = fermion_group(n, [models.chain_translation(n)], [k])
        Ms = numpy_fermion_sector(2 * n, perms, chars, Hfull, states)
        ev = np.linalg.eigvalsh(Ms)
        got.extend(ev)
        if k in (0, 3):
            f, _ = g.is_representative(states)
            reps = states[f == 1]
            M = O.matrix_dense(g, 2 * n, 0, reps, O.Terms(diag), O.Terms(off))
            assert np.allclose(np.linalg.eigvalsh(M), ev, atol=1e-10)
    assert np.allclose(np.sort(got), want, atol=1e-10)
