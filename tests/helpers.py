"""Shared builders for the tests: models, oracle groups and brute force."""
from __future__ import annotations

import itertools

import numpy as np

from lsamd import models
from lsamd.expr import Expr, Term
from oracle import oracle as O


def oracle_group(n, generators, sectors, spin_inversion):
    return O.Group.from_generators(n, generators, sectors, spin_inversion)


def weight_states(n, h):
    """All n-bit states of Hamming weight h, ascending (numpy brute force)."""
    xs = np.arange(1 << n, dtype=np.uint64)
    pc = np.zeros(xs.size, dtype=np.int64)
    for b in range(n):
        pc += ((xs >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return xs[pc == h]


def permute_array(perm, xs):
    """Vectorised permuteBits' (Benes.hs:338-345): bit i of g(x) = bit p[i] of x."""
    out = np.zeros_like(xs)
    for i, p in enumerate(perm):
        out |= ((xs >> np.uint64(p)) & np.uint64(1)) << np.uint64(i)
    return out


def brute_orbit_data(n, perms, chars, spin_inversion, xs):
    """For each x: min over images (and flipped images), and the stabiliser
    sum of chi (with -1 for flips in the antisymmetric sector)."""
    flip = np.uint64((1 << n) - 1)
    rep = xs.copy()
    stab = np.zeros(xs.size, dtype=np.complex128)
    for p, c in zip(perms, chars):
        y = permute_array(p, xs)
        rep = np.minimum(rep, y)
        stab += np.where(y == xs, c, 0)
        if spin_inversion:
            yf = y ^ flip
            rep = np.minimum(rep, yf)
            stab += np.where(yf == xs, spin_inversion * c, 0)
    return rep, stab


def burnside_dimension(n, h, perms, chars, spin_inversion):
    """Dimension of the symmetry sector: (1/|G|) sum_g conj(chi_g) fix_h(g)."""
    xs = weight_states(n, h)
    total = 0j
    flip = np.uint64((1 << n) - 1)
    for p, c in zip(perms, chars):
        y = permute_array(p, xs)
        total += np.conj(c) * np.count_nonzero(y == xs)
        if spin_inversion:
            total += spin_inversion * np.conj(c) * np.count_nonzero((y ^ flip) == xs)
    order = len(perms) * (2 if spin_inversion else 1)
    return total / order


def pauli_dense(n, expr_terms):
    """Dense matrix of a sum of Pauli strings built from numpy Kronecker
    products, independent of the term machinery: expr_terms is a list of
    (coefficient, [(kind, site), ...]) with kind in x, y, z, +, -."""
    I = np.eye(2)
    P = {
        "x": np.array([[0, 1], [1, 0]], dtype=complex),
        "y": np.array([[0, -1j], [1j, 0]], dtype=complex),
        "z": np.array([[1, 0], [0, -1]], dtype=complex),
        "+": np.array([[0, 1], [0, 0]], dtype=complex),
        "-": np.array([[0, 0], [1, 0]], dtype=complex),
    }
    # basis index = state integer; bit i = spin of site i; local basis |1> = up
    # ordered as (|1>, |0>) so sigma^z|1> = +|1>, sigma^+|0> = |1>.
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=complex)
    for coeff, ops in expr_terms:
        mats = [I.astype(complex)] * n
        for kind, site in ops:
            mats[site] = mats[site] @ P[kind]  # ops listed left to right, last acts first
        # kron with site n-1 as the most significant bit, local order (1, 0)
        full = np.array([[1.0 + 0j]])
        for site in reversed(range(n)):
            full = np.kron(full, mats[site])
        # reorder local basis (1,0) -> (0,1): index bits are inverted
        perm = np.arange(dim) ^ (dim - 1)
        M += coeff * full[np.ix_(perm, perm)]
    return M


def term_matrix(n_bits, terms):
    """Dense matrix of Term objects acting on all 2^n states."""
    dim = 1 << n_bits
    M = np.zeros((dim, dim), dtype=complex)
    for a in range(dim):
        for t in terms:
            r = t.apply(a)
            if r is not None:
                c, b = r
                M[b, a] += c
    return M


def sector_setup(n, h, gens, sectors, spin_inversion, expr):
    """Oracle group, representatives and term arrays for one sector."""
    g = oracle_group(n, gens, sectors, spin_inversion)
    lo = (1 << h) - 1
    hi = lo << (n - h) if not spin_inversion else lo << (n - h - 1)
    reps = g.enumerate(lo, hi)
    diag, off = expr.split()
    return g, reps, O.Terms(diag), O.Terms(off)


def sector_matrix_numpy(n, perms, chars, spin_inversion, Hfull, states, tol=1e-8):
    """Projected matrix P^dagger H P from explicit symmetrised vectors
    P_r = sum_g conj(chi_g) g|r> over orbits (numpy, independent of the
    reference's state_info formula)."""
    index = {int(s): i for i, s in enumerate(states)}
    flip = (1 << n) - 1
    seen = set()
    vecs = []
    for s in states:
        s = int(s)
        if s in seen:
            continue
        v = np.zeros(len(states), dtype=complex)
        for p, c in zip(perms, chars):
            y = int(permute_array(p, np.array([s], dtype=np.uint64))[0])
            v[index[y]] += np.conj(c)
            seen.add(y)
            if spin_inversion:
                v[index[y ^ flip]] += spin_inversion * np.conj(c)
                seen.add(y ^ flip)
        nv = np.linalg.norm(v)
        if nv > tol:
            vecs.append(v / nv)
    if not vecs:
        return np.zeros((0, 0))
    P = np.stack(vecs, axis=1)
    return P.conj().T @ Hfull @ P


def _cycles(perm):
    n = len(perm)
    seen = [False] * n
    out = []
    for i in range(n):
        if not seen[i]:
            length = 0
            j = i
            while not seen[j]:
                seen[j] = True
                j = int(perm[j])
                length += 1
            out.append(length)
    return out


def _poly_coeff(cycle_lengths, h):
    """Coefficient of x^h in prod_c (1 + x^len(c)): states of weight h fixed
    by a permutation with these cycles."""
    coeffs = [0] * (h + 1)
    coeffs[0] = 1
    for L in cycle_lengths:
        for w in range(h, L - 1, -1):
            coeffs[w] += coeffs[w - L]
    return coeffs[h]


def burnside_by_cycles(n, h, perms, chars, spin_inversion):
    """Sector dimension (1/|G|) sum_g conj(chi_g) fix_h(g) from cycle
    structures (no enumeration): fix(g) counts weight-h states constant on
    the cycles of g; fix(flip g) requires even cycles, alternating, each
    contributing half its length (2 choices per cycle)."""
    total = 0j
    for p, c in zip(perms, chars):
        cyc = _cycles(p)
        total += np.conj(c) * _poly_coeff(cyc, h)
        if spin_inversion:
            if all(L % 2 == 0 for L in cyc) and sum(L // 2 for L in cyc) == h:
                total += spin_inversion * np.conj(c) * (2 ** len(cyc))
    order = len(perms) * (2 if spin_inversion else 1)
    return total / order


def chain_sector_matrices(n, h, expr, sector_list, norm_eps):
    """Oracle representatives and dense projected matrices of chain sectors
    (momentum k, spin inversion +1 / -1 / None), with the given zero
    threshold for stabiliser sums."""
    O.lib().ora_group_set_norm_eps.argtypes = [O.C.c_void_p, O.C.c_double]
    diag, off = expr.split()
    out = []
    for k, si in sector_list:
        perms, num, den, re, im = O.group_closure(n, [models.chain_translation(n)], [k])
        g = O.Group(n, perms, re, im, si or 0)
        O.lib().ora_group_set_norm_eps(g.h, norm_eps)
        lo = (1 << h) - 1
        hi = lo << (n - h - 1) if si else lo << (n - h)
        reps = g.enumerate(lo, hi)
        M = O.matrix_dense(g, n, si or 0, reps, O.Terms(diag), O.Terms(off))
        out.append((reps, M))
    return out


def chain_sector_spectrum(n, h, expr, k, si):
    """Spectrum of a chain sector from the full weight-h Hamiltonian projected
    onto explicitly symmetrised vectors (numpy only)."""
    perms, num, den, re, im = O.group_closure(n, [models.chain_translation(n)], [k])
    states = weight_states(n, h)
    full = term_matrix(n, sum(expr.split(), []))
    M = sector_matrix_numpy(n, perms, re + 1j * im, si or 0, full[np.ix_(states, states)], states)
    return np.linalg.eigvalsh(M) if M.size else np.zeros(0)
