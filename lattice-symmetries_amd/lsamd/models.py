"""Lattices, symmetry generators and Hamiltonians of the benchmark workloads.

Site numbering and symmetry generators follow the reference's model files
(chapel/data/heisenberg_chain_32_symm.yaml, heisenberg_square_4x4.yaml,
heisenberg_square_6x6.yaml): site (row, col) of an L x L lattice is
row * L + col; Hamiltonians use Pauli matrices ("σˣ₀ σˣ₁ + σʸ₀ σʸ₁ + σᶻ₀ σᶻ₁"
per bond) unless a coupling scale is given.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .expr import (Expr, fermion_annihilate, fermion_create, fermion_number, heisenberg_bond,
                   sum_exprs)

Bond = Tuple[int, int]


# --- chains -----------------------------------------------------------------

def chain_bonds(n: int, distance: int = 1) -> List[Bond]:
    return [(i, (i + distance) % n) for i in range(n)]


def chain_translation(n: int) -> np.ndarray:
    return np.array([(i + 1) % n for i in range(n)], dtype=np.int32)


def chain_parity(n: int) -> np.ndarray:
    return np.array(list(range(n))[::-1], dtype=np.int32)


def heisenberg_chain(n: int, J: float = 1.0, scale: float = 1.0) -> Expr:
    """J * scale * sum_i sigma_i . sigma_{i+1}; scale = 1/4 gives S . S."""
    return sum_exprs(heisenberg_bond(i, j, J * scale) for i, j in chain_bonds(n))


def j1j2_chain(n: int, J1: float = 1.0, J2: float = 0.5, scale: float = 1.0) -> Expr:
    return sum_exprs([heisenberg_bond(i, j, J1 * scale) for i, j in chain_bonds(n, 1)] +
                     [heisenberg_bond(i, j, J2 * scale) for i, j in chain_bonds(n, 2)])


# --- square lattice -----------------------------------------------------------

def square_translation_x(L: int) -> np.ndarray:
    return np.array([r * L + (c + 1) % L for r in range(L) for c in range(L)], dtype=np.int32)


def square_translation_y(L: int) -> np.ndarray:
    return np.array([((r + 1) % L) * L + c for r in range(L) for c in range(L)], dtype=np.int32)


def square_reflection_x(L: int) -> np.ndarray:
    return np.array([r * L + (L - 1 - c) for r in range(L) for c in range(L)], dtype=np.int32)


def square_reflection_y(L: int) -> np.ndarray:
    return np.array([(L - 1 - r) * L + c for r in range(L) for c in range(L)], dtype=np.int32)


def square_rotation(L: int) -> np.ndarray:
    return np.array([c * L + (L - 1 - r) for r in range(L) for c in range(L)], dtype=np.int32)


def square_nn_bonds(L: int) -> List[Bond]:
    """Nearest-neighbour bonds in the order of heisenberg_square_6x6.yaml."""
    bonds = []
    for r in range(L):
        for c in range(L):
            i = r * L + c
            bonds.append((i, r * L + (c + 1) % L))
            bonds.append((i, ((r + 1) % L) * L + c))
    return bonds


def square_nnn_bonds(L: int) -> List[Bond]:
    bonds = []
    for r in range(L):
        for c in range(L):
            i = r * L + c
            bonds.append((i, ((r + 1) % L) * L + (c + 1) % L))
            bonds.append((i, ((r + 1) % L) * L + (c - 1) % L))
    return bonds


def j1j2_square(L: int, J1: float = 1.0, J2: float = 0.5, scale: float = 1.0) -> Expr:
    return sum_exprs([heisenberg_bond(i, j, J1 * scale) for i, j in square_nn_bonds(L)] +
                     [heisenberg_bond(i, j, J2 * scale) for i, j in square_nnn_bonds(L)])


def random_bond_square(L: int, seed: int = 7, J1: float = 1.0, J2: float = 0.5, spread: float = 0.5,
                       scale: float = 1.0) -> Expr:
    """J1-J2 bonds of the L x L lattice with seeded couplings J (1 + spread u),
    u uniform in [-1/2, 1/2) per bond: every bond its own coupling (the
    lattice symmetries are broken, and the stored matrix holds many more
    distinct values than a symmetric model)."""
    rng = np.random.default_rng(seed)
    nn, nnn = square_nn_bonds(L), square_nnn_bonds(L)
    u = rng.random(len(nn) + len(nnn)) - 0.5
    return sum_exprs([heisenberg_bond(i, j, J1 * scale * (1 + spread * u[k])) for k, (i, j) in enumerate(nn)] +
                     [heisenberg_bond(i, j, J2 * scale * (1 + spread * u[len(nn) + k]))
                      for k, (i, j) in enumerate(nnn)])


def square_generators(L: int):
    """Translations, both reflections and the rotation of the model files."""
    return [square_translation_x(L), square_translation_y(L), square_reflection_x(L),
            square_reflection_y(L), square_rotation(L)]


# --- Hubbard -------------------------------------------------------------------

def hubbard(n_sites: int, bonds: List[Bond], t: float = 1.0, U: float = 4.0) -> Expr:
    """-t sum_{<ij>, s} (c+_is c_js + h.c.) + U sum_i n_i,up n_i,down with the up
    species in bits 0..n-1 and the down species in bits n..2n-1."""
    terms = []
    for (i, j) in bonds:
        for off in (0, n_sites):
            a, b = i + off, j + off
            terms.append((fermion_create(a) * fermion_annihilate(b)) * (-t))
            terms.append((fermion_create(b) * fermion_annihilate(a)) * (-t))
    for i in range(n_sites):
        terms.append((fermion_number(i) * fermion_number(i + n_sites)) * U)
    return sum_exprs(terms)


def square_bonds_open_aware(L: int) -> List[Bond]:
    """Periodic nearest-neighbour bonds without duplicates (L >= 3)."""
    return square_nn_bonds(L)
