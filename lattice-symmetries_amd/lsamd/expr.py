"""Operator expressions as sums of non-branching terms.

A non-branching term (v, m, l, r, x, s) maps |a> to v (-1)^popcount(a & s)
|a ^ x> when (a & m) == r (haskell/src/LatticeSymmetries/NonbranchingTerm.hs
:17-33).  Products follow the composition rule of NonbranchingTerm.hs:43-58
and the generators of Generator.hs:229-247; sigma^x = sigma^+ + sigma^-,
sigma^y = -i sigma^+ + i sigma^-, S = sigma / 2 (Parser.hs:242-260).

This is the engine's small front-end: it produces the term arrays that the
reference's Haskell layer would hand to the C ABI for the models used here.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Iterable, Tuple

Key = Tuple[int, int, int, int, int]  # (m, l, r, x, s)


def _popcount(v: int) -> int:
    return bin(v).count("1")


@dataclass(frozen=True)
class Term:
    v: complex
    m: int
    l: int
    r: int
    x: int
    s: int

    def compose(self, b: "Term") -> "Term":
        """Operator product self * b (b acts first), as NonbranchingTerm.hs:43-58.

        b maps alpha to alpha ^ x_b with sign (-1)^(alpha . s_b), then self
        applies with sign (-1)^((alpha ^ x_b) . s_a); on the masked bits alpha
        equals the composite r, so the constant part of the sign is
        popcount(r & (s_a ^ s_b)) + popcount(x_b & s_a) (mod 2).  The
        reference writes the first term as popcount(r & s_a & s_b), which is
        the same for the index-ordered products its canonical form builds
        (r & s_a = 0 there); this form also holds for products such as
        c+_j c_i with j > i."""
        a = self
        m = a.m | b.m
        r = b.r | (a.r & ~b.m)
        l = a.l | (b.l & ~a.m)
        x = l ^ r
        s = (a.s ^ b.s) & ~m
        p = _popcount(r & (a.s ^ b.s)) + _popcount(b.x & a.s)
        if (a.r ^ b.l) & a.m & b.m:
            v = 0j
        else:
            v = ((-1) ** p) * a.v * b.v
        return Term(v, m, l, r, x, s)

    def apply(self, alpha: int):
        """(coefficient, beta) of this term on basis state alpha, or None."""
        if (alpha & self.m) != self.r:
            return None
        sign = -1 if _popcount(alpha & self.s) % 2 else 1
        return self.v * sign, alpha ^ self.x


class Expr:
    """A polynomial of non-branching terms."""

    def __init__(self, terms: Dict[Key, complex] | None = None):
        self.terms: Dict[Key, complex] = {}
        for k, v in (terms or {}).items():
            if v != 0:
                self.terms[k] = complex(v)

    @staticmethod
    def of(term: Term) -> "Expr":
        return Expr({(term.m, term.l, term.r, term.x, term.s): term.v})

    @staticmethod
    def identity() -> "Expr":
        return Expr({(0, 0, 0, 0, 0): 1.0})

    def __add__(self, other: "Expr") -> "Expr":
        out = dict(self.terms)
        for k, v in other.terms.items():
            out[k] = out.get(k, 0) + v
        return Expr(out)

    def __sub__(self, other: "Expr") -> "Expr":
        return self + other * (-1)

    def __mul__(self, other) -> "Expr":
        if isinstance(other, Expr):
            out: Dict[Key, complex] = {}
            for ka, va in self.terms.items():
                ta = Term(va, *ka)
                for kb, vb in other.terms.items():
                    t = ta.compose(Term(vb, *kb))
                    if t.v != 0:
                        k = (t.m, t.l, t.r, t.x, t.s)
                        out[k] = out.get(k, 0) + t.v
            return Expr(out)
        return Expr({k: v * other for k, v in self.terms.items()})

    __rmul__ = __mul__

    def __neg__(self):
        return self * (-1)

    def adjoint(self) -> "Expr":
        """Hermitian conjugate: swaps l and r, conjugates v; for fermionic
        terms the sign mask keeps its meaning (the strings are real)."""
        out = {}
        for (m, l, r, x, s), v in self.terms.items():
            out[(m, r, l, x, s)] = out.get((m, r, l, x, s), 0) + v.conjugate()
        return Expr(out)

    def is_hermitian(self, tol: float = 0.0) -> bool:
        a = self.adjoint().terms
        keys = set(a) | set(self.terms)
        return all(abs(self.terms.get(k, 0) - a.get(k, 0)) <= tol for k in keys)

    def split(self):
        """(diagonal, off-diagonal) lists of Term (nbtIsDiagonal: x == 0)."""
        diag = [Term(v, *k) for k, v in sorted(self.terms.items()) if k[3] == 0]
        off = [Term(v, *k) for k, v in sorted(self.terms.items()) if k[3] != 0]
        return diag, off

    def __len__(self):
        return len(self.terms)


def _bit(i: int) -> int:
    return 1 << i


# Spin generators (Generator.hs:229-238)
def sigma_z(i: int) -> Expr:
    return Expr.of(Term(-1.0, 0, 0, 0, 0, _bit(i)))


def sigma_plus(i: int) -> Expr:
    return Expr.of(Term(1.0, _bit(i), _bit(i), 0, _bit(i), 0))


def sigma_minus(i: int) -> Expr:
    return Expr.of(Term(1.0, _bit(i), 0, _bit(i), _bit(i), 0))


def sigma_x(i: int) -> Expr:
    return sigma_plus(i) + sigma_minus(i)


def sigma_y(i: int) -> Expr:
    return sigma_plus(i) * (-1j) + sigma_minus(i) * 1j


# Fermion generators on bit b (Generator.hs:240-247); spinful sites put the
# up species in bits 0..n-1 and the down species in bits n..2n-1
# (Basis.hs:621-628).
def fermion_number(b: int) -> Expr:
    return Expr.of(Term(1.0, _bit(b), _bit(b), _bit(b), 0, 0))


def fermion_create(b: int) -> Expr:
    return Expr.of(Term(1.0, _bit(b), _bit(b), 0, _bit(b), _bit(b) - 1))


def fermion_annihilate(b: int) -> Expr:
    return Expr.of(Term(1.0, _bit(b), 0, _bit(b), _bit(b), _bit(b) - 1))


def heisenberg_bond(i: int, j: int, coupling: float = 1.0) -> Expr:
    """coupling * sigma_i . sigma_j (Pauli matrices)."""
    return (sigma_x(i) * sigma_x(j) + sigma_y(i) * sigma_y(j) + sigma_z(i) * sigma_z(j)) * coupling


def sum_exprs(exprs: Iterable[Expr]) -> Expr:
    out = Expr()
    for e in exprs:
        out = out + e
    return out
