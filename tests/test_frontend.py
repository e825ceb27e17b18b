"""The C++ model front-end (ls_amd_model_*) on the reference's YAML model
files (fixtures copied from chapel/data/).

The YAML reader is checked against PyYAML (an independent parser) and the
expression lowering against the Python term algebra, whose generators and
products are pinned by the reference's golden term vectors and known-answer
energies elsewhere in the suite."""
import glob
import os

import numpy as np
import pytest
import yaml

import lsamd
from lsamd import models
from lsamd.expr import (Expr, fermion_annihilate, fermion_create, fermion_number, sigma_minus, sigma_plus, sigma_x,
                        sigma_y, sigma_z, sum_exprs)

MODELS = os.path.join(os.path.dirname(__file__), "golden", "models")
FILES = sorted(glob.glob(os.path.join(MODELS, "*.yaml")))


def read(name):
    with open(os.path.join(MODELS, name), encoding="utf-8") as f:
        return f.read()


def terms_dict(expr: Expr):
    return {k: v for k, v in expr.terms.items()}


def same_terms(a: Expr, b: Expr, tol=1e-14):
    ka, kb = terms_dict(a), terms_dict(b)
    assert set(ka) == set(kb), (sorted(set(ka) ^ set(kb))[:5])
    for k in ka:
        assert abs(ka[k] - kb[k]) <= tol, (k, ka[k], kb[k])


# Python rendering of the reference's expression strings used in the model files
PRIMS = {"σˣ": sigma_x, "σʸ": sigma_y, "σᶻ": sigma_z, "σ⁺": sigma_plus, "σ⁻": sigma_minus,
         "Sˣ": lambda i: sigma_x(i) * 0.5, "Sʸ": lambda i: sigma_y(i) * 0.5, "Sᶻ": lambda i: sigma_z(i) * 0.5}
SUB = {c: str(d) for d, c in enumerate("₀₁₂₃₄₅₆₇₈₉")}


def python_expr(expression: str, sites):
    """c × P_a P_b for the two-site products of the model files."""
    parts = expression.split()
    coeff = 1.0
    if parts[0] not in PRIMS and not parts[0][:2] in PRIMS:
        coeff = float(parts[0])
        parts = parts[2:] if parts[1] == "×" else parts[1:]
    prims = [(PRIMS[p[:2]], int("".join(SUB[c] for c in p[2:]))) for p in parts]
    old = sorted({i for _, i in prims})
    out = []
    for tup in sites:
        mp = dict(zip(old, tup))
        e = Expr.identity()
        for f, i in prims:
            e = e * f(mp[i])
        out.append(e * coeff)
    return sum_exprs(out)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_model_files_match_pyyaml_and_python_terms(path):
    text = open(path, encoding="utf-8").read()
    ref = yaml.safe_load(text)
    m = lsamd.Model(text)
    b = ref["basis"]
    assert m.particle == "spin"
    assert m.number_sites == b["number_spins"]
    want_h = b.get("hamming_weight")
    if want_h is None and b.get("spin_inversion"):
        want_h = b["number_spins"] // 2
    assert m.hamming_weight == want_h
    assert m.spin_inversion == b.get("spin_inversion")
    syms = b.get("symmetries") or []
    assert len(m.symmetries) == len(syms)
    for (perm, sector), s in zip(m.symmetries, syms):
        assert list(perm) == s["permutation"] and sector == s["sector"]
    want = sum_exprs([python_expr(t["expression"], t["sites"]) for t in ref["hamiltonian"]["terms"]])
    same_terms(m.expr(), want)
    # the diagonal / off-diagonal split follows x == 0
    assert all(t.x == 0 for t in m.terms(-1, True)) and all(t.x != 0 for t in m.terms(-1, False))


def test_6x6_model_is_the_benchmark_hamiltonian():
    m = lsamd.Model(read("heisenberg_square_6x6.yaml"))
    same_terms(m.expr(), models.j1j2_square(6, 1.0, 0.0))
    gens = models.square_generators(6)
    assert [list(p) for p, _ in m.symmetries] == [list(g) for g in gens]


def test_expression_grammar():
    def one(text, particle="spin", n=4, extra=""):
        header = {"spin": f"number_spins: {n}",
                  "spinless-fermion": f"particle: spinless-fermion\n  number_sites: {n}",
                  "spinful-fermion": f"particle: spinful-fermion\n  number_sites: {n}\n  number_particles: [1, 1]"}
        y = f"basis:\n  {header[particle]}\nhamiltonian:\n  terms:\n    - expression: \"{text}\"\n{extra}"
        return lsamd.Model(y).expr()

    same_terms(one("σˣ₀ σˣ₁"), sigma_x(0) * sigma_x(1))
    same_terms(one("2 σᶻ₀ - 0.5 × σʸ₁ σʸ₂ + σ⁺₃ σ⁻₀"),
               sigma_z(0) * 2 - sigma_y(1) * sigma_y(2) * 0.5 + sigma_plus(3) * sigma_minus(0))
    same_terms(one("(σˣ₀ + σʸ₀) σᶻ₁"), (sigma_x(0) + sigma_y(0)) * sigma_z(1))
    same_terms(one("\\\\sigma^x_0 S^z_1 + Sx2"), sigma_x(0) * sigma_z(1) * 0.5 + sigma_x(2) * 0.5)
    # same-site products stay in non-branching form (the two projectors of sigma^x sigma^x = 1)
    same_terms(one("σˣ₀ σˣ₀"), sigma_x(0) * sigma_x(0))
    assert sorted(one("σˣ₀ σˣ₀").terms.items()) == [((1, 0, 0, 0, 0), 1.0), ((1, 1, 1, 0, 0), 1.0)]
    same_terms(one("σ⁺₁ σ⁺₁ + σᶻ₂"), sigma_z(2))
    same_terms(one("Sˣ₁₀ Sˣ₁₁", n=12), sigma_x(10) * sigma_x(11) * 0.25)
    # fermions: the spinful string of the reference's own test (python/run_tests.py:123)
    same_terms(one("2.0 × n↑₀ n↓₀ + -1.0 × c†↑₀ c↑₁ + 1.0 × c↑₀ c†↑₁ + 2.0 × n↑₁ n↓₁ + -1.0 × c†↓₀ c↓₁ "
                   "+ 1.0 × c↓₀ c†↓₁", "spinful-fermion", n=2),
               fermion_number(0) * fermion_number(2) * 2 - fermion_create(0) * fermion_annihilate(1)
               + fermion_annihilate(0) * fermion_create(1) + fermion_number(1) * fermion_number(3) * 2
               - fermion_create(2) * fermion_annihilate(3) + fermion_annihilate(2) * fermion_create(3))
    # index before the spin label, and ASCII spellings
    same_terms(one("c†₀↑ c₁↓ + c^\\\\dagger_1\\\\down c_0\\\\up", "spinful-fermion", n=2),
               fermion_create(0) * fermion_annihilate(3) + fermion_create(3) * fermion_annihilate(0))
    same_terms(one("c†₀ c₂ + c†₂ c₀ - 0.5 n₁", "spinless-fermion", n=3),
               fermion_create(0) * fermion_annihilate(2) + fermion_create(2) * fermion_annihilate(0)
               - fermion_number(1) * 0.5)
    # site replication over the sorted distinct indices of the expression
    same_terms(one("σᶻ₃ σᶻ₁", extra="      sites: [[0, 1], [2, 3]]\n"),
               sigma_z(1) * sigma_z(0) + sigma_z(3) * sigma_z(2))


@pytest.mark.parametrize("text,match", [
    ("basis:\n  number_spins: 4\nhamiltonian:\n  terms:\n    - expression: \"σˣ₀ σ₁\"\n", "expression"),
    ("basis:\n  number_spins: 4\nhamiltonian:\n  terms:\n    - expression: \"σˣ₀ σˣ₇\"\n", "out of range"),
    ("basis:\n  number_spins: 4\nhamiltonian:\n  terms:\n    - expression: \"σˣ₀ σˣ₁\"\n"
     "      sites: [[0, 1, 2]]\n", "wrong number of site indices"),
    ("basis:\n  number_spins: 4\nhamiltonian:\n  terms:\n    - expression: \"c†₀ c₁\"\n", "particle"),
    ("basis:\n  number_spins: 4\n  hamming_weight: 5\n", "Hamming weight"),
    ("basis:\n  number_spins: 4\n  hamming_weight: 1\n  spin_inversion: 1\n", "spin inversion"),
    ("basis:\n  number_spins: 4\n  symmetries:\n    - permutation: [1, 2, 3]\n      sector: 0\n", "permutation"),
    ("basis:\n  number_spins: 4\nhamiltonian:\n  terms:\n    - expression: \"σˣ₀\"\n      sites: *nowhere\n",
     "alias"),
    ("basis:\n  number_spins: [4\n", "bracket"),
])
def test_model_errors(text, match):
    with pytest.raises(ValueError, match=match):
        lsamd.Model(text)


def test_hamiltonian_energy_from_model_file():
    """Dense ground state of the 4-site ring from its model file (sigma.sigma
    Heisenberg ring: E0 = -8)."""
    m = lsamd.Model(read("heisenberg_chain_4.yaml"))
    n = m.number_sites
    H = np.zeros((1 << n, 1 << n), dtype=complex)
    for t in m.terms(-1, True) + m.terms(-1, False):
        for a in range(1 << n):
            r = t.apply(a)
            if r is not None:
                H[r[1], a] += r[0]
    assert np.allclose(H, H.conj().T)
    assert abs(np.linalg.eigvalsh(H)[0] + 8.0) < 1e-12
