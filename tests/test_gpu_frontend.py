"""Model files -> device basis and operator (ls_amd_model_create_basis /
ls_amd_model_create_operator) against the oracle on the same terms."""
import os

import numpy as np
import pytest

import lsamd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(__file__), "golden", "models")


@pytest.mark.parametrize("name", ["heisenberg_kagome_12_symm.yaml", "issue_01.yaml", "heisenberg_square_4x4.yaml"])
def test_model_file_runs_on_the_device(engine, name):
    print("model", name, flush=True)
    with open(os.path.join(MODELS, name), encoding="utf-8") as f:
        m = lsamd.Model(f.read())
    b = m.device_basis()
    b.build()
    op = m.device_operator(b)
    reps = b.states
    gens = [p for p, _ in m.symmetries]
    sectors = [s for _, s in m.symmetries]
    S = lsamd.Symmetries([lsamd.Symmetry(p, s) for p, s in zip(gens, sectors)])
    og = O.Group(m.number_sites, S.elements(), S.characters.real, S.characters.imag, m.spin_inversion or 0)
    lo = (1 << m.hamming_weight) - 1
    hi = lo << (m.number_sites - m.hamming_weight - (1 if m.spin_inversion else 0))
    assert np.array_equal(reps, og.enumerate(lo, hi))
    diag, off = m.terms(-1, True), m.terms(-1, False)
    x = np.random.default_rng(3).standard_normal(reps.size)
    want = O.matvec_push(og, m.number_sites, m.spin_inversion or 0, reps, O.Terms(diag), O.Terms(off), x)
    y = op.matvec(x, "otf")
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    M = O.matrix_dense(og, m.number_sites, m.spin_inversion or 0, reps, O.Terms(diag), O.Terms(off))
    e0 = np.linalg.eigvalsh(M)[0]
    r = op.lanczos("otf", max_steps=200, k=1, tol=1e-13)
    assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)
