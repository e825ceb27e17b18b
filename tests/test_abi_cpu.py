"""CPU checks of the C-ABI library and the host layer (no GPU calls)."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import helpers as H
from lsamd import _abi, models
from lsamd.expr import (Expr, fermion_annihilate, fermion_create, sigma_minus, sigma_plus, sigma_x,
                        sigma_y, sigma_z, sum_exprs)
from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    declared = _abi.declared_functions()
    assert len(declared) > 60
    missing = [name for name in declared if not hasattr(lib, name)]
    assert missing == []
    # and the ctypes signatures cover the same set
    assert sorted(_abi.SIGNATURES) == declared


def test_reference_entry_points_are_all_present():
    # the symbol set a replacement of libkernels + the Chapel library exports
    names = """ls_hs_fatal_error ls_hs_set_exception_handler ls_hs_error
    ls_hs_internal_destroy_external_array ls_hs_internal_read_refcount ls_hs_internal_write_refcount
    ls_hs_internal_inc_refcount ls_hs_internal_dec_refcount ls_internal_operator_apply_diag_x1
    ls_internal_operator_apply_off_diag_x1 ls_hs_state_index ls_hs_is_representative ls_hs_state_info
    ls_hs_build_representatives ls_hs_unchecked_set_representatives ls_hs_internal_get_chpl_kernels
    ls_hs_internal_set_chpl_kernels ls_hs_create_state_index_binary_search_kernel_data
    ls_hs_destroy_state_index_binary_search_kernel_data ls_hs_state_index_binary_search_kernel
    ls_internal_create_halide_kernel_data ls_internal_destroy_halide_kernel_data
    ls_hs_is_representative_halide_kernel ls_hs_state_info_halide_kernel ls_chpl_init ls_chpl_finalize
    ls_chpl_enumerate_representatives ls_chpl_operator_apply_diag ls_chpl_operator_apply_off_diag
    ls_chpl_matrix_vector_product""".split()
    lib = _abi.load()
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_reference_header():
    # sizes / offsets of kernels/lattice_symmetries_types.h on x86-64 (gcc),
    # recorded from the reference header
    want = {
        "basis": (72, 24, 32, 40, 64),
        "group": (56, 16, 48),
        "terms": (56, 8, 48),
        "op": (40, 16, 32),
    }
    A = _abi
    got = {
        "basis": (C.sizeof(A.ls_hs_basis), A.ls_hs_basis.state_index_is_identity.offset,
                  A.ls_hs_basis.kernels.offset, A.ls_hs_basis.representatives.offset,
                  A.ls_hs_basis.haskell_payload.offset),
        "group": (C.sizeof(A.ls_hs_permutation_group), A.ls_hs_permutation_group.masks.offset,
                  A.ls_hs_permutation_group.haskell_payload.offset),
        "terms": (C.sizeof(A.ls_hs_nonbranching_terms), A.ls_hs_nonbranching_terms.v.offset,
                  A.ls_hs_nonbranching_terms.s.offset),
        "op": (C.sizeof(A.ls_hs_operator), A.ls_hs_operator.off_diag_terms.offset,
               A.ls_hs_operator.haskell_payload.offset),
    }
    assert got == want
    # the engine's own C header gives the same layout under gcc
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lattice_symmetries_types.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(ls_hs_basis),
         offsetof(ls_hs_basis, state_index_is_identity), offsetof(ls_hs_basis, kernels),
         offsetof(ls_hs_basis, representatives), offsetof(ls_hs_basis, haskell_payload),
         sizeof(ls_hs_permutation_group), offsetof(ls_hs_permutation_group, masks),
         offsetof(ls_hs_permutation_group, haskell_payload), sizeof(ls_hs_nonbranching_terms),
         offsetof(ls_hs_nonbranching_terms, v), offsetof(ls_hs_nonbranching_terms, s),
         sizeof(ls_hs_operator), offsetof(ls_hs_operator, haskell_payload));
  return 0;
}"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [72, 24, 32, 40, 64, 56, 16, 48, 56, 8, 48, 40, 32]


def test_struct_layouts_match_the_reference_header_member_by_member():
    """oracle/layout_probe.c compiled against include/lattice_symmetries_types.h
    prints the same sizes and offsets of every member as the same probe
    compiled against the reference's kernels/lattice_symmetries_types.h
    (oracle/_ref/layout_ref, built by `make -C oracle` where the reference
    tree is present)."""
    ref = os.path.join(REPO, "oracle", "_ref", "layout_ref")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref/layout_ref not built (the reference tree is absent)")
    want = subprocess.run([ref], check=True, capture_output=True, text=True).stdout
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "layout_ours")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"),
                        os.path.join(REPO, "oracle", "layout_probe.c"), "-o", exe], check=True)
        got = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert got.splitlines() == want.splitlines()
    assert len(want.splitlines()) > 50


@pytest.mark.parametrize("n,gens,sectors", [
    (32, [models.chain_translation(32), models.chain_parity(32)], [0, 0]),
    (36, models.square_generators(6), [0, 0, 0, 0, 0]),
    (12, [models.chain_translation(12), models.chain_parity(12)], [6, 1]),
    (28, [models.chain_translation(28)], [3]),
    (5, [models.chain_translation(5)], [2]),
])
def test_group_create_matches_oracle_closure(n, gens, sectors):
    """ls_amd_permutation_group_create (host C++) vs the oracle: same
    ascending elements, bit-identical characters, and Benes networks that
    realise every permutation."""
    import lsamd
    S = lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])
    perms, num, den, re, im = O.group_closure(n, gens, sectors)
    assert np.array_equal(S.elements(), perms)
    ch = S.characters
    assert np.array_equal(ch.real, re) and np.array_equal(ch.imag, im)
    L = 0
    while (1 << L) < n:
        L += 1
    assert S.depth == (2 * L - 1 if L else 0)
    masks, shifts = S.masks, S.shifts
    rng = np.random.default_rng(7)
    xs = rng.integers(0, 1 << n, size=64, dtype=np.uint64)
    for e in range(len(S)):
        me = np.ascontiguousarray(masks[:, e])
        for x in xs:
            a = O.lib().ora_benes_apply(S.depth, me.ctypes.data_as(C.POINTER(C.c_uint64)), 1,
                                        shifts.ctypes.data_as(C.POINTER(C.c_uint64)), int(x))
            assert a == O.permute_bits(perms[e], int(x))


def test_group_create_rejects_incompatible_sectors():
    import lsamd
    with pytest.raises(ValueError):
        lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(6), 0),
                          lsamd.Symmetry(models.chain_translation(6), 1)])
    with pytest.raises(ValueError):
        lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(6), 6)])  # sector >= period


def test_spin_terms_match_dense_pauli_matrices():
    n = 4
    expr = (sigma_x(0) * sigma_x(1) * 0.7 + sigma_y(1) * sigma_y(2) * 1.3 + sigma_z(2) * sigma_z(3) * -0.4 +
            sigma_plus(3) * sigma_minus(0) * (0.2 + 0.5j) + sigma_x(0) * sigma_y(0) * 0.9 + sigma_y(2) * 0.3)
    want = H.pauli_dense(n, [(0.7, [("x", 0), ("x", 1)]), (1.3, [("y", 1), ("y", 2)]),
                             (-0.4, [("z", 2), ("z", 3)]), (0.2 + 0.5j, [("+", 3), ("-", 0)]),
                             (0.9, [("x", 0), ("y", 0)]), (0.3, [("y", 2)])])
    diag, off = expr.split()
    got = H.term_matrix(n, diag + off)
    assert np.allclose(got, want, atol=1e-14)


def test_heisenberg_bond_closed_form():
    # sigma.sigma = 2 (s+ s- + s- s+) + sz sz (Parser.hs:249-257)
    e1 = models.heisenberg_bond(0, 1)
    e2 = (sigma_plus(0) * sigma_minus(1) + sigma_minus(0) * sigma_plus(1)) * 2 + sigma_z(0) * sigma_z(1)
    assert e1.terms.keys() == e2.terms.keys()
    for k in e1.terms:
        assert abs(e1.terms[k] - e2.terms[k]) < 1e-15


def test_term_products_match_matrix_products():
    """Term.compose (NonbranchingTerm.hs:43-58) against products of the
    generator matrices, for every ordering of fermionic and spin operators."""
    n = 4
    gens = ([fermion_create(i) for i in range(n)] + [fermion_annihilate(i) for i in range(n)])
    mats = [H.term_matrix(n, sum(g.split(), [])) for g in gens]
    for ga, ma in zip(gens, mats):
        for gb, mb in zip(gens, mats):
            prod = ga * gb
            got = H.term_matrix(n, sum(prod.split(), []))
            assert np.allclose(got, ma @ mb), (ga.terms, gb.terms)
    spins = [sigma_plus(i) for i in range(n)] + [sigma_minus(i) for i in range(n)] + [sigma_z(i) for i in range(n)]
    smats = [H.term_matrix(n, sum(g.split(), [])) for g in spins]
    for ga, ma in zip(spins, smats):
        for gb, mb in zip(spins, smats):
            assert np.allclose(H.term_matrix(n, sum((ga * gb).split(), [])), ma @ mb)


def test_hubbard_hamiltonian_is_hermitian():
    expr = models.hubbard(4, models.chain_bonds(4), t=1.0, U=4.0)
    M = H.term_matrix(8, sum(expr.split(), []))
    assert np.allclose(M, M.conj().T)
    assert expr.is_hermitian(1e-15)


def test_fermion_terms_anticommute():
    n = 4
    I = H.term_matrix(n, [t for t in Expr.identity().split()[0]])
    for i in range(n):
        for j in range(n):
            ci = H.term_matrix(n, sum(fermion_annihilate(i).split(), []))
            cdj = H.term_matrix(n, sum(fermion_create(j).split(), []))
            anti = ci @ cdj + cdj @ ci
            assert np.allclose(anti, I if i == j else 0 * I)
            cj = H.term_matrix(n, sum(fermion_annihilate(j).split(), []))
            assert np.allclose(ci @ cj + cj @ ci, 0 * I)


def test_hermiticity_check():
    assert models.j1j2_square(4).is_hermitian(1e-15)
    assert not (sigma_plus(0) * sigma_minus(1)).is_hermitian(1e-15)
