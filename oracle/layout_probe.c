/*
 * Test infrastructure: prints the sizes and member offsets of the C-ABI
 * structs of lattice_symmetries_types.h.  Compiled once against the
 * reference's kernels/lattice_symmetries_types.h (oracle/Makefile, target
 * ref, into _ref/layout_ref) and once against this repository's
 * include/lattice_symmetries_types.h (by the test), so the two layouts are
 * compared member by member on the same compiler.
 */
#include <stddef.h>
#include <stdio.h>

#include "lattice_symmetries_types.h"

#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
#define S(T) printf(#T " %zu\n", sizeof(T))

int main(void) {
  S(chpl_external_array);
  F(chpl_external_array, elts);
  F(chpl_external_array, num_elts);
  F(chpl_external_array, freer);
  S(ls_hs_scalar);
  S(ls_hs_basis_kernels);
  F(ls_hs_basis_kernels, state_info_kernel);
  F(ls_hs_basis_kernels, state_info_data);
  F(ls_hs_basis_kernels, is_representative_kernel);
  F(ls_hs_basis_kernels, is_representative_data);
  F(ls_hs_basis_kernels, state_index_kernel);
  F(ls_hs_basis_kernels, state_index_data);
  S(ls_hs_permutation_group);
  F(ls_hs_permutation_group, refcount);
  F(ls_hs_permutation_group, number_bits);
  F(ls_hs_permutation_group, number_shifts);
  F(ls_hs_permutation_group, number_masks);
  F(ls_hs_permutation_group, masks);
  F(ls_hs_permutation_group, shifts);
  F(ls_hs_permutation_group, eigvals_re);
  F(ls_hs_permutation_group, eigvals_im);
  F(ls_hs_permutation_group, haskell_payload);
  S(ls_hs_basis);
  F(ls_hs_basis, refcount);
  F(ls_hs_basis, number_sites);
  F(ls_hs_basis, number_particles);
  F(ls_hs_basis, number_up);
  F(ls_hs_basis, particle_type);
  F(ls_hs_basis, spin_inversion);
  F(ls_hs_basis, state_index_is_identity);
  F(ls_hs_basis, requires_projection);
  F(ls_hs_basis, kernels);
  F(ls_hs_basis, representatives);
  F(ls_hs_basis, haskell_payload);
  S(ls_hs_nonbranching_terms);
  F(ls_hs_nonbranching_terms, number_terms);
  F(ls_hs_nonbranching_terms, number_bits);
  F(ls_hs_nonbranching_terms, v);
  F(ls_hs_nonbranching_terms, m);
  F(ls_hs_nonbranching_terms, l);
  F(ls_hs_nonbranching_terms, r);
  F(ls_hs_nonbranching_terms, x);
  F(ls_hs_nonbranching_terms, s);
  S(ls_hs_operator);
  F(ls_hs_operator, refcount);
  F(ls_hs_operator, basis);
  F(ls_hs_operator, off_diag_terms);
  F(ls_hs_operator, diag_terms);
  F(ls_hs_operator, haskell_payload);
  S(ls_chpl_kernels);
  F(ls_chpl_kernels, enumerate_states);
  F(ls_chpl_kernels, operator_apply_off_diag);
  F(ls_chpl_kernels, operator_apply_diag);
  F(ls_chpl_kernels, matrix_vector_product);
  printf("LS_HS_SPINLESS_FERMION %d\n", (int)LS_HS_SPINLESS_FERMION);
  return 0;
}
