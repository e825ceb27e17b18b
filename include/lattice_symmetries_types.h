/*
 * lattice_symmetries_types.h -- data types of the lattice-symmetries C ABI,
 * restated for the MI355X engine (liblsamd.so).
 *
 * The engine is a drop-in under the reference's Haskell and Python layers,
 * so every struct below has exactly the member order, types and padding of
 * the reference's kernels/lattice_symmetries_types.h (lines cited per
 * item).  Only the types the hot path touches are declared.
 *
 * Refcount members are `_Atomic int` in the reference's C build.  C++ has
 * no `_Atomic`; the engine declares them as plain `int` (same size and
 * alignment) and updates them with __atomic builtins.
 */
#ifndef LSAMD_LATTICE_SYMMETRIES_TYPES_H
#define LSAMD_LATTICE_SYMMETRIES_TYPES_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
#define LSAMD_ATOMIC_INT int
extern "C" {
#else
#define LSAMD_ATOMIC_INT _Atomic int
#endif

/* Chapel's external array: elements, count, and the function that frees the
   elements (reference: kernels/lattice_symmetries_types.h:27-32). */
typedef struct {
  void *elts;
  uint64_t num_elts;
  void *freer;
} chpl_external_array;

/* Complex double as two adjacent doubles, real part first (types.h:43-56:
   std::complex<double> / _Complex double / {_real, _imag} are one layout). */
typedef struct ls_hs_scalar {
  double _real;
  double _imag;
} ls_hs_scalar;

typedef uint64_t ls_hs_bits;

/* types.h:81-85 */
typedef enum ls_hs_particle_type {
  LS_HS_SPIN,
  LS_HS_SPINFUL_FERMION,
  LS_HS_SPINLESS_FERMION
} ls_hs_particle_type;

/* Kernel signatures of the per-basis kernel table (types.h:87-98). */
typedef void (*ls_hs_internal_state_index_kernel_type)(
    ptrdiff_t batch_size, uint64_t const *alphas, ptrdiff_t alphas_stride,
    ptrdiff_t *indices, ptrdiff_t indices_stride, void const *private_data);

typedef void (*ls_hs_internal_is_representative_kernel_type)(
    ptrdiff_t batch_size, uint64_t const *alphas, ptrdiff_t alphas_stride,
    uint8_t *are_representatives, double *norms, void const *private_data);

typedef void (*ls_hs_internal_state_info_kernel_type)(
    ptrdiff_t batch_size, uint64_t const *alphas, ptrdiff_t alphas_stride,
    uint64_t *betas, ptrdiff_t betas_stride, ls_hs_scalar *characters,
    double *norms, void const *private_data);

/* types.h:100-107 */
typedef struct ls_hs_basis_kernels {
  ls_hs_internal_state_info_kernel_type state_info_kernel;
  void *state_info_data;
  ls_hs_internal_is_representative_kernel_type is_representative_kernel;
  void *is_representative_data;
  ls_hs_internal_state_index_kernel_type state_index_kernel;
  void *state_index_data;
} ls_hs_basis_kernels;

/* Permutation group as Benes networks (types.h:109-119):
   masks is [number_shifts][number_masks] row-major (one row per network
   stage, one column per group element), shifts is [number_shifts],
   eigvals_re/_im are the characters of the group elements. */
typedef struct ls_hs_permutation_group {
  LSAMD_ATOMIC_INT refcount;
  int number_bits;
  int number_shifts;
  int number_masks;
  uint64_t *masks;
  uint64_t *shifts;
  double *eigvals_re;
  double *eigvals_im;
  void *haskell_payload;
} ls_hs_permutation_group;

/* types.h:121-133 */
typedef struct ls_hs_basis {
  LSAMD_ATOMIC_INT refcount;
  int number_sites;
  int number_particles;
  int number_up;
  ls_hs_particle_type particle_type;
  int spin_inversion;
  bool state_index_is_identity;
  bool requires_projection;
  ls_hs_basis_kernels *kernels;
  chpl_external_array representatives;
  void *haskell_payload;
} ls_hs_basis;

/* Non-branching terms (types.h:140-151): term t maps |a> to
   v[t] (-1)^popcount(a & s[t]) |a ^ x[t]> when (a & m[t]) == r[t].
   m, l, r, x, s are [number_terms][number_words] row-major. */
typedef struct ls_hs_nonbranching_terms {
  int number_terms;
  int number_bits;
  ls_hs_scalar const *v;
  uint64_t const *m;
  uint64_t const *l;
  uint64_t const *r;
  uint64_t const *x;
  uint64_t const *s;
} ls_hs_nonbranching_terms;

/* types.h:153-161 */
typedef struct ls_hs_operator {
  LSAMD_ATOMIC_INT refcount;
  ls_hs_basis const *basis;
  ls_hs_nonbranching_terms const *off_diag_terms;
  ls_hs_nonbranching_terms const *diag_terms;
  void *haskell_payload;
} ls_hs_operator;

/* Global engine table (types.h:170-180). */
typedef struct ls_chpl_kernels {
  void (*enumerate_states)(ls_hs_basis const *, uint64_t, uint64_t,
                           chpl_external_array *);
  void (*operator_apply_off_diag)(ls_hs_operator *, int64_t, uint64_t *,
                                  chpl_external_array *, chpl_external_array *,
                                  chpl_external_array *, int64_t);
  void (*operator_apply_diag)(ls_hs_operator *, int64_t, uint64_t *,
                              chpl_external_array *, int64_t);
  void (*matrix_vector_product)(ls_hs_operator *, int, double const *,
                                double *);
} ls_chpl_kernels;

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* LSAMD_LATTICE_SYMMETRIES_TYPES_H */
