/*
 * lsamd.h -- C ABI of liblsamd.so, the MI355X (gfx950) engine for the
 * lattice-symmetries hot path: symmetry-adapted basis construction and the
 * sparse Hamiltonian product y = H x.
 *
 * Part 1 declares the reference's own entry points with the reference's
 * names and signatures.  liblsamd.so exports all of them, so it replaces the
 * reference's libkernels (kernels/reference.c, kernels/indexing.c,
 * kernels/kernels.c) and the four table functions of the Chapel library
 * (chapel/src/{StatesEnumeration,BatchedOperator,DistributedMatrixVector}.chpl)
 * without any change to the callers.  Arguments are
 * host pointers, as in the reference; the work runs on the GPU.
 *
 * Part 2 declares additive entry points (prefix ls_amd_) that keep data
 * resident in device memory: basis build, index, CSR cache, FP64 / FP32 /
 * block products and the Lanczos driver.  They have no reference
 * counterpart; each says which reference routine it generalises.
 *
 * Errors follow the reference: invalid arguments go to ls_hs_error (a
 * settable handler, abort by default); after the handler returns the entry
 * point returns without touching its outputs further.
 */
#ifndef LSAMD_H
#define LSAMD_H

#include "lattice_symmetries_types.h"

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================== */
/* Part 1: reference entry points                                            */
/* ======================================================================== */

/* --- errors, refcounts, external arrays: kernels/reference.c ------------- */

/* reference.c:11-15 -- print and abort. */
__attribute__((noreturn)) void ls_hs_fatal_error(char const *func, int line,
                                                 char const *message);
/* reference.c:25-30 -- NULL restores the default (print and abort). */
void ls_hs_set_exception_handler(void (*handler)(char const *message));
/* reference.c:32-36 */
void ls_hs_error(char const *message);
/* reference.c:40-48 */
void ls_hs_internal_destroy_external_array(chpl_external_array *arr);
/* reference.c:50-64 */
int ls_hs_internal_read_refcount(int const *refcount);
void ls_hs_internal_write_refcount(int *refcount, int value);
int ls_hs_internal_inc_refcount(int *refcount);
int ls_hs_internal_dec_refcount(int *refcount);

/* --- term application: kernels/reference.c:67-134 ------------------------ */

/* reference.c:67-95: ys[i] = sum over diagonal terms of Re(v) (-1)^.. xs[i]
   (xs may be NULL: factor 1). */
void ls_internal_operator_apply_diag_x1(ls_hs_operator const *op,
                                        ptrdiff_t batch_size,
                                        uint64_t const *alphas, double *ys,
                                        double const *xs);
/* reference.c:97-134: for each alpha, every matching off-diagonal term gives
   one (beta, coeff); offsets[batch_size + 1] delimits the rows. */
void ls_internal_operator_apply_off_diag_x1(ls_hs_operator const *op,
                                            ptrdiff_t batch_size,
                                            uint64_t const *alphas,
                                            uint64_t *betas,
                                            ls_hs_scalar *coeffs,
                                            ptrdiff_t *offsets,
                                            double const *xs);

/* --- basis dispatchers: kernels/reference.c:137-211 ---------------------- */

/* reference.c:137-148 */
void ls_hs_state_index(ls_hs_basis const *basis, ptrdiff_t batch_size,
                       uint64_t const *spins, ptrdiff_t spins_stride,
                       ptrdiff_t *indices, ptrdiff_t indices_stride);
/* reference.c:150-159 */
void ls_hs_is_representative(ls_hs_basis const *basis, ptrdiff_t batch_size,
                             uint64_t const *alphas, ptrdiff_t alphas_stride,
                             uint8_t *are_representatives, double *norms);
/* reference.c:161-171 */
void ls_hs_state_info(ls_hs_basis const *basis, ptrdiff_t batch_size,
                      uint64_t const *alphas, ptrdiff_t alphas_stride,
                      uint64_t *betas, ptrdiff_t betas_stride,
                      ls_hs_scalar *characters, double *norms);
/* reference.c:173-194: enumerate through the registered engine table, then
   build the state index with 22 prefix bits. */
void ls_hs_build_representatives(ls_hs_basis *basis, uint64_t lower,
                                 uint64_t upper);
/* reference.c:196-211 */
void ls_hs_unchecked_set_representatives(ls_hs_basis *basis,
                                         chpl_external_array const *states,
                                         int cache_bits);
/* reference.c:214-225 */
ls_chpl_kernels const *ls_hs_internal_get_chpl_kernels(void);
void ls_hs_internal_set_chpl_kernels(ls_chpl_kernels const *kernels);

/* --- state index: kernels/indexing.c ------------------------------------- */

typedef struct ls_hs_state_index_binary_search_data
    ls_hs_state_index_binary_search_data;

/* indexing.c:94-117.  The representatives are borrowed (host memory, sorted
   ascending); the engine also keeps a device copy plus a device prefix table
   until the data are destroyed. */
ls_hs_state_index_binary_search_data *
ls_hs_create_state_index_binary_search_kernel_data(
    chpl_external_array const *representatives, int number_bits,
    int prefix_bits);
/* indexing.c:119-125 */
void ls_hs_destroy_state_index_binary_search_kernel_data(
    ls_hs_state_index_binary_search_data *cache);
/* indexing.c:273-325: index of each spin, -1 when absent; strides must be 1. */
void ls_hs_state_index_binary_search_kernel(ptrdiff_t batch_size,
                                            uint64_t const *spins,
                                            ptrdiff_t spins_stride,
                                            ptrdiff_t *indices,
                                            ptrdiff_t indices_stride,
                                            void const *private_kernel_data);

/* --- orbit kernels: kernels/kernels.c + kernels/generator.cpp ------------ */

typedef struct ls_internal_halide_kernel_data ls_internal_halide_kernel_data;

/* kernels.c:112-189.  Copies the group (masks, shifts, characters) to the
   device; the group may be destroyed afterwards.  spin_inversion must be
   -1, 0 or 1. */
ls_internal_halide_kernel_data *
ls_internal_create_halide_kernel_data(ls_hs_permutation_group const *g,
                                      int spin_inversion);
/* kernels.c:191-193 */
void ls_internal_destroy_halide_kernel_data(ls_internal_halide_kernel_data *p);
/* kernels.c:195-244 + generator.cpp:168-272 */
void ls_hs_is_representative_halide_kernel(ptrdiff_t batch_size,
                                           uint64_t const *alphas,
                                           ptrdiff_t alphas_stride,
                                           uint8_t *are_representatives,
                                           double *norms,
                                           void const *private_data);
/* kernels.c:246-319 + generator.cpp:10-166 */
void ls_hs_state_info_halide_kernel(ptrdiff_t batch_size,
                                    uint64_t const *alphas,
                                    ptrdiff_t alphas_stride, uint64_t *betas,
                                    ptrdiff_t betas_stride,
                                    ls_hs_scalar *characters, double *norms,
                                    void const *private_data);

/* --- engine table functions: chapel/src/ --------------------------------- */

/* chapel/src/library.c:19-34 and LatticeSymmetries.chpl:17-32: initialise
   the device and register the four functions below in the global table. */
void ls_chpl_init(void);
/* Frees the engine's cached operators (those built for
   ls_chpl_matrix_vector_product) and the scratch kept between basis builds. */
void ls_chpl_finalize(void);
/* StatesEnumeration.chpl:692-709.  Representatives of `p` in [lower, upper],
   ascending; `dest` receives a malloc'ed array and its freer. */
void ls_chpl_enumerate_representatives(ls_hs_basis const *p, uint64_t lower,
                                       uint64_t upper,
                                       chpl_external_array *dest);
/* BatchedOperator.chpl:298-316 */
void ls_chpl_operator_apply_diag(ls_hs_operator *matrixPtr, int64_t count,
                                 uint64_t *alphas, chpl_external_array *coeffs,
                                 int64_t numTasks);
/* BatchedOperator.chpl:318-357 */
void ls_chpl_operator_apply_off_diag(ls_hs_operator *matrixPtr, int64_t count,
                                     uint64_t *alphas,
                                     chpl_external_array *betas,
                                     chpl_external_array *coeffs,
                                     chpl_external_array *offsets,
                                     int64_t numTasks);
/* DistributedMatrixVector.chpl:1090-1105: y = H x for one real vector over
   the representatives of op->basis (numVectors must be 1).  The diagonal
   assigns y; without diagonal terms y is accumulated into (:1054-1057). */
void ls_chpl_matrix_vector_product(ls_hs_operator *matrixPtr, int numVectors,
                                   double const *xPtr, double *yPtr);
/* The product above keeps a device copy of each operator it serves (the
   CSR cache).  An entry is rebuilt when the basis, its representatives,
   the engine's group / index data behind it or the terms change; it is
   dropped when that group or index data is destroyed, and the entries
   together hold at most LSAMD_CACHE_BYTES (default: half the device
   memory) of device memory, least recently used first out.  No counterpart
   in the reference, which recomputes every product from scratch:
   ls_amd_forget_operator drops the entry of an operator its caller is
   about to destroy; ls_amd_operator_cache_entries counts the entries. */
void ls_amd_forget_operator(ls_hs_operator const *op);
int64_t ls_amd_operator_cache_entries(void);

/* ======================================================================== */
/* Part 2: additive entry points (device-resident data)                      */
/* ======================================================================== */

/* Library status.  Returns 0 when a gfx950 device is usable, else nonzero. */
int ls_amd_device_available(void);
/* Last error message set by the engine (empty string when none). */
char const *ls_amd_last_error(void);
void ls_amd_clear_error(void);

/* Device memory for callers without their own allocator. */
void *ls_amd_device_malloc(size_t bytes);
void ls_amd_device_free(void *ptr);
int ls_amd_memcpy_h2d(void *dst, void const *src, size_t bytes);
int ls_amd_memcpy_d2h(void *dst, void const *src, size_t bytes);
int ls_amd_synchronize(void);

/* Group set-up (Group.hs:106-132, Benes.hs:316-377 restated in C++): closure
   of `number_generators` permutations of n sites (bit i of g(x) = bit
   p[i] of x) with momentum sectors, as the reference's
   ls_hs_permutation_group (ascending group order, characters
   cos/sin(-2 pi phase), Benes networks of depth 2 log2(next pow2 >= n) - 1).
   Returns NULL (and sets the last error) for inconsistent input. */
ls_hs_permutation_group *ls_amd_permutation_group_create(
    int number_sites, int number_generators, int32_t const *permutations,
    int32_t const *sectors);
void ls_amd_permutation_group_destroy(ls_hs_permutation_group *g);
/* Group elements as permutations, [number_masks][number_sites]. */
int ls_amd_permutation_group_elements(ls_hs_permutation_group const *g,
                                      int32_t *out);

/* Basis handle: a spin basis (or a basis without permutations) whose
   representatives live in device memory.  hamming_weight = -1 for none.
   group may be NULL (no permutations); spin_inversion in {-1, 0, 1}. */
typedef struct ls_amd_basis ls_amd_basis;
ls_amd_basis *ls_amd_basis_create(int number_sites, int hamming_weight,
                                  int spin_inversion,
                                  ls_hs_permutation_group const *group);
/* Fermion basis with permutation symmetries (no reference counterpart: the
   reference has no fermionic symmetries, Basis.hs:941-948).  Spinful when
   number_down >= 0: bits 0..n-1 are the up species (number_up particles)
   and bits n..2n-1 the down species (Basis.hs:621-628); spinless when
   number_down < 0 (number_up particles, or -1 for any number).  The group
   acts on all bits (the same site permutation on both species); group
   elements carry the sign of the relabelled creation operators in the
   Jordan-Wigner order of the term sign masks. */
ls_amd_basis *ls_amd_fermion_basis_create(int number_sites, int number_up,
                                          int number_down,
                                          ls_hs_permutation_group const *group);
void ls_amd_basis_destroy(ls_amd_basis *b);
/* Opt-in tolerance for numerically zero norms (no reference counterpart).
   With eps = 0 (the default) the reference's rule holds: a state is a
   representative when its stabiliser sum of characters is > 0, and the
   norm is sqrt(sum / |G|) even when the sum is a tiny negative (NaN).  In
   sectors with characters other than +-1 such sums are +-1e-16 for states
   whose exact norm is zero; with eps > 0 sums <= eps count as zero, which
   keeps those states out of the basis and gives their images norm 0.
   Call before ls_amd_basis_build. */
int ls_amd_basis_set_norm_tolerance(ls_amd_basis *b, double eps);
/* Basis.hs:734-774 (min/max state estimates). */
uint64_t ls_amd_basis_min_state(ls_amd_basis const *b);
uint64_t ls_amd_basis_max_state(ls_amd_basis const *b);
/* GPU enumeration of representatives in [lower, upper] (ascending, as
   StatesEnumeration.chpl:242-326); builds the device index as well.
   Returns the number of states or -1. */
int64_t ls_amd_basis_build(ls_amd_basis *b, uint64_t lower, uint64_t upper);
int64_t ls_amd_basis_number_states(ls_amd_basis const *b);
/* Number of candidate states examined by the last build. */
int64_t ls_amd_basis_number_candidates(ls_amd_basis const *b);
uint64_t const *ls_amd_basis_states_device(ls_amd_basis const *b);
int ls_amd_basis_copy_states(ls_amd_basis const *b, uint64_t *host_out);
/* Adopt a host list of sorted representatives instead of building. */
int ls_amd_basis_set_states(ls_amd_basis *b, uint64_t const *host_states,
                            int64_t count);
/* Device versions of state_info / is_representative / index on device
   arrays (strides 1). */
int ls_amd_basis_state_info_device(ls_amd_basis const *b, int64_t count,
                                   uint64_t const *d_alphas, uint64_t *d_betas,
                                   double *d_characters, double *d_norms);
int ls_amd_basis_is_representative_device(ls_amd_basis const *b, int64_t count,
                                          uint64_t const *d_alphas,
                                          uint8_t *d_flags, double *d_norms);
int ls_amd_basis_index_device(ls_amd_basis const *b, int64_t count,
                              uint64_t const *d_states, int64_t *d_indices);

/* state_info without device allocations.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  Same
   arguments and results as ls_hs_state_info_halide_kernel (host arrays,
   strides in elements, private_data from
   ls_internal_create_halide_kernel_data), but the device staging lives in a
   fixed 2.5 MiB array of the library, so a call takes at most 65536 states
   (larger batches return -1 with ls_amd_last_error set).  Calls in one
   process share that array and are serialised.  Returns 0 or -1. */
int ls_amd_state_info_noalloc(ptrdiff_t batch_size, uint64_t const *alphas,
                              ptrdiff_t alphas_stride, uint64_t *betas,
                              ptrdiff_t betas_stride, ls_hs_scalar *characters,
                              double *norms, void const *private_data);

/* Operator handle over an ls_amd_basis: diagonal and off-diagonal
   non-branching terms (one 64-bit word per mask, as the reference's
   matvec requires, BatchedOperator.chpl:327-328). */
typedef struct ls_amd_operator ls_amd_operator;
ls_amd_operator *ls_amd_operator_create(ls_amd_basis const *b,
                                        ls_hs_nonbranching_terms const *diag,
                                        ls_hs_nonbranching_terms const *off_diag);
void ls_amd_operator_destroy(ls_amd_operator *op);

/* On-the-fly product, FP64, real x / y on the device: the reference's
   ls_chpl_matrix_vector_product with device pointers, in its push form
   (each alpha_i adds Re(c) into y at index(rep beta),
   DistributedMatrixVector.chpl:91-143), for any operator, Hermitian or
   not.  accumulate = 1 keeps y as the base of the sum when there are no
   diagonal terms. */
int ls_amd_matvec_f64(ls_amd_operator *op, double const *d_x, double *d_y,
                      int accumulate);
/* Complex x / y (interleaved re, im), y = M x with the full complex matrix
   (the reference casts coefficients to real; this keeps them complex). */
int ls_amd_matvec_c128(ls_amd_operator *op, double const *d_x, double *d_y);

/* CSR cache: off-diagonal part as rows of (column, value) plus a dense
   diagonal; row j holds every (i, C) that alpha_i emits with target
   index(rep beta) = j, in ascending i, so the CSR products compute the same
   push-form y = M x as the reference for any operator.  precision: 64 stores FP64 values, 32 stores FP32 values, 96
   stores both; 128 stores complex values M[i, j] as (re, im) pairs in FP64
   and FP32, for sector batches (the real-valued CSR products reject it).
   Returns the number of stored off-diagonal nonzeros or -1. */
int64_t ls_amd_operator_build_csr(ls_amd_operator *op, int precision);
/* Value table (default on): a real CSR with at most 256 distinct stored
   values and fewer than 2^24 rows keeps each entry as one 32-bit word,
   column | value index << 24, and the values in a table; one with at most
   4096 distinct values (any number of rows) keeps the 32-bit column and a
   16-bit value index (the wide table).  Set before
   ls_amd_operator_build_csr; 0 keeps (column, value) pairs. */
int ls_amd_operator_set_csr_packing(ls_amd_operator *op, int enable);
/* Orbit search of the fused kernels (CSR build and on-the-fly product).
   When the group is the set of translations of an lx x ly torus (site
   r * lx + c) times the elements fixing site 0, and every character is real
   +-1, each point element's image is translated by word rotations instead
   of running one network per element: the same representatives, norms and
   coefficients (the stored CSR is bit-identical).  On by default;
   ls_amd_operator_set_factored(op, 0) keeps the element-by-element search,
   ls_amd_operator_uses_factored reports which one runs (1 / 0). */
int ls_amd_operator_set_factored(ls_amd_operator *op, int enable);
int ls_amd_operator_uses_factored(ls_amd_operator const *op);
/* out[4]: stored off-diagonal nonzeros, value table (0 none, 1 packed into
   the column words, 2 wide), value-table size, complex (0/1). */
int ls_amd_operator_csr_info(ls_amd_operator const *op, int64_t *out);
int64_t ls_amd_operator_csr_nnz(ls_amd_operator const *op);
/* Stored CSR (off-diagonal) rows: row_ptr[n+1] (int64), cols, values
   (2 nnz interleaved re, im for a complex CSR). */
int ls_amd_operator_copy_csr(ls_amd_operator const *op, int64_t *row_ptr,
                             uint32_t *cols, double *values, double *diag);
/* y = H x from the cached CSR: FP64 one vector. */
int ls_amd_csr_matvec_f64(ls_amd_operator *op, double const *d_x, double *d_y);
/* FP32 block of k right-hand sides, X and Y row-major [n][k] (k <= 16). */
int ls_amd_csr_matvec_f32(ls_amd_operator *op, int k, float const *d_x,
                          float *d_y);
/* FP64 block of k right-hand sides, row-major [n][k] (k <= 16). */
int ls_amd_csr_matvec_f64_block(ls_amd_operator *op, int k, double const *d_x,
                                double *d_y);
/* Choose the CSR kernel schedules by timing the candidates on the current
   matrix; returns the chosen FP64 lanes per row. */
int ls_amd_operator_autotune(ls_amd_operator *op);
/* Whether the product of a mode (0 FP64 CSR, 1 FP32 CSR, 2 on-the-fly,
   3 on-the-fly complex) satisfies <u, M v> = <M u, v> on seeded vectors:
   1 yes (symmetric Re(M) for real modes, Hermitian M for mode 3), 0 no,
   -1 on error.  The Lanczos solvers refuse operators that fail it. */
int ls_amd_operator_is_hermitian(ls_amd_operator *op, int mode);
/* Current CSR product schedule: out[5] = lanes per row (FP64), flags (FP64),
   lanes per row (FP32 block), flags (FP32 block), persistent grid size.
   The FP64 pair serves every product that gathers at most 16 bytes of x per
   stored entry (the FP64 vector, FP64 blocks of 1-2, FP32 blocks of 1-4),
   the FP32 pair the wider blocks (FP32 of 8-16, FP64 of 4-16); the
   autotuner times the first on the FP64 vector, the second on the FP32
   block of 8.
   flags: bit 0 non-temporal loads of the CSR stream, bit 1 die-aware
   mapping (each accelerator die's workgroups walk one contiguous range of
   rows) on the persistent grid, bit 2 (with bit 1) the die-aware mapping
   over one workgroup per tile instead.  Unless set (ls_amd_operator_set_schedule)
   or autotuned, the CSR build picks lanes so that each lane sums about four
   entries of a mean row, and flags 6 from a million rows on (else 0). */
int ls_amd_operator_schedule(ls_amd_operator const *op, int *out);
/* Process-wide: whether basis builds may use the translation-factorised
   filter for groups of torus translations times point elements with
   characters +-1 (1, the default) or always test element by element (0).
   Both give the same representatives. */
int ls_amd_set_factored_enumeration(int enable);
int ls_amd_operator_set_schedule(ls_amd_operator *op, int lanes_f64,
                                 int flags_f64, int lanes_f32, int flags_f32);

/* Eigensolvers (the consumers of the product; the reference drives its
   product from scipy eigsh, python/lattice_symmetries/__init__.py:591-752,
   or PRIMME, chapel/src/Diagonalize.chpl:134-162).  FP64 Krylov vectors,
   orthogonalisation and Ritz solve; the product in the requested form. */
typedef struct ls_amd_lanczos_result {
  int steps;              /* Lanczos steps (blocks for block Lanczos) */
  int refine_steps;       /* FP64 correction steps after the main run */
  int number_eigenvalues; /* valid entries of eigenvalues */
  double eigenvalues[8];  /* lowest Ritz values, ascending (the first one
                             after FP64 correction when requested) */
  double rayleigh;        /* FP64 Rayleigh quotient of the lowest Ritz vector */
  double residual_norm;   /* FP64 || H u - rayleigh u || for that vector */
  double matvec_seconds;  /* wall time spent in products */
  int64_t matvec_calls;
  double total_seconds;
} ls_amd_lanczos_result;

/* Single-vector Lanczos with full re-orthogonalisation.  mode: 0 FP64 CSR,
   1 FP32 CSR (FP32 vectors in the product only), 2 on-the-fly FP64, 3
   on-the-fly FP64 with complex vectors (Krylov vectors are the real
   embedding, 2 n doubles, interleaved re, im; a single start vector sees
   each eigenvalue of the complex matrix once).
   Seeded N(0,1) start vector.  Stops when the k lowest Ritz pairs have
   relative residual bound < tol or after max_steps.  refine_steps > 0 runs
   that many FP64 Lanczos steps restarted from the lowest Ritz vector.
   d_ritz_out (device, may be NULL) receives the lowest Ritz vector. */
int ls_amd_lanczos(ls_amd_operator *op, int mode, int max_steps, int k,
                   double tol, uint64_t seed, int refine_steps,
                   double *d_ritz_out, ls_amd_lanczos_result *out);
/* Block Lanczos with block_size (1, 2, 4, 8) right-hand sides per product,
   through ls_amd_block_eigensolve: max_blocks Krylov blocks per restart
   cycle, the k lowest eigenvalues (k <= block_size).  mode: 0 FP64 block
   CSR, 1 FP32 block CSR until the FP32 residual floor, then FP64 (needs the
   CSR in both precisions).  refine_steps > 0 adds an FP64 Lanczos run
   restarted from the lowest Ritz vector. */
int ls_amd_block_lanczos(ls_amd_operator *op, int mode, int block_size,
                         int max_blocks, int k, double tol, uint64_t seed,
                         int refine_steps, double *d_ritz_out,
                         ls_amd_lanczos_result *out);

typedef struct ls_amd_eig_info {
  int cycles;            /* restart cycles run */
  int fp64_cycles;       /* of which with the FP64 product */
  int blocks;            /* block products in the Krylov cycles */
  int converged;         /* 1 if every segment met tol: the residual bounds
                            fell below tol and the explicit final residuals
                            are within 10 tol */
  int64_t products;      /* all block products, the final check included */
  double product_seconds;
  double total_seconds;
  double max_residual;   /* largest FP64 relative residual reported */
} ls_amd_eig_info;

/* Thick-restart block Lanczos on one operator (Krylov blocks stored in the
   product's precision, FP64 orthogonalisation arithmetic: full
   re-orthogonalisation, a second pass when some direction of the block
   nearly vanishes, CholQR2, fresh random directions for exhausted ones).
   block_size in {1, 2, 4, 8}; nev <= block_size eigenvalues; max_blocks
   Krylov blocks per cycle; keep_blocks Ritz blocks kept at a restart.
   mode 1: FP32 block product until the relative residual falls below
   switch_tol (or stalls), then FP64 until tol; mode 0: FP64 throughout.
   evals[nev] ascending; residuals[nev] (may be NULL) are FP64 relative
   residual norms || H u - lambda u || / max(1, |lambda|); d_ritz_out (device,
   may be NULL) receives the lowest Ritz vector. */
int ls_amd_block_eigensolve(ls_amd_operator *op, int mode, int block_size,
                            int nev, int max_blocks, int keep_blocks,
                            int max_cycles, double tol, double switch_tol,
                            uint64_t seed, double *evals, double *residuals,
                            double *d_ritz_out, ls_amd_eig_info *info);

/* The same solve, returning all nev Ritz vectors: d_evecs (device) receives
   nev rows of number_states values, row e the eigenvector of evals[e] (unit
   norm), the [number of eigenvalues][dimension] layout of the eigenvectors
   the reference's Diagonalize.chpl:246-255 writes. */
int ls_amd_block_eigenvectors(ls_amd_operator *op, int mode, int block_size,
                              int nev, int max_blocks, int keep_blocks,
                              int max_cycles, double tol, double switch_tol,
                              uint64_t seed, double *evals, double *residuals,
                              double *d_evecs, ls_amd_eig_info *info);

/* Host utility: eigen-decomposition of a dense symmetric n x n matrix A
   (row-major) by Householder tridiagonalisation and implicit QR with
   Wilkinson shifts, as the Krylov solvers use for their projected
   matrices.  w[n] ascending; Z (may be NULL) the eigenvectors as columns. */
int ls_amd_dense_eigh(int n, double const *A, double *w, double *Z);

/* Host utility: the same for a dense complex Hermitian n x n matrix A
   (row-major, interleaved re, im), by complex Householder tridiagonalisation
   and the same implicit QR, as the block solver uses for the projected
   matrices of complex sectors.
   w[n] ascending; Z (may be NULL) the eigenvectors as columns, interleaved. */
int ls_amd_dense_eigh_hermitian(int n, double const *A, double *w, double *Z);

/* ---- Many symmetry sectors in one product --------------------------------
   No counterpart in the reference, which builds one basis per sector and
   diagonalises the sectors one after another (chapel/src/Diagonalize.chpl:
   134-162).  A sector batch concatenates the complex CSR caches of several
   operators (each built with ls_amd_operator_build_csr(op, 128)) into one
   block-diagonal matrix.  Vectors hold k complex right-hand sides per state,
   laid out [state][re | im][k] over all sectors in order (sector s starts at
   state offsets[s]).  Read as [2 state + part][k] this is the real
   embedding [[A, -B], [B, A]] of each sector's M = A + iB.  The product is
   one persistent launch over every sector, captured in a HIP graph on first
   use of a given (x, y, k) and replayed afterwards. */
typedef struct ls_amd_sector_batch ls_amd_sector_batch;
ls_amd_sector_batch *ls_amd_sector_batch_create(ls_amd_operator *const *ops,
                                                int count);
void ls_amd_sector_batch_destroy(ls_amd_sector_batch *batch);
int ls_amd_sector_batch_count(ls_amd_sector_batch const *batch);
int64_t ls_amd_sector_batch_number_states(ls_amd_sector_batch const *batch);
int64_t ls_amd_sector_batch_nnz(ls_amd_sector_batch const *batch);
/* offsets[count + 1] (host): first state of every sector. */
int ls_amd_sector_batch_offsets(ls_amd_sector_batch const *batch,
                                int64_t *offsets);
/* Product schedule: lanes per row (1, 2, 4, 8, 16), workgroups in the grid
   (8 .. 2^26; workgroups loop over row tiles, so a grid of at least one
   workgroup per tile of 256 / lanes rows gives each workgroup one tile), and
   whether each die's workgroups own one contiguous range of rows
   (die_aware = 1, grid a multiple of 8).  One schedule serves every k. */
int ls_amd_sector_batch_set_schedule(ls_amd_sector_batch *batch, int lanes,
                                     int blocks, int die_aware);
int ls_amd_sector_batch_get_schedule(ls_amd_sector_batch const *batch,
                                     int *lanes, int *blocks, int *die_aware);
/* Times every schedule (lanes x grid of 2, 4, 8, 16 workgroups per compute
   unit or one per tile x die-aware) for the FP32 product with k columns and
   keeps the fastest. */
int ls_amd_sector_batch_autotune(ls_amd_sector_batch *batch, int k);
/* y = M x on every sector, k in {1, 2, 4, 8} complex columns. */
int ls_amd_sector_batch_matvec_f32(ls_amd_sector_batch *batch, int k,
                                   float const *d_x, float *d_y);
int ls_amd_sector_batch_matvec_f64(ls_amd_sector_batch *batch, int k,
                                   double const *d_x, double *d_y);

/* Thick-restart block Lanczos (as ls_amd_block_eigensolve) on every sector
   at once, one product launch per block step, in complex arithmetic with
   blocks of block_size (1, 2, 4) complex columns; nev eigenvalues per
   sector (nev <= block_size).  max_blocks Krylov blocks per cycle, the
   lowest keep_blocks Ritz blocks kept at a restart.  The product is FP32,
   on Krylov blocks stored as FP32 with FP64 orthogonalisation sums, until
   the relative residual falls below switch_tol or stalls, then FP64 on
   FP64 blocks until tol.  Sectors with fewer than 2 (max_blocks + 1)
   block_size states (up to 512) are diagonalised densely instead.
   evals[count * nev] receives each sector's lowest eigenvalues ascending
   (NaN past a sector's dimension); residuals (may be NULL) their FP64
   relative residual norms || M u - lambda u || / max(1, |lambda|), computed
   with the FP64 product at the end. */
int ls_amd_sector_batch_eigensolve(ls_amd_sector_batch *batch, int block_size,
                                   int nev, int max_blocks, int keep_blocks,
                                   int max_cycles, double tol,
                                   double switch_tol, uint64_t seed,
                                   double *evals, double *residuals,
                                   ls_amd_eig_info *info);

/* ---- Model files --------------------------------------------------------
   The reference's YAML model format (chapel/data/<model>.yaml: basis +
   hamiltonian + observables, Yaml.hs:14-48; expression grammar of
   Parser.hs:60-170 with "sites" replication, Expr.hs:266-286), read without
   Haskell.  text is the file's content (UTF-8). */
typedef struct ls_amd_model ls_amd_model;
ls_amd_model *ls_amd_model_from_yaml(char const *text);
void ls_amd_model_destroy(ls_amd_model *model);
/* info[10]: particle (0 spin, 1 spinless fermion, 2 spinful fermion),
   number of sites, Hamming weight (-1 none), spin inversion (0, +1, -1),
   number up (or spinless particles; -1 none), number down (-1 none), total
   spinful particles (-1 none), number of symmetries, has hamiltonian,
   number of observables. */
int ls_amd_model_basis_info(ls_amd_model const *model, int *info);
/* Symmetry `index`: permutation[number_sites] and sector. */
int ls_amd_model_symmetry(ls_amd_model const *model, int index,
                          int32_t *permutation, int32_t *sector);
/* Terms of an operator: which = -1 the hamiltonian, k >= 0 observable k.
   diagonal selects the terms with x = 0.  Terms are sorted by
   (m, l, r, x, s); v holds (re, im) pairs. */
int64_t ls_amd_model_number_terms(ls_amd_model const *model, int which,
                                  int diagonal);
int ls_amd_model_terms(ls_amd_model const *model, int which, int diagonal,
                       double *v, uint64_t *m, uint64_t *l, uint64_t *r,
                       uint64_t *x, uint64_t *s);
/* Device basis of the model (not yet built) and an operator on it. */
ls_amd_basis *ls_amd_model_create_basis(ls_amd_model const *model);
ls_amd_operator *ls_amd_model_create_operator(ls_amd_model const *model,
                                              int which,
                                              ls_amd_basis const *basis);

/* Timing of the last launched kernel of the given family, in milliseconds
   (HIP events on the engine's stream). family: 0 csr f64, 1 csr f32 block,
   2 on-the-fly f64, 3 enumeration, 4 csr build, 5 csr f64 block, 6 sector
   batch product (graph replay). */
double ls_amd_last_kernel_ms(int family);

#ifdef __cplusplus
} /* extern "C" */
#endif

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#endif /* LSAMD_H */
