// operator_internal.hpp -- operator objects shared by operator.hip and abi.cpp.
#pragma once

#include <complex>
#include <functional>

#include "lsamd_internal.hpp"
#include "spmv_kernels.hpp"

namespace lsamd {

// Host + device copy of one set of non-branching terms (single word).
struct Terms {
  int count = 0;
  std::vector<double> h_v; // interleaved re, im
  std::vector<uint64_t> h_m, h_r, h_x, h_s;
  DeviceArray<double> d_v;
  DeviceArray<uint64_t> d_m, d_r, d_x, d_s;
  bool load(ls_hs_nonbranching_terms const *t);
  TermsView view() const {
    return TermsView{d_v.ptr, d_m.ptr, d_r.ptr, d_x.ptr, d_s.ptr, count};
  }
  uint64_t fingerprint() const;
};

// What the product has to do with every beta (BatchedOperator.chpl:264-282).
// kFactored: kPermutations through the translation factorisation of the
// group (state_info_factored).
enum ProjectionKind { kNoProjection = 0, kInversionOnly = 1, kPermutations = 2, kFactored = 3 };

struct Csr {
  int64_t rows = 0;
  int64_t nnz = 0;
  DeviceArray<int64_t> row_ptr;
  DeviceArray<uint32_t> cols;
  DeviceArray<double> vals64;
  DeviceArray<float> vals32;
  DeviceArray<double> diag64;
  DeviceArray<float> diag32;
  bool has64 = false, has32 = false;
  bool is_complex = false; // values are (re, im) pairs of M[i, j]
  // Value table: when the matrix has at most 256 distinct off-diagonal
  // values and fewer than 2^24 rows, cols[k] holds column | (value index
  // << 24) and vals64 / vals32 are replaced by table64 / table32.  Wide
  // table (at most kWideTableSize values, any number of rows): cols[k] is the
  // column, vidx[k] the value's index in table64 / table32.
  bool packed = false;
  bool wide = false;
  int table_size = 0;
  DeviceArray<double> table64;
  DeviceArray<float> table32;
  DeviceArray<uint16_t> vidx;
  bool has_table() const { return packed || wide; }
};

struct OperatorCore {
  // basis side
  int number_bits = 0;
  int spin_inversion = 0;
  ProjectionKind kind = kNoProjection;
  GroupData const *group = nullptr; // kPermutations only
  IndexView index{};                // representatives + prefix table
  bool identity_index = false;      // state index is the state itself
  int64_t number_states = 0;
  double const *d_norms = nullptr;  // state_info norms of the representatives
  DeviceArray<double> owned_norms;
  // terms
  Terms diag, off;
  // CSR cache
  Csr csr;
  // The FP64 schedule serves every product that gathers at most 16 B of x
  // per entry (the FP64 vector, FP64 blocks of 2, FP32 blocks up to 4), the
  // FP32 one the wider blocks (FP32 from 8 columns, FP64 from 4).
  int lanes_f64 = 8;  // lanes per row of the FP64 SpMV
  int lanes_f32 = 8;  // lanes per row of the FP32 block-of-8 SpMV
  int flags_f64 = 0;  // bit 0 non-temporal stream loads, bit 1 die-aware persistent grid
  int flags_f32 = 0;
  int persistent_blocks = 1024; // grid of the die-aware schedule (a multiple of 8)
  bool schedule_set = false;    // set by the caller or the autotuner: the CSR build keeps it
  bool allow_packing = true;    // value-table CSR when possible
  bool allow_factored = true;   // factorised orbit search when the group has one
  DeviceArray<int> d_err; // error flag + offending state
  DeviceArray<uint64_t> d_err_state;
  // <u, M v> = <M u, v> for real / complex vectors: -1 untested, 0 no, 1 yes
  int hermitian[2] = {-1, -1};
};

bool operator_prepare_norms(OperatorCore &op);
bool operator_matvec_otf(OperatorCore &op, double const *d_x, double *d_y, bool accumulate,
                         bool complex_values);
bool operator_build_csr(OperatorCore &op, int precision);
bool operator_csr_f64(OperatorCore &op, double const *d_x, double *d_y, bool accumulate);
// The same product in pieces, for overlapping it with copies of y: first
// y = diag x over all rows (operator_csr_f64_diag, when the operator has
// diagonal terms), then y[row] += the row's sum for rows [r0, r1)
// (operator_csr_f64_rows).  Each row gets the bits operator_csr_f64 gives it.
bool operator_csr_f64_diag(OperatorCore &op, double const *d_x, double *d_y);
bool operator_csr_f64_rows(OperatorCore &op, double const *d_x, double *d_y, int64_t r0, int64_t r1);
bool operator_csr_f32_block(OperatorCore &op, int k, float const *d_x, float *d_y);
bool operator_csr_f64_block(OperatorCore &op, int k, double const *d_x, double *d_y);
int operator_autotune(OperatorCore &op);

// Relative asymmetry of a product on seeded vectors (lanczos.hip), and the
// operator's cached answer for a product mode (0 FP64 CSR, 1 FP32 CSR,
// 2 on-the-fly, 3 on-the-fly complex).
bool asymmetry_test(int64_t n_reals, bool complex_values, std::function<bool(double const *, double *)> const &prod,
                    uint64_t seed, double &rel);
bool operator_is_hermitian(OperatorCore &op, int mode, bool &hermitian);

// Term application on device arrays (reference.c:67-134).
bool apply_diag_device(Terms const &t, int64_t count, uint64_t const *d_alphas, double *d_ys,
                       double const *d_xs);
// Two passes: counts -> offsets, then betas / coeffs.  d_offsets has count+1
// entries (int64).
bool apply_off_diag_device(Terms const &t, int64_t count, uint64_t const *d_alphas,
                           double const *d_xs, DeviceArray<uint64_t> &betas,
                           DeviceArray<double> &coeffs, DeviceArray<int64_t> &offsets,
                           int64_t &total);

bool exclusive_scan(uint32_t const *d_counts, int64_t n, uint64_t *d_offsets, uint64_t &total);

// Thick-restart block Lanczos on FP64 row-major [rows][block] blocks over
// independent segments of rows (block_krylov.hip).
struct BlockKrylovOptions {
  int block = 8;       // columns per block
  int nev = 4;         // eigenvalues wanted per segment
  int max_blocks = 12; // Krylov blocks per cycle
  int keep = 2;        // Ritz blocks kept at a restart
  int max_cycles = 40;
  double tol = 1e-7;        // relative residual for convergence with the FP64 product
  double switch_tol = 1e-5; // relative residual at which FP32 gives way to FP64
  uint64_t seed = 1;
  bool fp32 = true;         // start with the FP32 product
  // blocks of `block` complex columns, each row stored as [re | im] (2 * block
  // reals), with complex (Hermitian) arithmetic on the host side
  bool complex_values = false;
};

struct BlockKrylovResult {
  std::vector<double> evals;     // segments * nev, ascending per segment
  std::vector<double> residuals; // FP64 relative residual norms
  std::vector<double> history;   // largest residual bound after every cycle
  int cycles = 0, fp64_cycles = 0, blocks = 0;
  bool converged = false;
  int64_t products = 0;
  double product_seconds = 0.0, total_seconds = 0.0, max_residual = 0.0;
  double first_rayleigh = 0.0, first_residual = 0.0; // lowest Ritz vector of segment 0
};

// y = A x on a block; with fp32 the FP32 product on float blocks, otherwise
// the FP64 product on double blocks.
using BlockProduct = std::function<bool(bool fp32, void const *x, void *y)>;

// Each eigenvalue appears `multiplicity` times in the segments' spectra (2 for
// the real embedding of complex sectors); d_ritz_block (may be null) receives
// the lowest Ritz block.  Segment s spans rows [seg_rows[s], seg_rows[s + 1]);
// only segments with active[s] != 0 are solved (all of them if `active` is
// empty), the others' results are NaN and their rows are left to the product.
bool check_block_krylov_options(BlockKrylovOptions const &o, int multiplicity);
bool block_krylov_solve(std::vector<int64_t> const &seg_rows, std::vector<char> const &active, int multiplicity,
                        BlockProduct const &product, BlockKrylovOptions const &o, BlockKrylovResult &res,
                        double *d_ritz_block);

// Dense symmetric eigen-decomposition (dense_eigh.cpp): A is n x n row-major;
// w receives the eigenvalues ascending and Z the eigenvectors as columns.
bool symmetric_eigh(std::vector<double> A, int n, std::vector<double> &w, std::vector<double> &Z);
// Complex Hermitian eigen-decomposition (complex Householder + the same QR), same conventions.
bool hermitian_eigh(std::vector<std::complex<double>> A, int n, std::vector<double> &w,
                    std::vector<std::complex<double>> &Z);
bool block_column(double const *X, int b, int col, double *out, int64_t rows);

} // namespace lsamd

struct ls_amd_operator {
  ls_amd_basis const *basis = nullptr;
  lsamd::OperatorCore core;
};
