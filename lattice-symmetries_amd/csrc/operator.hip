// operator.hip -- term application, fused product, CSR cache and SpMV.
//
// The reference applies H in push form: every representative alpha_i emits
// (beta, c) pairs, c is rescaled by chi_beta n_beta / n_alpha after
// state_info, and y[index(rep beta)] += Re(c) x_i with relaxed atomics
// (chapel/src/BatchedOperator.chpl:207-253,
// chapel/src/DistributedMatrixVector.chpl:91-143).  The engine computes the
// same matrix M, M[index(rep beta), i] = C, for every operator, Hermitian or
// not:
// - the fused on-the-fly product is push form too: one lane per alpha_i,
//   the diagonal pass first, then atomic adds of Re(c) (or c x_i for complex
//   vectors) into y;
// - the CSR build emits the pairs of each alpha_i as a source row, then
//   transposes them on the device (count -> scan -> scatter -> per-row sort),
//   so that row j of the cached matrix holds every (i, C) with target j, in
//   ascending i.  Products then gather x along a row and write y once,
//   without atomics, and every build stores the same bits.
#include <algorithm>
#include <cstring>

#include "operator_internal.hpp"
#include "orbit_device.hpp"

namespace lsamd {

// ---------------------------------------------------------------------------
// Terms
// ---------------------------------------------------------------------------

bool Terms::load(ls_hs_nonbranching_terms const *t) {
  count = 0;
  h_v.clear();
  h_m.clear();
  h_r.clear();
  h_x.clear();
  h_s.clear();
  if (t == nullptr || t->number_terms == 0) return true;
  if (t->number_bits > 64) return fail("bases with more than 64 bits are not yet implemented");
  count = t->number_terms;
  h_v.resize(2 * static_cast<size_t>(count));
  std::memcpy(h_v.data(), t->v, sizeof(double) * 2 * count);
  h_m.assign(t->m, t->m + count);
  h_r.assign(t->r, t->r + count);
  h_x.assign(t->x, t->x + count);
  h_s.assign(t->s, t->s + count);
  return d_v.upload(h_v.data(), h_v.size()) && d_m.upload(h_m.data(), count) &&
         d_r.upload(h_r.data(), count) && d_x.upload(h_x.data(), count) &&
         d_s.upload(h_s.data(), count);
}

uint64_t Terms::fingerprint() const {
  uint64_t h = 1469598103934665603ull ^ static_cast<uint64_t>(count);
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for (int i = 0; i < count; ++i) {
    uint64_t vb[2];
    std::memcpy(vb, &h_v[2 * i], 16);
    mix(vb[0]);
    mix(vb[1]);
    mix(h_m[i]);
    mix(h_r[i]);
    mix(h_x[i]);
    mix(h_s[i]);
  }
  return h;
}

// ---------------------------------------------------------------------------
// reference.c:67-134 on device arrays
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void apply_diag_kernel(TermsView t, int64_t count,
                                                         uint64_t const *__restrict__ alphas,
                                                         double *__restrict__ ys,
                                                         double const *__restrict__ xs) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t const a = alphas[i];
  double acc = 0;
  for (int k = 0; k < t.count; ++k) {
    int const delta = (a & t.m[k]) == t.r[k];
    if (delta) {
      int const sign = 1 - 2 * (__popcll(a & t.s[k]) % 2);
      double const factor =
          xs != nullptr ? static_cast<double>(delta * sign) * xs[i] : static_cast<double>(delta * sign);
      acc += t.v[2 * k] * factor;
    }
  }
  ys[i] = acc;
}

__global__ __launch_bounds__(256) void count_applicable_kernel(TermsView t, int64_t count,
                                                               uint64_t const *__restrict__ alphas,
                                                               uint32_t *__restrict__ counts) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t const a = alphas[i];
  uint32_t c = 0;
  for (int k = 0; k < t.count; ++k) c += (a & t.m[k]) == t.r[k];
  counts[i] = c;
}

__global__ __launch_bounds__(256) void apply_off_diag_kernel(TermsView t, int64_t count,
                                                             uint64_t const *__restrict__ alphas,
                                                             double const *__restrict__ xs,
                                                             int64_t const *__restrict__ offsets,
                                                             uint64_t *__restrict__ betas,
                                                             double *__restrict__ coeffs) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t const a = alphas[i];
  int64_t o = offsets[i];
  for (int k = 0; k < t.count; ++k) {
    int const delta = (a & t.m[k]) == t.r[k];
    if (delta) {
      int const sign = 1 - 2 * (__popcll(a & t.s[k]) % 2);
      double const factor =
          xs != nullptr ? static_cast<double>(delta * sign) * xs[i] : static_cast<double>(delta * sign);
      coeffs[2 * o] = t.v[2 * k] * factor;
      coeffs[2 * o + 1] = t.v[2 * k + 1] * factor;
      betas[o] = a ^ t.x[k];
      ++o;
    }
  }
}

bool apply_diag_device(Terms const &t, int64_t count, uint64_t const *d_alphas, double *d_ys,
                       double const *d_xs) {
  if (count <= 0) return true;
  if (t.count == 0) {
    LSAMD_HIP(hipMemsetAsync(d_ys, 0, sizeof(double) * count, stream()));
    return true;
  }
  apply_diag_kernel<<<grid_for(count, 256), 256, 0, stream()>>>(t.view(), count, d_alphas, d_ys, d_xs);
  LSAMD_HIP(hipGetLastError());
  return true;
}

bool apply_off_diag_device(Terms const &t, int64_t count, uint64_t const *d_alphas,
                           double const *d_xs, DeviceArray<uint64_t> &betas,
                           DeviceArray<double> &coeffs, DeviceArray<int64_t> &offsets,
                           int64_t &total) {
  total = 0;
  if (!offsets.allocate(static_cast<size_t>(count + 1))) return false;
  if (t.count == 0 || count == 0) {
    LSAMD_HIP(hipMemsetAsync(offsets.ptr, 0, sizeof(int64_t) * (count + 1), stream()));
    betas.reset();
    coeffs.reset();
    return true;
  }
  DeviceArray<uint32_t> counts;
  if (!counts.allocate(static_cast<size_t>(count))) return false;
  count_applicable_kernel<<<grid_for(count, 256), 256, 0, stream()>>>(t.view(), count, d_alphas,
                                                                     counts.ptr);
  LSAMD_HIP(hipGetLastError());
  uint64_t tot = 0;
  if (!exclusive_scan(counts.ptr, count, reinterpret_cast<uint64_t *>(offsets.ptr), tot)) return false;
  int64_t const tot_i = static_cast<int64_t>(tot);
  LSAMD_HIP(hipMemcpyAsync(offsets.ptr + count, &tot_i, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
  if (!betas.allocate(static_cast<size_t>(tot > 0 ? tot : 1))) return false;
  if (!coeffs.allocate(static_cast<size_t>(2 * (tot > 0 ? tot : 1)))) return false;
  apply_off_diag_kernel<<<grid_for(count, 256), 256, 0, stream()>>>(t.view(), count, d_alphas, d_xs,
                                                                   offsets.ptr, betas.ptr, coeffs.ptr);
  LSAMD_HIP(hipGetLastError());
  LSAMD_HIP(hipStreamSynchronize(stream()));
  total = tot_i;
  return true;
}

// ---------------------------------------------------------------------------
// Fused product
// ---------------------------------------------------------------------------

struct OpView {
  TermsView diag;
  TermsView off;
  IndexView idx;
  double const *norms;
  int64_t n;
  uint64_t flip;
  int identity_index;
};

OpView make_view(OperatorCore const &op) {
  return OpView{op.diag.view(), op.off.view(), op.index, op.d_norms, op.number_states,
                op.number_bits >= 64 ? ~uint64_t{0} : ((uint64_t{1} << op.number_bits) - 1),
                op.identity_index ? 1 : 0};
}

// Off-diagonal terms staged in LDS: m, r, x, s then v (re, im).
struct TermsLds {
  uint64_t *m, *r, *x, *s;
  double *v;
};

__device__ __forceinline__ TermsLds stage_terms(TermsView const &t, uint64_t *smem) {
  int const T = t.count;
  TermsLds l{smem, smem + T, smem + 2 * T, smem + 3 * T, reinterpret_cast<double *>(smem + 4 * T)};
  for (int k = threadIdx.x; k < T; k += blockDim.x) {
    l.m[k] = t.m[k];
    l.r[k] = t.r[k];
    l.x[k] = t.x[k];
    l.s[k] = t.s[k];
    l.v[2 * k] = t.v[2 * k];
    l.v[2 * k + 1] = t.v[2 * k + 1];
  }
  __syncthreads();
  return l;
}

inline size_t terms_lds_bytes(int T) { return static_cast<size_t>(T) * (4 * 8 + 16); }

__device__ __forceinline__ int next_applicable(TermsLds const &l, int T, uint64_t a, int k) {
  while (k < T && (a & l.m[k]) != l.r[k]) ++k;
  return k;
}

// Coefficient of term k for alpha with the factor `scale` (x_alpha, or 1):
// returns the representative of beta in `beta` and
// c = chi_beta v (-1)^popcount(alpha & s) scale n_beta / n_alpha in
// (c_re, c_im), with the reference's order of operations
// (reference.c:97-134 multiplies v by sign * x; BatchedOperator.chpl:241-247
// scales by n_beta / n_alpha and then multiplies by chi).
template <typename W, int D, int SI, int KIND>
__device__ __forceinline__ void off_diag_entry(GroupView const &g, OpView const &op,
                                               TermsLds const &l, int k, uint64_t alpha,
                                               double n_alpha, double scale, uint64_t &beta,
                                               double &c_re, double &c_im, RowTables const *rt) {
  double const factor = ((__popcll(alpha & l.s[k]) & 1) ? -1.0 : 1.0) * scale;
  double re = l.v[2 * k] * factor, im = l.v[2 * k + 1] * factor;
  beta = alpha ^ l.x[k];
  if constexpr (KIND == kInversionOnly) {
    uint64_t const inv = beta ^ op.flip; // BatchedOperator.chpl:192-199
    if (inv < beta) {
      beta = inv;
      re *= static_cast<double>(SI);
      im *= static_cast<double>(SI);
    }
  } else if constexpr (KIND == kPermutations || KIND == kFactored) {
    // BatchedOperator.chpl:232-247
    StateInfo const s =
        KIND == kFactored ? state_info_factored<W, D, SI>(g, beta, rt) : state_info_one<W, D, SI>(g, beta);
    double const t_re = (re * s.norm) / n_alpha, t_im = (im * s.norm) / n_alpha;
    re = s.c_re * t_re - s.c_im * t_im;
    im = s.c_re * t_im + s.c_im * t_re;
    beta = s.rep;
  }
  c_re = re;
  c_im = im;
}

__device__ __forceinline__ void record_error(int *err, uint64_t *err_state, uint64_t beta) {
  if (atomicCAS(err, 0, 1) == 0) *err_state = beta;
}

// Diagonal pass of the fused product (DistributedMatrixVector.chpl:34-71):
// y_i = sum_k Re(v_k) (-1)^.. x_i in the term order of reference.c:67-95;
// without diagonal terms y is zeroed, or left as it is when accumulating.
template <bool CPLX>
__global__ __launch_bounds__(256) void otf_diag_kernel(OpView op, double const *__restrict__ x,
                                                       double *__restrict__ y, int accumulate) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= op.n) return;
  if (op.diag.count == 0) {
    if (!accumulate) {
      if constexpr (CPLX) {
        y[2 * i] = 0.0;
        y[2 * i + 1] = 0.0;
      } else {
        y[i] = 0.0;
      }
    }
    return;
  }
  uint64_t const alpha = op.idx.reps[i];
  double const xr = CPLX ? x[2 * i] : x[i];
  double const xi = CPLX ? x[2 * i + 1] : 0.0;
  double base_re = 0, base_im = 0;
  for (int k = 0; k < op.diag.count; ++k) {
    if ((alpha & op.diag.m[k]) != op.diag.r[k]) continue;
    double const sign = (__popcll(alpha & op.diag.s[k]) & 1) ? -1.0 : 1.0;
    base_re += op.diag.v[2 * k] * (sign * xr);
    if constexpr (CPLX) base_im += op.diag.v[2 * k] * (sign * xi);
  }
  if constexpr (CPLX) {
    y[2 * i] = base_re;
    y[2 * i + 1] = base_im;
  } else {
    y[i] = base_re;
  }
}

// Off-diagonal pass, push form (DistributedMatrixVector.chpl:91-143): lane i
// emits the pairs of alpha_i and adds them into y at index(rep beta).  Real
// vectors follow the reference exactly: c includes x_i, and a pair whose
// real part is 0 is skipped before the index lookup.  Complex vectors add
// c x_i for every nonzero c.
template <typename W, int D, int SI, int KIND, bool CPLX>
__global__ __launch_bounds__(256) void otf_push_kernel(GroupView g, OpView op, double const *__restrict__ x,
                                                       double *__restrict__ y, int *err, uint64_t *err_state) {
  extern __shared__ uint64_t smem[];
  TermsLds const l = stage_terms(op.off, smem);
  __shared__ RowTables rt;
  if constexpr (KIND == kFactored) row_tables_init(g, rt);
  int const T = op.off.count;
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool const valid = i < op.n;
  uint64_t const alpha = valid ? op.idx.reps[i] : 0;
  double const xr = valid ? (CPLX ? x[2 * i] : x[i]) : 0.0;
  double const xi = (CPLX && valid) ? x[2 * i + 1] : 0.0;
  double const n_alpha = ((KIND == kPermutations || KIND == kFactored) && valid) ? op.norms[i] : 1.0;
  int k = valid ? next_applicable(l, T, alpha, 0) : T;
  while (__any(k < T)) {
    if (k < T) {
      uint64_t beta;
      double c_re, c_im;
      off_diag_entry<W, D, SI, KIND>(g, op, l, k, alpha, n_alpha, CPLX ? 1.0 : xr, beta, c_re, c_im, &rt);
      bool const nz = CPLX ? (c_re != 0 || c_im != 0) : (c_re != 0);
      if (nz) {
        int64_t const j = op.identity_index ? static_cast<int64_t>(beta) : index_lookup(op.idx, beta);
        if (j < 0 || j >= op.n) {
          record_error(err, err_state, beta);
        } else if constexpr (CPLX) {
          unsafeAtomicAdd(&y[2 * j], c_re * xr - c_im * xi);
          unsafeAtomicAdd(&y[2 * j + 1], c_re * xi + c_im * xr);
        } else {
          unsafeAtomicAdd(&y[j], c_re);
        }
      }
      k = next_applicable(l, T, alpha, k + 1);
    }
  }
}

// Dispatch over (W, D, SI, KIND): permutation kernels follow the group,
// the others run on 64-bit words with the spin inversion as SI.
template <typename F> bool dispatch_op(OperatorCore const &op, F &&f) {
  if (op.kind == kPermutations) {
    bool const factored = op.allow_factored && op.group->factored();
    return dispatch_group(*op.group, [&]<typename W, int D, int SI>() -> bool {
      if constexpr (SI != kFermionic) {
        if (factored) return f.template operator()<W, D, SI, kFactored>();
      }
      return f.template operator()<W, D, SI, kPermutations>();
    });
  }
  if (op.kind == kInversionOnly) {
    if (op.spin_inversion == 1) return f.template operator()<uint64_t, 0, 1, kInversionOnly>();
    return f.template operator()<uint64_t, 0, -1, kInversionOnly>();
  }
  return f.template operator()<uint64_t, 0, 0, kNoProjection>();
}

GroupView group_view_or_empty(OperatorCore const &op) {
  if (op.group != nullptr) return op.group->view();
  return GroupView{nullptr, nullptr, nullptr, nullptr, 0, 0, 0};
}

bool check_error(OperatorCore &op) {
  int err = 0;
  uint64_t state = 0;
  LSAMD_HIP(hipMemcpyAsync(&err, op.d_err.ptr, sizeof(int), hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipMemcpyAsync(&state, op.d_err_state.ptr, sizeof(uint64_t), hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  if (err != 0) {
    LSAMD_HIP(hipMemsetAsync(op.d_err.ptr, 0, sizeof(int), stream()));
    // DistributedMatrixVector.chpl:133-135
    return fail("invalid index: -1 for state " + std::to_string(state) +
                " (the operator does not respect the basis symmetries)");
  }
  return true;
}

bool ensure_err(OperatorCore &op) {
  if (op.d_err.ptr == nullptr) {
    if (!op.d_err.allocate(1) || !op.d_err_state.allocate(1)) return false;
    LSAMD_HIP(hipMemsetAsync(op.d_err.ptr, 0, sizeof(int), stream()));
  }
  return true;
}

// state_info norm of every representative (n_alpha of BatchedOperator.chpl:244).
bool operator_prepare_norms(OperatorCore &op) {
  if (op.kind != kPermutations || op.d_norms != nullptr) return true;
  int64_t const n = op.number_states;
  DeviceArray<uint64_t> reps_out;
  DeviceArray<double> chars;
  if (!op.owned_norms.allocate(static_cast<size_t>(n > 0 ? n : 1))) return false;
  if (!reps_out.allocate(static_cast<size_t>(n > 0 ? n : 1))) return false;
  if (!chars.allocate(static_cast<size_t>(2 * (n > 0 ? n : 1)))) return false;
  if (!launch_state_info(*op.group, n, op.index.reps, 1, reps_out.ptr, 1, chars.ptr, op.owned_norms.ptr))
    return false;
  LSAMD_HIP(hipStreamSynchronize(stream()));
  op.d_norms = op.owned_norms.ptr;
  return true;
}

bool operator_matvec_otf(OperatorCore &op, double const *d_x, double *d_y, bool accumulate,
                         bool complex_values) {
  if (!ensure_err(op) || !operator_prepare_norms(op)) return false;
  int64_t const n = op.number_states;
  if (n == 0) return true;
  OpView const v = make_view(op);
  GroupView const g = group_view_or_empty(op);
  size_t const smem = terms_lds_bytes(op.off.count);
  timer_begin(kTimerOtfF64);
  if (complex_values)
    otf_diag_kernel<true><<<grid_for(n, 256), 256, 0, stream()>>>(v, d_x, d_y, accumulate ? 1 : 0);
  else
    otf_diag_kernel<false><<<grid_for(n, 256), 256, 0, stream()>>>(v, d_x, d_y, accumulate ? 1 : 0);
  LSAMD_HIP(hipGetLastError());
  bool ok = op.off.count == 0 || dispatch_op(op, [&]<typename W, int D, int SI, int KIND>() -> bool {
    if (complex_values)
      otf_push_kernel<W, D, SI, KIND, true><<<grid_for(n, 256), 256, smem, stream()>>>(
          g, v, d_x, d_y, op.d_err.ptr, op.d_err_state.ptr);
    else
      otf_push_kernel<W, D, SI, KIND, false><<<grid_for(n, 256), 256, smem, stream()>>>(
          g, v, d_x, d_y, op.d_err.ptr, op.d_err_state.ptr);
    LSAMD_HIP(hipGetLastError());
    return true;
  });
  timer_end(kTimerOtfF64);
  if (!ok) return false;
  return check_error(op);
}

// ---------------------------------------------------------------------------
// CSR cache
// ---------------------------------------------------------------------------

constexpr uint32_t kDropped = 0xffffffffu;

template <typename W, int D, int SI, int KIND, bool CPLX>
__global__ __launch_bounds__(256) void csr_fill_kernel(GroupView g, OpView op,
                                                       int64_t const *__restrict__ row_ptr,
                                                       uint32_t *__restrict__ cols,
                                                       double *__restrict__ vals,
                                                       double *__restrict__ diag,
                                                       uint32_t *__restrict__ kept,
                                                       int *err, uint64_t *err_state) {
  extern __shared__ uint64_t smem[];
  TermsLds const l = stage_terms(op.off, smem);
  __shared__ RowTables rt;
  if constexpr (KIND == kFactored) row_tables_init(g, rt);
  int const T = op.off.count;
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool const valid = i < op.n;
  uint64_t const alpha = valid ? op.idx.reps[i] : 0;
  if (valid) {
    double d = 0;
    for (int k = 0; k < op.diag.count; ++k) {
      if ((alpha & op.diag.m[k]) != op.diag.r[k]) continue;
      double const sign = (__popcll(alpha & op.diag.s[k]) & 1) ? -1.0 : 1.0;
      d += op.diag.v[2 * k] * sign;
    }
    diag[i] = d;
  }
  double const n_alpha = ((KIND == kPermutations || KIND == kFactored) && valid) ? op.norms[i] : 1.0;
  int64_t o = valid ? row_ptr[i] : 0;
  uint32_t nkept = 0;
  int k = valid ? next_applicable(l, T, alpha, 0) : T;
  while (__any(k < T)) {
    if (k < T) {
      uint64_t beta;
      double c_re, c_im;
      off_diag_entry<W, D, SI, KIND>(g, op, l, k, alpha, n_alpha, 1.0, beta, c_re, c_im, &rt);
      uint32_t col = kDropped;
      double val_re = 0.0, val_im = 0.0;
      if (CPLX ? (c_re != 0 || c_im != 0) : (c_re != 0)) {
        int64_t const j = op.identity_index ? static_cast<int64_t>(beta) : index_lookup(op.idx, beta);
        if (j < 0 || j >= op.n) {
          record_error(err, err_state, beta);
        } else {
          col = static_cast<uint32_t>(j);
          val_re = c_re;
          val_im = c_im;
          ++nkept;
        }
      }
      cols[o] = col;
      if constexpr (CPLX) {
        vals[2 * o] = val_re;
        vals[2 * o + 1] = val_im;
      } else {
        vals[o] = val_re;
      }
      ++o;
      k = next_applicable(l, T, alpha, k + 1);
    }
  }
  if (valid) kept[i] = nkept;
}

__global__ __launch_bounds__(256) void csr_compact_kernel(int64_t n, int64_t const *__restrict__ old_ptr,
                                                          uint32_t const *__restrict__ old_cols,
                                                          double const *__restrict__ old_vals,
                                                          uint64_t const *__restrict__ new_ptr,
                                                          uint32_t *__restrict__ cols,
                                                          double *__restrict__ vals, int width) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t o = new_ptr[i];
  for (int64_t k = old_ptr[i]; k < old_ptr[i + 1]; ++k) {
    if (old_cols[k] == kDropped) continue;
    cols[o] = old_cols[k];
    for (int q = 0; q < width; ++q) vals[o * width + q] = old_vals[k * width + q];
    ++o;
  }
}

// ---------------------------------------------------------------------------
// Source rows -> target rows.  Entry t of source row i (column j, value C)
// becomes an entry of row j with column i.  Keys i << 24 | t are scattered
// into the target rows through atomic cursors, each target row's keys are
// sorted, and the sorted keys give the final columns and values: row j then
// lists its sources in ascending i (in term order within one i), whatever
// order the atomics took.
// ---------------------------------------------------------------------------

constexpr int kKeyLocalBits = 24; // entries per source row < 2^24 (one per term at most)
constexpr int kSortCap = 1024;    // longest row sorted in LDS; longer rows are heap-sorted in place

// The count and scatter passes give kPushLanes lanes to a source row: lane l
// takes its entries b + l, b + l + kPushLanes, ..., so a wave reads the
// column stream coalesced and has a row's atomics in flight together.  Slot
// order within a target row follows the atomics, and the sort that comes
// next fixes it.
constexpr int kPushLanes = 16;

__global__ __launch_bounds__(256) void push_count_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                         uint32_t const *__restrict__ cols,
                                                         uint32_t *__restrict__ counts) {
  int64_t const i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kPushLanes;
  int const l = static_cast<int>(threadIdx.x) % kPushLanes;
  if (i >= n) return;
  int64_t const e = rp[i + 1];
  for (int64_t p = rp[i] + l; p < e; p += kPushLanes) atomicAdd(&counts[cols[p]], 1u);
}

// Sort keys of the transpose: the source row in the high bits (the column of
// M, the order within a target row) and, below it, the entry's global
// position when every position fits 32 bits (wide keys), else its position
// within the source row (kKeyLocalBits).  Both sort a row the same way.
__global__ __launch_bounds__(256) void push_scatter_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                           uint32_t const *__restrict__ cols,
                                                           int64_t const *__restrict__ target_ptr,
                                                           uint32_t *__restrict__ cursor,
                                                           uint64_t *__restrict__ keys, int wide_keys) {
  int64_t const i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kPushLanes;
  int const l = static_cast<int>(threadIdx.x) % kPushLanes;
  if (i >= n) return;
  int64_t const b = rp[i], e = rp[i + 1];
  for (int64_t p = b + l; p < e; p += kPushLanes) {
    uint32_t const j = cols[p];
    uint32_t const slot = atomicAdd(&cursor[j], 1u);
    keys[target_ptr[j] + slot] = wide_keys ? (static_cast<uint64_t>(i) << 32) | static_cast<uint64_t>(p)
                                           : (static_cast<uint64_t>(i) << kKeyLocalBits) | static_cast<uint64_t>(p - b);
  }
}

__device__ void heap_sort(uint64_t *a, int64_t n) {
  auto sift = [a](int64_t root, int64_t end) {
    while (true) {
      int64_t child = 2 * root + 1;
      if (child >= end) return;
      if (child + 1 < end && a[child] < a[child + 1]) ++child;
      if (a[root] >= a[child]) return;
      uint64_t const t = a[root];
      a[root] = a[child];
      a[child] = t;
      root = child;
    }
  };
  for (int64_t r = n / 2 - 1; r >= 0; --r) sift(r, n);
  for (int64_t end = n - 1; end > 0; --end) {
    uint64_t const t = a[0];
    a[0] = a[end];
    a[end] = t;
    sift(0, end);
  }
}

// Largest entry of counts[0, n) into *out (which starts at zero).
__global__ __launch_bounds__(256) void max_count_kernel(uint32_t const *__restrict__ counts, int64_t n,
                                                        uint32_t *__restrict__ out) {
  uint32_t m = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    m = max(m, counts[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), off, 64)));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// One wave per row: bitonic sort of the row's keys in LDS (dynamic, `cap`
// keys: the longest row's power of two, at most kSortCap, so that more
// workgroups fit a CU), padded to a power of two with all-ones keys.  Longer
// rows are heap-sorted in place.  The row index is uniform over the
// workgroup, so every barrier is reached by all its threads.
__global__ __launch_bounds__(64) void sort_rows_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                       uint64_t *__restrict__ keys, int cap) {
  extern __shared__ uint64_t s[];
  for (int64_t row = blockIdx.x; row < n; row += gridDim.x) {
    int64_t const b = rp[row];
    int64_t const len = rp[row + 1] - b;
    if (len <= 1) continue;
    if (len > cap) {
      if (threadIdx.x == 0) heap_sort(keys + b, len);
      continue;
    }
    int P = 2;
    while (P < len) P <<= 1;
    for (int q = threadIdx.x; q < P; q += 64) s[q] = q < len ? keys[b + q] : ~uint64_t{0};
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int q = threadIdx.x; q < P / 2; q += 64) {
          int const lo = 2 * stride * (q / stride) + (q % stride);
          int const hi = lo + stride;
          bool const ascending = (lo & size) == 0;
          uint64_t const a = s[lo], c = s[hi];
          if ((a > c) == ascending) {
            s[lo] = c;
            s[hi] = a;
          }
        }
        __syncthreads();
      }
    }
    for (int q = threadIdx.x; q < len; q += 64) keys[b + q] = s[q];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void push_finalize_kernel(int64_t nnz, uint64_t const *__restrict__ keys,
                                                            int64_t const *__restrict__ src_ptr,
                                                            double const *__restrict__ src_vals, int width,
                                                            uint32_t *__restrict__ cols,
                                                            double *__restrict__ vals, int wide_keys) {
  int64_t const q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  uint64_t const key = keys[q];
  int64_t i, p;
  if (wide_keys) {
    i = static_cast<int64_t>(key >> 32);
    p = static_cast<int64_t>(key & 0xffffffffull);
  } else {
    i = static_cast<int64_t>(key >> kKeyLocalBits);
    p = src_ptr[i] + static_cast<int64_t>(key & ((uint64_t{1} << kKeyLocalBits) - 1));
  }
  cols[q] = static_cast<uint32_t>(i);
  for (int w = 0; w < width; ++w) vals[q * width + w] = src_vals[p * width + w];
}

// Replaces the source-row CSR (row_ptr, cols, vals of `width` doubles per
// entry, nnz entries over n rows) by its target-row form.
bool transpose_to_targets(int64_t n, int64_t nnz, int width, DeviceArray<int64_t> &row_ptr,
                          DeviceArray<uint32_t> &cols, DeviceArray<double> &vals) {
  if (n == 0 || nnz == 0) return true;
  DeviceArray<uint32_t> counts;
  DeviceArray<int64_t> target_ptr;
  DeviceArray<uint64_t> keys;
  if (!counts.allocate(static_cast<size_t>(n)) || !target_ptr.allocate(static_cast<size_t>(n + 1)) ||
      !keys.allocate(static_cast<size_t>(nnz)))
    return false;
  LSAMD_HIP(hipMemsetAsync(counts.ptr, 0, sizeof(uint32_t) * n, stream()));
  push_count_kernel<<<grid_for(n * kPushLanes, 256), 256, 0, stream()>>>(n, row_ptr.ptr, cols.ptr, counts.ptr);
  LSAMD_HIP(hipGetLastError());
  // the longest target row sizes the sort's LDS
  DeviceArray<uint32_t> longest;
  if (!longest.allocate(1)) return false;
  LSAMD_HIP(hipMemsetAsync(longest.ptr, 0, sizeof(uint32_t), stream()));
  max_count_kernel<<<static_cast<int>(std::min<int64_t>(grid_for(n, 256), 1024)), 256, 0, stream()>>>(counts.ptr, n,
                                                                                                  longest.ptr);
  LSAMD_HIP(hipGetLastError());
  uint32_t max_len = 0;
  LSAMD_HIP(hipMemcpyAsync(&max_len, longest.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  int sort_cap = 2;
  while (sort_cap < kSortCap && sort_cap < static_cast<int>(std::min<uint32_t>(max_len, kSortCap))) sort_cap <<= 1;
  uint64_t total = 0;
  if (!exclusive_scan(counts.ptr, n, reinterpret_cast<uint64_t *>(target_ptr.ptr), total)) return false;
  if (static_cast<int64_t>(total) != nnz) return fail("CSR transpose: entry count mismatch");
  int64_t const last = nnz;
  LSAMD_HIP(hipMemcpyAsync(target_ptr.ptr + n, &last, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
  LSAMD_HIP(hipMemsetAsync(counts.ptr, 0, sizeof(uint32_t) * n, stream()));
  int const wide_keys = nnz <= int64_t{0xffffffff} ? 1 : 0;
  push_scatter_kernel<<<grid_for(n * kPushLanes, 256), 256, 0, stream()>>>(n, row_ptr.ptr, cols.ptr,
                                                                           target_ptr.ptr, counts.ptr, keys.ptr,
                                                                           wide_keys);
  LSAMD_HIP(hipGetLastError());
  int const sort_grid = static_cast<int>(std::min<int64_t>(n, 1 << 16));
  sort_rows_kernel<<<sort_grid, 64, sizeof(uint64_t) * sort_cap, stream()>>>(n, target_ptr.ptr, keys.ptr, sort_cap);
  LSAMD_HIP(hipGetLastError());
  LSAMD_HIP(hipStreamSynchronize(stream()));
  counts.reset();
  cols.reset();
  DeviceArray<uint32_t> cols2;
  DeviceArray<double> vals2;
  if (!cols2.allocate(static_cast<size_t>(nnz)) || !vals2.allocate(static_cast<size_t>(nnz) * width)) return false;
  push_finalize_kernel<<<grid_for(nnz, 256), 256, 0, stream()>>>(nnz, keys.ptr, row_ptr.ptr, vals.ptr, width,
                                                                 cols2.ptr, vals2.ptr, wide_keys);
  LSAMD_HIP(hipGetLastError());
  LSAMD_HIP(hipStreamSynchronize(stream()));
  row_ptr = std::move(target_ptr);
  cols = std::move(cols2);
  vals = std::move(vals2);
  return true;
}

__global__ __launch_bounds__(256) void to_f32_kernel(int64_t n, double const *__restrict__ in,
                                                     float *__restrict__ out) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<float>(in[i]);
}

// ---------------------------------------------------------------------------
// Value table: distinct off-diagonal values (bit patterns) through a per-block
// LDS hash merged into a global hash, then column | index << 24 packing.
// ---------------------------------------------------------------------------

constexpr uint64_t kEmptySlot = ~uint64_t{0}; // a NaN pattern, never a stored value
constexpr int kLocalSlots = 2 * kWideTableSize; // per-block LDS hash (64 KB)
constexpr int kGlobalSlots = 4 * kWideTableSize;

__device__ __forceinline__ uint32_t slot_hash(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  return static_cast<uint32_t>(k);
}

__global__ __launch_bounds__(256) void distinct_values_kernel(double const *__restrict__ vals, int64_t nnz,
                                                              unsigned long long *__restrict__ table,
                                                              int *__restrict__ overflow) {
  __shared__ unsigned long long local[kLocalSlots];
  __shared__ int local_overflow;
  for (int i = threadIdx.x; i < kLocalSlots; i += blockDim.x) local[i] = kEmptySlot;
  if (threadIdx.x == 0) local_overflow = 0;
  __syncthreads();
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    unsigned long long const key = static_cast<unsigned long long>(__double_as_longlong(vals[k]));
    uint32_t slot = slot_hash(key) & (kLocalSlots - 1);
    int probes = 0;
    while (true) {
      unsigned long long const prev = atomicCAS(&local[slot], static_cast<unsigned long long>(kEmptySlot), key);
      if (prev == kEmptySlot || prev == key) break;
      slot = (slot + 1) & (kLocalSlots - 1);
      if (++probes >= kLocalSlots) {
        local_overflow = 1;
        break;
      }
    }
    if (local_overflow) break;
  }
  __syncthreads();
  if (local_overflow) {
    if (threadIdx.x == 0) atomicExch(overflow, 1);
    return;
  }
  for (int i = threadIdx.x; i < kLocalSlots; i += blockDim.x) {
    unsigned long long const key = local[i];
    if (key == kEmptySlot) continue;
    uint32_t slot = slot_hash(key) & (kGlobalSlots - 1);
    int probes = 0;
    while (true) {
      unsigned long long const prev = atomicCAS(&table[slot], static_cast<unsigned long long>(kEmptySlot), key);
      if (prev == kEmptySlot || prev == key) break;
      slot = (slot + 1) & (kGlobalSlots - 1);
      if (++probes >= kGlobalSlots) {
        atomicExch(overflow, 1);
        break;
      }
    }
  }
}

// cols[k] |= index of vals[k] in the sorted table << 24
__global__ __launch_bounds__(256) void pack_columns_kernel(double const *__restrict__ vals, int64_t nnz,
                                                           double const *__restrict__ table, int size,
                                                           uint32_t *__restrict__ cols) {
  __shared__ double t[kValueTableSize];
  for (int i = threadIdx.x; i < size; i += blockDim.x) t[i] = table[i];
  __syncthreads();
  int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  double const v = vals[k];
  int lo = 0, hi = size - 1;
  while (lo < hi) {
    int const mid = (lo + hi) >> 1;
    if (t[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  cols[k] = (cols[k] & kPackedColumnMask) | (static_cast<uint32_t>(lo) << 24);
}

// vidx[k] = index of vals[k] in the sorted table (wide value table)
__global__ __launch_bounds__(256) void pack_wide_kernel(double const *__restrict__ vals, int64_t nnz,
                                                        double const *__restrict__ table, int size,
                                                        uint16_t *__restrict__ vidx) {
  __shared__ double t[kWideTableSize];
  for (int i = threadIdx.x; i < size; i += blockDim.x) t[i] = table[i];
  __syncthreads();
  int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  double const v = vals[k];
  int lo = 0, hi = size - 1;
  while (lo < hi) {
    int const mid = (lo + hi) >> 1;
    if (t[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  vidx[k] = static_cast<uint16_t>(lo);
}

// Replaces the value arrays of a real CSR by a value table: packed into the
// column words when it has at most 256 distinct values and fewer than 2^24
// rows, else a wide table with 16-bit indices when it has at most
// kWideTableSize.  vals is the FP64 value array (kept by the caller until
// this returns).
bool try_pack_values(Csr &c, double const *vals) {
  if (c.nnz == 0) return true;
  DeviceArray<unsigned long long> table;
  DeviceArray<int> overflow;
  if (!table.allocate(kGlobalSlots) || !overflow.allocate(1)) return false;
  LSAMD_HIP(hipMemsetAsync(table.ptr, 0xFF, sizeof(unsigned long long) * kGlobalSlots, stream()));
  LSAMD_HIP(hipMemsetAsync(overflow.ptr, 0, sizeof(int), stream()));
  distinct_values_kernel<<<2048, 256, 0, stream()>>>(vals, c.nnz, table.ptr, overflow.ptr);
  LSAMD_HIP(hipGetLastError());
  std::vector<unsigned long long> h(kGlobalSlots);
  int ovf = 0;
  LSAMD_HIP(hipMemcpyAsync(h.data(), table.ptr, sizeof(unsigned long long) * kGlobalSlots, hipMemcpyDeviceToHost,
                           stream()));
  LSAMD_HIP(hipMemcpyAsync(&ovf, overflow.ptr, sizeof(int), hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  if (ovf) return true;
  std::vector<double> values;
  for (unsigned long long k : h)
    if (k != kEmptySlot) {
      double v;
      std::memcpy(&v, &k, sizeof(v));
      values.push_back(v);
    }
  bool const fits_word = values.size() <= static_cast<size_t>(kValueTableSize) && c.rows < (int64_t{1} << 24);
  if (values.empty() || (!fits_word && values.size() > static_cast<size_t>(kWideTableSize))) return true;
  std::sort(values.begin(), values.end());
  std::vector<float> values32(values.size());
  for (size_t i = 0; i < values.size(); ++i) values32[i] = static_cast<float>(values[i]);
  if (!c.table64.upload(values.data(), values.size()) || !c.table32.upload(values32.data(), values32.size()))
    return false;
  int const size = static_cast<int>(values.size());
  if (fits_word) {
    pack_columns_kernel<<<grid_for(c.nnz, 256), 256, 0, stream()>>>(vals, c.nnz, c.table64.ptr, size, c.cols.ptr);
  } else {
    if (!c.vidx.allocate(static_cast<size_t>(c.nnz))) return false;
    pack_wide_kernel<<<grid_for(c.nnz, 256), 256, 0, stream()>>>(vals, c.nnz, c.table64.ptr, size, c.vidx.ptr);
  }
  LSAMD_HIP(hipGetLastError());
  LSAMD_HIP(hipStreamSynchronize(stream()));
  c.packed = fits_word;
  c.wide = !fits_word;
  c.table_size = size;
  return true;
}

// The product schedule when neither the caller nor the autotuner chose one:
// lanes per row so that each lane sums about four entries of a mean row, and
// for a million rows or more the die-aware mapping with one workgroup per
// tile, the autotuner's choice for such matrices (16 lanes and that mapping
// on 6x6, against 8 lanes on the plain grid otherwise).
void default_schedule(OperatorCore &op) {
  Csr const &c = op.csr;
  double const mean = c.rows > 0 ? static_cast<double>(c.nnz) / static_cast<double>(c.rows) : 0.0;
  int lanes = 2;
  while (lanes < 64 && 8.0 * lanes <= mean) lanes *= 2;
  op.lanes_f64 = lanes;
  op.flags_f64 = c.rows >= (int64_t{1} << 20) ? 6 : 0;
  op.lanes_f32 = lanes;
  op.flags_f32 = 0;
}

bool operator_build_csr(OperatorCore &op, int precision) {
  if (precision != 32 && precision != 64 && precision != 96 && precision != 128)
    return fail("precision must be 32, 64, 96 (real) or 128 (complex)");
  bool const cplx = precision == 128;
  int const width = cplx ? 2 : 1;
  if (op.number_states >= (int64_t{1} << 32) - 1) return fail("CSR columns are 32-bit: too many states");
  if (op.off.count >= (1 << kKeyLocalBits)) return fail("CSR build: too many off-diagonal terms");
  if (!ensure_err(op) || !operator_prepare_norms(op)) return false;
  int64_t const n = op.number_states;
  Csr &c = op.csr;
  c = Csr{};
  c.rows = n;
  DeviceArray<uint32_t> counts;
  DeviceArray<int64_t> row_ptr;
  if (!counts.allocate(static_cast<size_t>(n > 0 ? n : 1))) return false;
  if (!row_ptr.allocate(static_cast<size_t>(n + 1))) return false;
  if (!c.diag64.allocate(static_cast<size_t>(n > 0 ? n : 1))) return false;
  uint64_t upper = 0;
  if (n > 0) {
    if (op.off.count > 0) {
      count_applicable_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(op.off.view(), n, op.index.reps,
                                                                     counts.ptr);
      LSAMD_HIP(hipGetLastError());
    } else {
      LSAMD_HIP(hipMemsetAsync(counts.ptr, 0, sizeof(uint32_t) * n, stream()));
    }
    if (!exclusive_scan(counts.ptr, n, reinterpret_cast<uint64_t *>(row_ptr.ptr), upper)) return false;
  }
  int64_t const up = static_cast<int64_t>(upper);
  LSAMD_HIP(hipMemcpyAsync(row_ptr.ptr + n, &up, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
  DeviceArray<uint32_t> cols;
  DeviceArray<double> vals;
  if (!cols.allocate(static_cast<size_t>(up > 0 ? up : 1))) return false;
  if (!vals.allocate(static_cast<size_t>(width * (up > 0 ? up : 1)))) return false;
  OpView const v = make_view(op);
  GroupView const g = group_view_or_empty(op);
  size_t const smem = terms_lds_bytes(op.off.count);
  timer_begin(kTimerCsrBuild);
  if (n > 0) {
    bool ok = dispatch_op(op, [&]<typename W, int D, int SI, int KIND>() -> bool {
      if (cplx)
        csr_fill_kernel<W, D, SI, KIND, true><<<grid_for(n, 256), 256, smem, stream()>>>(
            g, v, row_ptr.ptr, cols.ptr, vals.ptr, c.diag64.ptr, counts.ptr, op.d_err.ptr,
            op.d_err_state.ptr);
      else
        csr_fill_kernel<W, D, SI, KIND, false><<<grid_for(n, 256), 256, smem, stream()>>>(
            g, v, row_ptr.ptr, cols.ptr, vals.ptr, c.diag64.ptr, counts.ptr, op.d_err.ptr,
            op.d_err_state.ptr);
      LSAMD_HIP(hipGetLastError());
      return true;
    });
    if (!ok) return false;
  }
  timer_end(kTimerCsrBuild);
  if (!check_error(op)) return false;
  // Drop zero coefficients (the reference skips c == 0,
  // DistributedMatrixVector.chpl:127).
  uint64_t kept = 0;
  DeviceArray<uint64_t> new_ptr;
  if (!new_ptr.allocate(static_cast<size_t>(n + 1))) return false;
  if (n > 0 && !exclusive_scan(counts.ptr, n, new_ptr.ptr, kept)) return false;
  if (static_cast<int64_t>(kept) != up) {
    DeviceArray<uint32_t> cols2;
    DeviceArray<double> vals2;
    if (!cols2.allocate(static_cast<size_t>(kept > 0 ? kept : 1))) return false;
    if (!vals2.allocate(static_cast<size_t>(width * (kept > 0 ? kept : 1)))) return false;
    csr_compact_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(n, row_ptr.ptr, cols.ptr, vals.ptr,
                                                               new_ptr.ptr, cols2.ptr, vals2.ptr, width);
    LSAMD_HIP(hipGetLastError());
    int64_t const kk = static_cast<int64_t>(kept);
    LSAMD_HIP(hipMemcpyAsync(new_ptr.ptr + n, &kk, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
    LSAMD_HIP(hipStreamSynchronize(stream()));
    DeviceArray<int64_t> rp2;
    rp2.ptr = reinterpret_cast<int64_t *>(new_ptr.ptr);
    rp2.count = new_ptr.count;
    new_ptr.ptr = nullptr;
    new_ptr.count = 0;
    row_ptr = std::move(rp2);
    cols = std::move(cols2);
    vals = std::move(vals2);
  }
  c.nnz = static_cast<int64_t>(kept);
  // Source rows (the pairs each alpha_i emits) -> rows of M.
  if (!transpose_to_targets(n, c.nnz, width, row_ptr, cols, vals)) return false;
  c.row_ptr = std::move(row_ptr);
  c.cols = std::move(cols);
  if (cplx) {
    // complex values (re, im) in FP64 and FP32 for the embedded products
    int64_t const m = 2 * c.nnz;
    if (!c.vals32.allocate(static_cast<size_t>(m > 0 ? m : 1))) return false;
    if (!c.diag32.allocate(static_cast<size_t>(n > 0 ? n : 1))) return false;
    if (m > 0) to_f32_kernel<<<grid_for(m, 256), 256, 0, stream()>>>(m, vals.ptr, c.vals32.ptr);
    if (n > 0) to_f32_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(n, c.diag64.ptr, c.diag32.ptr);
    LSAMD_HIP(hipGetLastError());
    c.vals64 = std::move(vals);
    c.has32 = c.has64 = true;
    c.is_complex = true;
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
  if (op.allow_packing && !try_pack_values(c, vals.ptr)) return false;
  if (precision == 32 || precision == 96) {
    if (!c.diag32.allocate(static_cast<size_t>(n > 0 ? n : 1))) return false;
    if (!c.has_table()) {
      if (!c.vals32.allocate(static_cast<size_t>(c.nnz > 0 ? c.nnz : 1))) return false;
      if (c.nnz > 0) to_f32_kernel<<<grid_for(c.nnz, 256), 256, 0, stream()>>>(c.nnz, vals.ptr, c.vals32.ptr);
    }
    if (n > 0) to_f32_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(n, c.diag64.ptr, c.diag32.ptr);
    LSAMD_HIP(hipGetLastError());
    c.has32 = true;
  }
  if (precision == 64 || precision == 96) {
    if (!c.has_table()) c.vals64 = std::move(vals);
    c.has64 = true;
  } else {
    c.diag64.reset();
  }
  if (!op.schedule_set) default_schedule(op);
  LSAMD_HIP(hipStreamSynchronize(stream()));
  return true;
}

// ---------------------------------------------------------------------------
// SpMV kernels: LPR lanes share one row
// ---------------------------------------------------------------------------

template <typename F> bool dispatch_lpr(int lpr, F &&f) {
  switch (lpr) {
  case 1: return f.template operator()<1>();
  case 2: return f.template operator()<2>();
  case 4: return f.template operator()<4>();
  case 8: return f.template operator()<8>();
  case 16: return f.template operator()<16>();
  case 32: return f.template operator()<32>();
  case 64: return f.template operator()<64>();
  default: return fail("lanes per row must be a power of two <= 64");
  }
}

// flags: bit 0 non-temporal stream loads, bit 1 die-aware persistent grid
template <typename F> bool dispatch_flags(int flags, F &&f) {
  switch (flags & 3) {
  case 0: return f.template operator()<false, false>();
  case 1: return f.template operator()<true, false>();
  case 2: return f.template operator()<false, true>();
  default: return f.template operator()<true, true>();
  }
}

// Grid of a product launch: one workgroup per tile; with the die-aware
// mapping (flags bit 1) either the persistent grid or, with bit 2, one
// workgroup per tile rounded up to a multiple of 8, where each workgroup
// takes one tile of its die's contiguous range.
int product_grid(OperatorCore const &op, int64_t n, int lpr, int flags) {
  int const tiles = grid_for(n * lpr, 256);
  if (flags & 2) return (flags & 4) ? (tiles + 7) / 8 * 8 : op.persistent_blocks;
  return tiles;
}

template <typename F> bool dispatch_k(int k, F &&f) {
  switch (k) {
  case 1: return f.template operator()<1>();
  case 2: return f.template operator()<2>();
  case 4: return f.template operator()<4>();
  case 8: return f.template operator()<8>();
  case 16: return f.template operator()<16>();
  default: return fail("block size must be 1, 2, 4, 8 or 16");
  }
}

// The schedule of a block product, by the bytes of x it gathers per stored
// entry: up to 16 B (FP32 blocks of 1, 2, 4; FP64 blocks of 1, 2) the one
// tuned on the FP64 vector (8 B), above that the one tuned on the FP32 block
// of 8 (32 B).  On 6x6 the die-aware full grid that suits the narrow
// gathers takes the FP32 block of 2 from 6.85 to 6.01 ms, while the block
// of 8 is fastest on the plain grid.
struct Schedule {
  int lanes, flags;
};
Schedule block_schedule(OperatorCore const &op, size_t row_bytes) {
  return row_bytes <= 16 ? Schedule{op.lanes_f64, op.flags_f64} : Schedule{op.lanes_f32, op.flags_f32};
}

// The FP64 CSR product over `n` rows starting at row_ptr `rp` and output
// `y` (offsets of the whole arrays; x and the columns stay global).
bool launch_csr_f64(OperatorCore &op, int64_t const *rp, int64_t n, double const *d_x, double *d_y, int accumulate,
                    int has_diag) {
  Csr const &c = op.csr;
  int const grid = product_grid(op, n, op.lanes_f64, op.flags_f64);
  return dispatch_lpr(op.lanes_f64, [&]<int LPR>() -> bool {
    return dispatch_flags(op.flags_f64, [&]<bool NT, bool DIE>() -> bool {
      if (c.wide)
        csr_f64_wide_kernel<LPR, NT, DIE><<<grid, 256, sizeof(double) * c.table_size, stream()>>>(
            n, rp, c.cols.ptr, c.vidx.ptr, c.table64.ptr, c.table_size, c.diag64.ptr, d_x, d_y, accumulate, has_diag);
      else if (c.packed)
        csr_f64_kernel<LPR, NT, DIE, true><<<grid, 256, 0, stream()>>>(
            n, rp, c.cols.ptr, nullptr, c.table64.ptr, c.table_size, c.diag64.ptr, d_x, d_y, accumulate, has_diag);
      else
        csr_f64_kernel<LPR, NT, DIE, false><<<grid, 256, 0, stream()>>>(
            n, rp, c.cols.ptr, c.vals64.ptr, nullptr, 0, c.diag64.ptr, d_x, d_y, accumulate, has_diag);
      LSAMD_HIP(hipGetLastError());
      return true;
    });
  });
}

bool operator_csr_f64(OperatorCore &op, double const *d_x, double *d_y, bool accumulate) {
  Csr const &c = op.csr;
  if (c.is_complex) return fail("the CSR holds complex values: use a sector batch for its products");
  if (!c.has64) return fail("FP64 CSR has not been built");
  int64_t const n = c.rows;
  if (n == 0) return true;
  timer_begin(kTimerCsrF64);
  bool const ok = launch_csr_f64(op, c.row_ptr.ptr, n, d_x, d_y, accumulate ? 1 : 0, op.diag.count > 0 ? 1 : 0);
  timer_end(kTimerCsrF64);
  return ok;
}

__global__ __launch_bounds__(256) void diag_times_kernel(int64_t n, double const *__restrict__ diag,
                                                         double const *__restrict__ x, double *__restrict__ y) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) y[i] = diag[i] * x[i];
}

bool operator_csr_f64_diag(OperatorCore &op, double const *d_x, double *d_y) {
  Csr const &c = op.csr;
  if (c.is_complex || !c.has64) return fail("FP64 CSR has not been built");
  if (c.rows == 0) return true;
  diag_times_kernel<<<grid_for(c.rows, 256), 256, 0, stream()>>>(c.rows, c.diag64.ptr, d_x, d_y);
  LSAMD_HIP(hipGetLastError());
  return true;
}

bool operator_csr_f64_rows(OperatorCore &op, double const *d_x, double *d_y, int64_t r0, int64_t r1) {
  Csr const &c = op.csr;
  if (c.is_complex || !c.has64) return fail("FP64 CSR has not been built");
  if (r0 < 0 || r1 > c.rows || r0 > r1) return fail("row range outside the CSR");
  if (r1 == r0) return true;
  return launch_csr_f64(op, c.row_ptr.ptr + r0, r1 - r0, d_x, d_y + r0, 1, 0);
}

bool operator_csr_f32_block(OperatorCore &op, int k, float const *d_x, float *d_y) {
  Csr const &c = op.csr;
  if (c.is_complex) return fail("the CSR holds complex values: use a sector batch for its products");
  if (!c.has32) return fail("FP32 CSR has not been built");
  int64_t const n = c.rows;
  if (n == 0) return true;
  Schedule const s = block_schedule(op, sizeof(float) * static_cast<size_t>(k));
  timer_begin(kTimerCsrF32);
  bool ok = dispatch_k(k, [&]<int K>() -> bool {
    return dispatch_lpr(s.lanes, [&]<int LPR>() -> bool {
      int const grid = product_grid(op, n, LPR, s.flags);
      return dispatch_flags(s.flags, [&]<bool NT, bool DIE>() -> bool {
        if (c.wide)
          csr_block_wide_kernel<K, LPR, float, float, NT, DIE><<<grid, 256, sizeof(float) * c.table_size, stream()>>>(
              n, c.row_ptr.ptr, c.cols.ptr, c.vidx.ptr, c.table32.ptr, c.table_size, c.diag32.ptr, d_x, d_y);
        else if (c.packed)
          csr_block_kernel<K, LPR, float, float, NT, DIE, true><<<grid, 256, 0, stream()>>>(
              n, c.row_ptr.ptr, c.cols.ptr, nullptr, c.table32.ptr, c.table_size, c.diag32.ptr, d_x, d_y);
        else
          csr_block_kernel<K, LPR, float, float, NT, DIE, false><<<grid, 256, 0, stream()>>>(
              n, c.row_ptr.ptr, c.cols.ptr, c.vals32.ptr, nullptr, 0, c.diag32.ptr, d_x, d_y);
        LSAMD_HIP(hipGetLastError());
        return true;
      });
    });
  });
  timer_end(kTimerCsrF32);
  return ok;
}

bool operator_csr_f64_block(OperatorCore &op, int k, double const *d_x, double *d_y) {
  Csr const &c = op.csr;
  if (c.is_complex) return fail("the CSR holds complex values: use a sector batch for its products");
  if (!c.has64) return fail("FP64 CSR has not been built");
  int64_t const n = c.rows;
  if (n == 0) return true;
  Schedule const s = block_schedule(op, sizeof(double) * static_cast<size_t>(k));
  timer_begin(kTimerCsrF64Block);
  bool ok = dispatch_k(k, [&]<int K>() -> bool {
    return dispatch_lpr(s.lanes, [&]<int LPR>() -> bool {
      int const grid = product_grid(op, n, LPR, s.flags);
      return dispatch_flags(s.flags, [&]<bool NT, bool DIE>() -> bool {
        if (c.wide)
          csr_block_wide_kernel<K, LPR, double, double, NT, DIE><<<grid, 256, sizeof(double) * c.table_size,
                                                                   stream()>>>(
              n, c.row_ptr.ptr, c.cols.ptr, c.vidx.ptr, c.table64.ptr, c.table_size, c.diag64.ptr, d_x, d_y);
        else if (c.packed)
          csr_block_kernel<K, LPR, double, double, NT, DIE, true><<<grid, 256, 0, stream()>>>(
              n, c.row_ptr.ptr, c.cols.ptr, nullptr, c.table64.ptr, c.table_size, c.diag64.ptr, d_x, d_y);
        else
          csr_block_kernel<K, LPR, double, double, NT, DIE, false><<<grid, 256, 0, stream()>>>(
              n, c.row_ptr.ptr, c.cols.ptr, c.vals64.ptr, nullptr, 0, c.diag64.ptr, d_x, d_y);
        LSAMD_HIP(hipGetLastError());
        return true;
      });
    });
  });
  timer_end(kTimerCsrF64Block);
  return ok;
}

// Times every product schedule (lanes per row x stream-load kind x grid
// mapping) on the cached matrix and keeps the fastest, separately for the
// FP64 kernel and the FP32 block-of-8 kernel: the schedules of the narrow
// and the wide gathers (block_schedule).
// plain, non-temporal, die-aware persistent (with / without non-temporal),
// die-aware over the full grid (with / without non-temporal)
constexpr int kScheduleFlags[] = {0, 1, 2, 3, 6, 7};

int operator_autotune(OperatorCore &op) {
  Csr const &c = op.csr;
  int64_t const n = c.rows;
  if (n == 0 || c.is_complex) return op.lanes_f64;
  op.schedule_set = true;
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    // four workgroups per CU; with the wide table as many as fit beside their
    // LDS copies of it (up to eight), so each CU stages the table a few times
    int per_cu = 4;
    if (c.wide)
      per_cu = static_cast<int>(std::min<size_t>(
          8, std::max<size_t>(1, (160u << 10) / (sizeof(double) * static_cast<size_t>(c.table_size)))));
    op.persistent_blocks = std::max(8, (prop.multiProcessorCount * per_cu / 8) * 8);
  }
  int const cand[] = {2, 4, 8, 16, 32, 64};
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto time_it = [&](auto &&launch) -> float {
    launch();
    (void)hipEventRecord(e0, stream());
    for (int r = 0; r < 3; ++r) launch();
    (void)hipEventRecord(e1, stream());
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms;
  };
  if (c.has64) {
    DeviceArray<double> x, y;
    if (x.allocate(n) && y.allocate(n)) {
      (void)hipMemsetAsync(x.ptr, 0, sizeof(double) * n, stream());
      float best = 1e30f;
      int best_l = op.lanes_f64, best_f = op.flags_f64;
      for (int f : kScheduleFlags)
        for (int l : cand) {
          op.lanes_f64 = l;
          op.flags_f64 = f;
          float const ms = time_it([&] { (void)operator_csr_f64(op, x.ptr, y.ptr, false); });
          if (ms < best) {
            best = ms;
            best_l = l;
            best_f = f;
          }
        }
      op.lanes_f64 = best_l;
      op.flags_f64 = best_f;
    }
  }
  if (c.has32) {
    DeviceArray<float> x, y;
    if (x.allocate(static_cast<size_t>(n) * 8) && y.allocate(static_cast<size_t>(n) * 8)) {
      (void)hipMemsetAsync(x.ptr, 0, sizeof(float) * n * 8, stream());
      float best = 1e30f;
      int best_l = op.lanes_f32, best_f = op.flags_f32;
      for (int f : kScheduleFlags)
        for (int l : cand) {
          op.lanes_f32 = l;
          op.flags_f32 = f;
          float const ms = time_it([&] { (void)operator_csr_f32_block(op, 8, x.ptr, y.ptr); });
          if (ms < best) {
            best = ms;
            best_l = l;
            best_f = f;
          }
        }
      op.lanes_f32 = best_l;
      op.flags_f32 = best_f;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(stream());
  return op.lanes_f64;
}

} // namespace lsamd
