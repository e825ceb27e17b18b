// lanczos.hip -- Lanczos eigensolvers on the GPU.
//
// Consumer of the product (the reference drives it from scipy's eigsh,
// python/lattice_symmetries/__init__.py:591-752, or PRIMME,
// chapel/src/Diagonalize.chpl:134-162).  Krylov vectors live in HBM in FP64;
// orthogonalisation (with full re-orthogonalisation, twice) and all inner
// products are FP64 with deterministic two-pass reductions.  The product
// itself is FP64 CSR, FP32 CSR (vectors converted on the fly) or the
// on-the-fly fused kernel, and the FP32 block kernel for block Lanczos.
// Ritz values come from a dense symmetric Jacobi solve on the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <vector>

#include "operator_internal.hpp"
#include "orbit_device.hpp"

namespace lsamd {
namespace {

constexpr int kBlock = 256;
constexpr int kDotBlocks = 1024;
constexpr int kTile = 8;

// partial[j * kDotBlocks + b] = sum over the rows of block b of V_j . w
__global__ __launch_bounds__(kBlock) void multi_dot_kernel(double const *__restrict__ V, int64_t ld, int m,
                                                           double const *__restrict__ w, int64_t n,
                                                           double *__restrict__ partial) {
  __shared__ double s[kTile][kBlock / 64];
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = 0; j0 < m; j0 += kTile) {
    double acc[kTile];
#pragma unroll
    for (int t = 0; t < kTile; ++t) acc[t] = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
      double const wi = w[i];
#pragma unroll
      for (int t = 0; t < kTile; ++t)
        if (j0 + t < m) acc[t] = fma(V[static_cast<int64_t>(j0 + t) * ld + i], wi, acc[t]);
    }
#pragma unroll
    for (int t = 0; t < kTile; ++t) {
      double v = acc[t];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) s[t][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < kTile && j0 + static_cast<int>(threadIdx.x) < m) {
      double v = 0.0;
      for (int q = 0; q < kBlock / 64; ++q) v += s[threadIdx.x][q];
      partial[static_cast<int64_t>(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = v;
    }
    __syncthreads();
  }
}

// out[j] = sum_b partial[j * nb + b], fixed order
__global__ __launch_bounds__(kBlock) void reduce_partials_kernel(double const *__restrict__ partial, int nb,
                                                                 double *__restrict__ out) {
  __shared__ double s[kBlock];
  int const j = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) v += partial[static_cast<int64_t>(j) * nb + b];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = s[0];
}

// w -= sum_j V_j h_j
__global__ __launch_bounds__(kBlock) void gemv_sub_kernel(double const *__restrict__ V, int64_t ld, int m,
                                                          double const *__restrict__ h, double *__restrict__ w,
                                                          int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  double acc = w[i];
  for (int j = 0; j < m; ++j) acc = fma(-V[static_cast<int64_t>(j) * ld + i], h[j], acc);
  w[i] = acc;
}

// out = sum_j V_j s_j
__global__ __launch_bounds__(kBlock) void gemv_kernel(double const *__restrict__ V, int64_t ld, int m,
                                                      double const *__restrict__ s, double *__restrict__ out,
                                                      int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < m; ++j) acc = fma(V[static_cast<int64_t>(j) * ld + i], s[j], acc);
  out[i] = acc;
}

__global__ __launch_bounds__(kBlock) void scale_kernel(double const *__restrict__ x, double a,
                                                       double *__restrict__ y, int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ __launch_bounds__(kBlock) void axpy_kernel(double a, double const *__restrict__ x,
                                                      double *__restrict__ y, int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] = fma(a, x[i], y[i]);
}

__global__ __launch_bounds__(kBlock) void f64_to_f32_kernel(double const *__restrict__ x, float *__restrict__ y,
                                                            int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] = static_cast<float>(x[i]);
}

__global__ __launch_bounds__(kBlock) void f32_to_f64_kernel(float const *__restrict__ x, double *__restrict__ y,
                                                            int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) y[i] = static_cast<double>(x[i]);
}

// out = i v for a complex vector stored as interleaved (re, im) pairs.
__global__ __launch_bounds__(kBlock) void times_i_kernel(double const *__restrict__ v, double *__restrict__ out,
                                                         int64_t pairs) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= pairs) return;
  double const re = v[2 * i], im = v[2 * i + 1];
  out[2 * i] = -im;
  out[2 * i + 1] = re;
}

// Seeded N(0, 1) entries: splitmix64 of (seed, i) and Box-Muller.
__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void random_normal_kernel(uint64_t seed, double *__restrict__ x, int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  uint64_t const a = splitmix(seed ^ splitmix(2 * static_cast<uint64_t>(i)));
  uint64_t const b = splitmix(seed ^ splitmix(2 * static_cast<uint64_t>(i) + 1));
  double const u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
  double const u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  x[i] = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

struct Blas {
  int64_t n;
  DeviceArray<double> partial, small;
  bool init(int64_t n_, int max_m) {
    n = n_;
    return partial.allocate(static_cast<size_t>(max_m + kTile) * kDotBlocks) &&
           small.allocate(static_cast<size_t>(max_m + kTile) + 16 * 16);
  }
  int grid() const { return grid_for(n, kBlock); }
  // h[0..m) = V^T w  (device), copied to host_out when given
  bool multi_dot(double const *V, int64_t ld, int m, double const *w, double *d_h) {
    multi_dot_kernel<<<kDotBlocks, kBlock, 0, stream()>>>(V, ld, m, w, n, partial.ptr);
    reduce_partials_kernel<<<m, kBlock, 0, stream()>>>(partial.ptr, kDotBlocks, d_h);
    LSAMD_HIP(hipGetLastError());
    return true;
  }
  bool dot(double const *a, double const *b, double &out) {
    if (!multi_dot(a, n, 1, b, small.ptr)) return false;
    LSAMD_HIP(hipMemcpyAsync(&out, small.ptr, sizeof(double), hipMemcpyDeviceToHost, stream()));
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
};

// Dense symmetric eigen-decomposition (cyclic Jacobi).  A is m x m row-major;
// on return w holds the eigenvalues ascending and Z the eigenvectors as
// columns.
void jacobi_eigh(std::vector<double> A, int m, std::vector<double> &w, std::vector<double> &Z) {
  Z.assign(static_cast<size_t>(m) * m, 0.0);
  for (int i = 0; i < m; ++i) Z[static_cast<size_t>(i) * m + i] = 1.0;
  auto at = [&](int i, int j) -> double & { return A[static_cast<size_t>(i) * m + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        tot += at(i, j) * at(i, j);
        if (i != j) off += at(i, j) * at(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        double const apq = at(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        double const theta = (at(q, q) - at(p, p)) / (2.0 * apq);
        double const t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double const c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          double const akp = at(k, p), akq = at(k, q);
          at(k, p) = c * akp - s * akq;
          at(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          double const apk = at(p, k), aqk = at(q, k);
          at(p, k) = c * apk - s * aqk;
          at(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          double const zkp = Z[static_cast<size_t>(k) * m + p], zkq = Z[static_cast<size_t>(k) * m + q];
          Z[static_cast<size_t>(k) * m + p] = c * zkp - s * zkq;
          Z[static_cast<size_t>(k) * m + q] = s * zkp + c * zkq;
        }
      }
  }
  std::vector<int> order(m);
  for (int i = 0; i < m; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return at(a, a) < at(b, b); });
  w.resize(m);
  std::vector<double> Zs(static_cast<size_t>(m) * m);
  for (int c = 0; c < m; ++c) {
    w[c] = at(order[c], order[c]);
    for (int r = 0; r < m; ++r) Zs[static_cast<size_t>(r) * m + c] = Z[static_cast<size_t>(r) * m + order[c]];
  }
  Z.swap(Zs);
}

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

// Length of the Krylov vectors for a product mode: the real embedding of a
// complex vector (interleaved re, im) for mode 3.
int64_t vector_length(OperatorCore const &op, int mode) {
  return mode == 3 ? 2 * op.number_states : op.number_states;
}

// y = H x in the chosen form (FP64 vectors in and out).
struct Product {
  OperatorCore *op;
  int mode; // 0 FP64 CSR, 1 FP32 CSR, 2 on-the-fly FP64, 3 on-the-fly FP64 complex
  DeviceArray<float> xf, yf;
  double seconds = 0.0;
  int64_t calls = 0;
  bool init(OperatorCore &o, int m) {
    op = &o;
    mode = m;
    int64_t const n = o.number_states;
    if (mode == 1) return xf.allocate(static_cast<size_t>(n)) && yf.allocate(static_cast<size_t>(n));
    return true;
  }
  bool apply(double const *x, double *y) {
    int64_t const n = op->number_states;
    auto t0 = Clock::now();
    bool ok = true;
    if (mode == 0) {
      ok = operator_csr_f64(*op, x, y, false);
    } else if (mode == 1) {
      f64_to_f32_kernel<<<grid_for(n, kBlock), kBlock, 0, stream()>>>(x, xf.ptr, n);
      ok = operator_csr_f32_block(*op, 1, xf.ptr, yf.ptr);
      f32_to_f64_kernel<<<grid_for(n, kBlock), kBlock, 0, stream()>>>(yf.ptr, y, n);
    } else {
      ok = operator_matvec_otf(*op, x, y, false, mode == 3);
    }
    LSAMD_HIP(hipStreamSynchronize(stream()));
    seconds += seconds_since(t0);
    ++calls;
    return ok;
  }
};

} // namespace

// Single-vector Lanczos.  v0 (device, may be null: seeded random) is the
// start vector.  Returns the k lowest Ritz values in evals and, when d_ritz
// is given, the lowest Ritz vector (normalised).
bool lanczos_run(OperatorCore &op, int mode, int max_steps, int k, double tol, uint64_t seed,
                 double const *d_v0, double *d_ritz, std::vector<double> &evals, int &steps,
                 double &matvec_seconds, int64_t &matvec_calls) {
  int64_t const n = vector_length(op, mode);
  if (n <= 0) return fail("empty basis");
  // Complex vectors (mode 3): the Krylov vectors are kept orthogonal to i V
  // as well, which is Gram-Schmidt in the complex inner product; without it
  // the real embedding's doubled eigenvalues reappear as ghost copies.
  bool const cplx = mode == 3;
  int const m_max = static_cast<int>(std::min<int64_t>(max_steps, cplx ? n / 2 : n));
  DeviceArray<double> V, w, h, JV;
  if (!V.allocate(static_cast<size_t>(n) * (m_max + 1)) || !w.allocate(n) || !h.allocate(m_max + 8)) return false;
  if (cplx && !JV.allocate(static_cast<size_t>(n) * (m_max + 1))) return false;
  Blas blas;
  if (!blas.init(n, m_max + 1)) return false;
  Product prod;
  if (!prod.init(op, mode)) return false;
  int const g = grid_for(n, kBlock);
  if (d_v0 != nullptr) {
    LSAMD_HIP(hipMemcpyAsync(V.ptr, d_v0, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
  } else {
    random_normal_kernel<<<g, kBlock, 0, stream()>>>(seed, V.ptr, n);
  }
  double nrm2 = 0;
  if (!blas.dot(V.ptr, V.ptr, nrm2)) return false;
  scale_kernel<<<g, kBlock, 0, stream()>>>(V.ptr, 1.0 / std::sqrt(nrm2), V.ptr, n);
  if (cplx) times_i_kernel<<<grid_for(n / 2, kBlock), kBlock, 0, stream()>>>(V.ptr, JV.ptr, n / 2);
  std::vector<double> alpha, beta;
  std::vector<double> hh(m_max + 8);
  std::vector<double> ev, Z;
  int m = 0;
  for (int j = 0; j < m_max; ++j) {
    double *vj = V.ptr + static_cast<int64_t>(j) * n;
    if (!prod.apply(vj, w.ptr)) return false;
    double a = 0;
    if (!blas.dot(vj, w.ptr, a)) return false;
    alpha.push_back(a);
    // w -= a v_j + b v_{j-1}, then two passes of full re-orthogonalisation
    axpy_kernel<<<g, kBlock, 0, stream()>>>(-a, vj, w.ptr, n);
    if (j > 0) axpy_kernel<<<g, kBlock, 0, stream()>>>(-beta.back(), vj - n, w.ptr, n);
    for (int pass = 0; pass < 2; ++pass) {
      if (!blas.multi_dot(V.ptr, n, j + 1, w.ptr, h.ptr)) return false;
      gemv_sub_kernel<<<g, kBlock, 0, stream()>>>(V.ptr, n, j + 1, h.ptr, w.ptr, n);
      if (cplx) {
        if (!blas.multi_dot(JV.ptr, n, j + 1, w.ptr, h.ptr)) return false;
        gemv_sub_kernel<<<g, kBlock, 0, stream()>>>(JV.ptr, n, j + 1, h.ptr, w.ptr, n);
      }
    }
    double bb = 0;
    if (!blas.dot(w.ptr, w.ptr, bb)) return false;
    double const b = std::sqrt(bb);
    m = j + 1;
    bool const last = (j + 1 == m_max) || b <= 1e-13 * std::max(1.0, std::fabs(a));
    bool stop = last;
    if (!last && (m % 5 == 0)) {
      std::vector<double> T(static_cast<size_t>(m) * m, 0.0);
      for (int i = 0; i < m; ++i) {
        T[static_cast<size_t>(i) * m + i] = alpha[i];
        if (i + 1 < m) T[static_cast<size_t>(i) * m + i + 1] = T[static_cast<size_t>(i + 1) * m + i] = beta[i];
      }
      jacobi_eigh(T, m, ev, Z);
      int const kk = std::min(k, m);
      double worst = 0;
      for (int q = 0; q < kk; ++q)
        worst = std::max(worst, b * std::fabs(Z[static_cast<size_t>(m - 1) * m + q]) / std::max(1.0, std::fabs(ev[q])));
      if (m >= k && worst < tol) stop = true;
    }
    if (stop) break;
    beta.push_back(b);
    scale_kernel<<<g, kBlock, 0, stream()>>>(w.ptr, 1.0 / b, V.ptr + static_cast<int64_t>(j + 1) * n, n);
    if (cplx)
      times_i_kernel<<<grid_for(n / 2, kBlock), kBlock, 0, stream()>>>(V.ptr + static_cast<int64_t>(j + 1) * n,
                                                                       JV.ptr + static_cast<int64_t>(j + 1) * n,
                                                                       n / 2);
  }
  std::vector<double> T(static_cast<size_t>(m) * m, 0.0);
  for (int i = 0; i < m; ++i) {
    T[static_cast<size_t>(i) * m + i] = alpha[i];
    if (i + 1 < m) T[static_cast<size_t>(i) * m + i + 1] = T[static_cast<size_t>(i + 1) * m + i] = beta[i];
  }
  jacobi_eigh(T, m, ev, Z);
  evals.assign(ev.begin(), ev.begin() + std::min(k, m));
  steps = m;
  if (d_ritz != nullptr) {
    for (int i = 0; i < m; ++i) hh[i] = Z[static_cast<size_t>(i) * m + 0];
    LSAMD_HIP(hipMemcpyAsync(h.ptr, hh.data(), sizeof(double) * m, hipMemcpyHostToDevice, stream()));
    gemv_kernel<<<g, kBlock, 0, stream()>>>(V.ptr, n, m, h.ptr, d_ritz, n);
    double rr = 0;
    if (!blas.dot(d_ritz, d_ritz, rr)) return false;
    scale_kernel<<<g, kBlock, 0, stream()>>>(d_ritz, 1.0 / std::sqrt(rr), d_ritz, n);
  }
  LSAMD_HIP(hipStreamSynchronize(stream()));
  matvec_seconds = prod.seconds;
  matvec_calls = prod.calls;
  return true;
}

// FP64 Rayleigh quotient and residual norm of a normalised vector (the real
// embedding of a complex vector when complex_vectors).
bool rayleigh_residual(OperatorCore &op, double const *d_u, double &theta, double &resid, bool complex_vectors) {
  int64_t const n = complex_vectors ? 2 * op.number_states : op.number_states;
  DeviceArray<double> w;
  Blas blas;
  if (!w.allocate(n) || !blas.init(n, 1)) return false;
  bool const ok = (op.csr.has64 && !complex_vectors && !op.csr.is_complex)
                      ? operator_csr_f64(op, d_u, w.ptr, false)
                      : operator_matvec_otf(op, d_u, w.ptr, false, complex_vectors);
  if (!ok) return false;
  double uu = 0;
  if (!blas.dot(d_u, d_u, uu) || !blas.dot(d_u, w.ptr, theta)) return false;
  theta /= uu;
  axpy_kernel<<<grid_for(n, kBlock), kBlock, 0, stream()>>>(-theta, d_u, w.ptr, n);
  double rr = 0;
  if (!blas.dot(w.ptr, w.ptr, rr)) return false;
  resid = std::sqrt(rr / uu);
  return true;
}

// Seeded test of <u, M v> = <M u, v> through a product y = M x on FP64
// vectors of n_reals doubles (interleaved re, im when complex_values): the
// relative asymmetry |<u, M v> - <M u, v>| / (|u| |M v| + |M u| |v|), which
// is rounding-sized for a symmetric (real) or Hermitian (complex) M and of
// order 1 / sqrt(n) or more otherwise.
bool asymmetry_test(int64_t n_reals, bool complex_values, std::function<bool(double const *, double *)> const &prod,
                    uint64_t seed, double &rel) {
  rel = 0.0;
  if (n_reals <= 0) return true;
  DeviceArray<double> u, v, Mu, Mv, t;
  Blas blas;
  if (!u.allocate(n_reals) || !v.allocate(n_reals) || !Mu.allocate(n_reals) || !Mv.allocate(n_reals) ||
      !t.allocate(n_reals) || !blas.init(n_reals, 1))
    return false;
  int const g = grid_for(n_reals, kBlock);
  random_normal_kernel<<<g, kBlock, 0, stream()>>>(seed, u.ptr, n_reals);
  random_normal_kernel<<<g, kBlock, 0, stream()>>>(seed ^ 0x9e3779b97f4a7c15ull, v.ptr, n_reals);
  LSAMD_HIP(hipGetLastError());
  if (!prod(u.ptr, Mu.ptr) || !prod(v.ptr, Mv.ptr)) return false;
  double a_re = 0, b_re = 0, a_im = 0, b_im = 0, uu = 0, vv = 0, mumu = 0, mvmv = 0;
  if (!blas.dot(u.ptr, Mv.ptr, a_re) || !blas.dot(Mu.ptr, v.ptr, b_re)) return false;
  if (complex_values) {
    // u . (i w) = -Im(u^H w) for interleaved complex vectors
    times_i_kernel<<<grid_for(n_reals / 2, kBlock), kBlock, 0, stream()>>>(Mv.ptr, t.ptr, n_reals / 2);
    if (!blas.dot(u.ptr, t.ptr, a_im)) return false;
    times_i_kernel<<<grid_for(n_reals / 2, kBlock), kBlock, 0, stream()>>>(v.ptr, t.ptr, n_reals / 2);
    if (!blas.dot(Mu.ptr, t.ptr, b_im)) return false;
  }
  if (!blas.dot(u.ptr, u.ptr, uu) || !blas.dot(v.ptr, v.ptr, vv) || !blas.dot(Mu.ptr, Mu.ptr, mumu) ||
      !blas.dot(Mv.ptr, Mv.ptr, mvmv))
    return false;
  double const scale = std::sqrt(uu * mvmv) + std::sqrt(mumu * vv);
  double const diff = std::hypot(a_re - b_re, a_im - b_im);
  rel = scale > 0.0 ? diff / scale : 0.0;
  return true;
}

// Whether the product of `mode` (0 FP64 CSR, 1 FP32 CSR, 2 on-the-fly,
// 3 on-the-fly complex) is symmetric (real modes) or Hermitian (mode 3);
// the answer is kept on the operator.
bool operator_is_hermitian(OperatorCore &op, int mode, bool &hermitian) {
  bool const cplx = mode == 3;
  int &known = op.hermitian[cplx ? 1 : 0];
  if (known < 0) {
    Product p;
    if (!p.init(op, mode)) return false;
    double rel = 0.0;
    auto prod = [&](double const *x, double *y) { return p.apply(x, y); };
    if (!asymmetry_test(vector_length(op, mode), cplx, prod, 0x4e524d4cull, rel)) return false;
    known = rel <= (mode == 1 ? 1e-5 : 1e-9) ? 1 : 0;
  }
  hermitian = known == 1;
  return true;
}

bool require_hermitian(OperatorCore &op, int mode) {
  bool h = false;
  if (!operator_is_hermitian(op, mode, h)) return false;
  if (!h)
    return fail(std::string("the operator is not ") + (mode == 3 ? "Hermitian" : "symmetric") +
                ": the Lanczos solvers need <u, H v> = <H u, v>");
  return true;
}

} // namespace lsamd

using namespace lsamd;

namespace {
void fill_result(ls_amd_lanczos_result *out, std::vector<double> const &evals, int steps, double mv_s,
                 int64_t mv_calls, double total_s) {
  out->steps = steps;
  out->number_eigenvalues = static_cast<int>(std::min<size_t>(evals.size(), 8));
  for (int i = 0; i < 8; ++i) out->eigenvalues[i] = i < static_cast<int>(evals.size()) ? evals[i] : 0.0;
  out->matvec_seconds = mv_s;
  out->matvec_calls = mv_calls;
  out->total_seconds = total_s;
}
} // namespace

extern "C" int ls_amd_lanczos(ls_amd_operator *op, int mode, int max_steps, int k, double tol, uint64_t seed,
                              int refine_steps, double *d_ritz_out, ls_amd_lanczos_result *out) {
  if (op == nullptr || out == nullptr) return fail("null argument") ? 0 : -1;
  if (mode < 0 || mode > 3)
    return fail("mode must be 0 (FP64 CSR), 1 (FP32 CSR), 2 (on-the-fly) or 3 (on-the-fly, complex)") ? 0 : -1;
  if ((mode == 0 && !op->core.csr.has64) || (mode == 1 && !op->core.csr.has32))
    return fail("build the CSR cache in the requested precision first") ? 0 : -1;
  if ((mode == 0 || mode == 1) && op->core.csr.is_complex)
    return fail("the CSR holds complex values: use mode 3") ? 0 : -1;
  if (!require_hermitian(op->core, mode)) return -1;
  auto t0 = Clock::now();
  int64_t const n = vector_length(op->core, mode);
  DeviceArray<double> u;
  if (!u.allocate(static_cast<size_t>(n))) return -1;
  std::vector<double> evals;
  int steps = 0;
  double mv = 0;
  int64_t calls = 0;
  if (!lanczos_run(op->core, mode, max_steps, std::max(1, std::min(k, 8)), tol, seed, nullptr, u.ptr, evals, steps,
                   mv, calls))
    return -1;
  out->refine_steps = 0;
  if (refine_steps > 0) {
    // FP64 correction: a short FP64 Krylov run restarted from the Ritz vector
    std::vector<double> ev2;
    int st2 = 0;
    double mv2 = 0;
    int64_t c2 = 0;
    int const form = mode == 3 ? 3 : (op->core.csr.has64 ? 0 : 2);
    if (!lanczos_run(op->core, form, refine_steps, 1, 0.0, seed, u.ptr, u.ptr, ev2, st2, mv2, c2)) return -1;
    evals[0] = ev2[0];
    mv += mv2;
    calls += c2;
    out->refine_steps = st2;
  }
  double theta = 0, resid = 0;
  if (!rayleigh_residual(op->core, u.ptr, theta, resid, mode == 3)) return -1;
  fill_result(out, evals, steps, mv, calls, seconds_since(t0));
  out->rayleigh = theta;
  out->residual_norm = resid;
  if (d_ritz_out != nullptr)
    if (hipMemcpy(d_ritz_out, u.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice) != hipSuccess) return -1;
  return 0;
}

namespace {
// Block Lanczos on one operator through the segmented solver (one segment).
bool operator_block_solve(OperatorCore &core, BlockKrylovOptions const &o, BlockKrylovResult &r,
                          double *d_ritz_block) {
  int64_t const n = core.number_states;
  BlockProduct prod = [&](bool fp32, void const *x, void *y) -> bool {
    if (fp32) return operator_csr_f32_block(core, o.block, static_cast<float const *>(x), static_cast<float *>(y));
    return operator_csr_f64_block(core, o.block, static_cast<double const *>(x), static_cast<double *>(y));
  };
  std::vector<int64_t> const seg{0, n};
  return block_krylov_solve(seg, {}, 1, prod, o, r, d_ritz_block);
}

void fill_info(ls_amd_eig_info *info, BlockKrylovResult const &r) {
  if (info == nullptr) return;
  info->cycles = r.cycles;
  info->fp64_cycles = r.fp64_cycles;
  info->blocks = r.blocks;
  info->converged = r.converged ? 1 : 0;
  info->products = r.products;
  info->product_seconds = r.product_seconds;
  info->total_seconds = r.total_seconds;
  info->max_residual = r.max_residual;
}
} // namespace

// The thick-restart block solve behind ls_amd_block_eigensolve and
// ls_amd_block_eigenvectors: the first n_out Ritz vectors (ascending
// eigenvalues) go to d_out, n_out rows of number_states values.
static int block_eigensolve_into(ls_amd_operator *op, int mode, int block_size, int nev, int max_blocks,
                                 int keep_blocks, int max_cycles, double tol, double switch_tol, uint64_t seed,
                                 double *evals, double *residuals, double *d_out, int n_out, ls_amd_eig_info *info) {
  if (op == nullptr || evals == nullptr) return fail("null argument") ? 0 : -1;
  if (mode != 0 && mode != 1) return fail("mode must be 0 (FP64 block CSR) or 1 (FP32 then FP64)") ? 0 : -1;
  Csr const &c = op->core.csr;
  if (c.is_complex) return fail("the CSR holds complex values: use a sector batch") ? 0 : -1;
  if (!c.has64 || (mode == 1 && !c.has32))
    return fail("build the CSR cache in both precisions first (precision 96)") ? 0 : -1;
  if (!require_hermitian(op->core, 0)) return -1;
  BlockKrylovOptions o;
  o.block = block_size;
  o.nev = nev;
  o.max_blocks = max_blocks;
  o.keep = keep_blocks;
  o.max_cycles = max_cycles;
  o.tol = tol;
  o.switch_tol = switch_tol;
  o.seed = seed;
  o.fp32 = mode == 1;
  int64_t const n = op->core.number_states;
  DeviceArray<double> Yb;
  if (d_out != nullptr && !Yb.allocate(static_cast<size_t>(n) * block_size)) return -1;
  BlockKrylovResult r;
  if (!operator_block_solve(op->core, o, r, d_out != nullptr ? Yb.ptr : nullptr)) return -1;
  std::copy(r.evals.begin(), r.evals.end(), evals);
  if (residuals != nullptr) std::copy(r.residuals.begin(), r.residuals.end(), residuals);
  if (d_out != nullptr) {
    for (int e = 0; e < n_out; ++e)
      if (!block_column(Yb.ptr, block_size, e, d_out + static_cast<size_t>(e) * n, n)) return -1;
    if (hipStreamSynchronize(stream()) != hipSuccess) return fail("stream synchronisation failed") ? 0 : -1;
  }
  fill_info(info, r);
  return 0;
}

extern "C" int ls_amd_block_eigensolve(ls_amd_operator *op, int mode, int block_size, int nev, int max_blocks,
                                       int keep_blocks, int max_cycles, double tol, double switch_tol, uint64_t seed,
                                       double *evals, double *residuals, double *d_ritz_out, ls_amd_eig_info *info) {
  return block_eigensolve_into(op, mode, block_size, nev, max_blocks, keep_blocks, max_cycles, tol, switch_tol, seed,
                               evals, residuals, d_ritz_out, 1, info);
}

extern "C" int ls_amd_block_eigenvectors(ls_amd_operator *op, int mode, int block_size, int nev, int max_blocks,
                                         int keep_blocks, int max_cycles, double tol, double switch_tol,
                                         uint64_t seed, double *evals, double *residuals, double *d_evecs,
                                         ls_amd_eig_info *info) {
  if (d_evecs == nullptr) return fail("null argument") ? 0 : -1;
  return block_eigensolve_into(op, mode, block_size, nev, max_blocks, keep_blocks, max_cycles, tol, switch_tol, seed,
                               evals, residuals, d_evecs, nev, info);
}

extern "C" int ls_amd_block_lanczos(ls_amd_operator *op, int mode, int block_size, int max_blocks, int k, double tol,
                                    uint64_t seed, int refine_steps, double *d_ritz_out, ls_amd_lanczos_result *out) {
  if (op == nullptr || out == nullptr) return fail("null argument") ? 0 : -1;
  if (mode != 0 && mode != 1) return fail("mode must be 0 (FP64 block CSR) or 1 (FP32 block CSR)") ? 0 : -1;
  if (op->core.csr.is_complex) return fail("the CSR holds complex values: use a sector batch") ? 0 : -1;
  if ((mode == 0 && !op->core.csr.has64) || (mode == 1 && !op->core.csr.has32))
    return fail("build the CSR cache in the requested precision first") ? 0 : -1;
  if (!op->core.csr.has64) return fail("the FP64 CSR is needed for the final FP64 cycles (precision 96)") ? 0 : -1;
  if (!require_hermitian(op->core, 0)) return -1;
  auto t0 = Clock::now();
  int64_t const n = op->core.number_states;
  BlockKrylovOptions o;
  o.block = block_size;
  o.nev = std::max(1, std::min(k, block_size));
  o.max_blocks = std::max(4, max_blocks);
  o.keep = 2;
  o.max_cycles = 60;
  o.tol = tol;
  o.switch_tol = std::max(tol, 1e-5);
  o.seed = seed;
  o.fp32 = mode == 1;
  DeviceArray<double> Yb, u;
  if (!Yb.allocate(static_cast<size_t>(n) * block_size) || !u.allocate(static_cast<size_t>(n))) return -1;
  BlockKrylovResult r;
  if (!operator_block_solve(op->core, o, r, Yb.ptr)) return -1;
  if (!block_column(Yb.ptr, block_size, 0, u.ptr, n)) return -1;
  std::vector<double> evals = r.evals;
  double mv = r.product_seconds;
  int64_t calls = r.products;
  out->refine_steps = 0;
  if (refine_steps > 0) {
    std::vector<double> ev2;
    int st2 = 0;
    double mv2 = 0;
    int64_t c2 = 0;
    if (!lanczos_run(op->core, 0, refine_steps, 1, 0.0, seed, u.ptr, u.ptr, ev2, st2, mv2, c2)) return -1;
    evals[0] = ev2[0];
    mv += mv2;
    calls += c2;
    out->refine_steps = st2;
  }
  double theta = 0, resid = 0;
  if (!rayleigh_residual(op->core, u.ptr, theta, resid, false)) return -1;
  fill_result(out, evals, r.blocks, mv, calls, seconds_since(t0));
  out->rayleigh = theta;
  out->residual_norm = resid;
  if (d_ritz_out != nullptr)
    if (hipMemcpy(d_ritz_out, u.ptr, sizeof(double) * n, hipMemcpyDeviceToDevice) != hipSuccess) return -1;
  return 0;
}
