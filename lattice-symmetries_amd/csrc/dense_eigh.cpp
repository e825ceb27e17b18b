// dense_eigh.cpp -- eigen-decomposition of the small dense symmetric
// matrices the Krylov solvers project onto (host side).
//
// Householder reduction to tridiagonal form with the transformations
// accumulated, then the implicit symmetric QR iteration with Wilkinson
// shifts on the tridiagonal matrix (Golub & Van Loan, Matrix Computations,
// 4th ed., algorithms 8.3.1 and 8.3.2), with eigenvectors carried along.
// About 8 n^3 flops, against roughly 10 sweeps x 6 n^3 for cyclic Jacobi.
//
// Complex Hermitian matrices (the projected matrices of complex sectors) are
// reduced the same way with complex Householder reflections (H = I - beta v
// v^H, H x = -e^{i arg x_0} |x| e_1); a diagonal phase D makes the
// tridiagonal real, the same implicit QR follows, and the eigenvectors are
// Q D Z.
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <vector>

#include "operator_internal.hpp"

namespace lsamd {

namespace {

// A (n x n, row-major, symmetric) -> tridiagonal (d, e) with A = Q T Q^T.
void tridiagonalise(std::vector<double> &A, int n, std::vector<double> &d, std::vector<double> &e,
                    std::vector<double> &Q) {
  auto a = [&](int i, int j) -> double & { return A[static_cast<size_t>(i) * n + j]; };
  Q.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) Q[static_cast<size_t>(i) * n + i] = 1.0;
  std::vector<double> v(n), p(n), q(n), qv(n);
  for (int k = 0; k + 2 < n; ++k) {
    int const len = n - k - 1;
    double norm2 = 0.0;
    for (int i = 0; i < len; ++i) norm2 += a(k + 1 + i, k) * a(k + 1 + i, k);
    double const x0 = a(k + 1, k);
    double const tail = norm2 - x0 * x0;
    if (tail == 0.0) continue; // column already reduced
    double const norm = std::sqrt(norm2);
    double const alpha = x0 > 0.0 ? -norm : norm;
    for (int i = 0; i < len; ++i) v[i] = a(k + 1 + i, k);
    v[0] -= alpha;
    double vn2 = 0.0;
    for (int i = 0; i < len; ++i) vn2 += v[i] * v[i];
    double const beta = 2.0 / vn2;
    // p = beta S v over the trailing block S = A[k+1:, k+1:]
    for (int i = 0; i < len; ++i) {
      double s = 0.0;
      for (int j = 0; j < len; ++j) s += a(k + 1 + i, k + 1 + j) * v[j];
      p[i] = beta * s;
    }
    double vp = 0.0;
    for (int i = 0; i < len; ++i) vp += v[i] * p[i];
    double const K = 0.5 * beta * vp;
    for (int i = 0; i < len; ++i) q[i] = p[i] - K * v[i];
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j) a(k + 1 + i, k + 1 + j) -= v[i] * q[j] + q[i] * v[j];
    a(k + 1, k) = a(k, k + 1) = alpha;
    for (int i = 1; i < len; ++i) a(k + 1 + i, k) = a(k, k + 1 + i) = 0.0;
    // Q <- Q H on columns k+1 ..
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int j = 0; j < len; ++j) s += Q[static_cast<size_t>(r) * n + k + 1 + j] * v[j];
      qv[r] = beta * s;
    }
    for (int r = 0; r < n; ++r)
      for (int j = 0; j < len; ++j) Q[static_cast<size_t>(r) * n + k + 1 + j] -= qv[r] * v[j];
  }
  d.resize(n);
  e.assign(std::max(n - 1, 0), 0.0);
  for (int i = 0; i < n; ++i) d[i] = a(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
}

// One implicit QR step with a Wilkinson shift on the unreduced block [lo, hi]
// of the tridiagonal (d, e), applied to the columns of Z.
void qr_step(std::vector<double> &d, std::vector<double> &e, int lo, int hi, std::vector<double> &Z, int n) {
  double const a = d[hi - 1], b = e[hi - 1], c = d[hi];
  double const half = 0.5 * (a - c);
  double mu;
  if (half == 0.0) {
    mu = c - std::fabs(b);
  } else {
    mu = c - b * b / (half + std::copysign(std::hypot(half, b), half));
  }
  double x = d[lo] - mu, z = e[lo];
  for (int k = lo; k < hi; ++k) {
    double const r = std::hypot(x, z);
    double cs = 1.0, sn = 0.0;
    if (r != 0.0) {
      cs = x / r;
      sn = z / r;
    }
    if (k > lo) e[k - 1] = r;
    double const dk = d[k], dk1 = d[k + 1], ek = e[k];
    d[k] = cs * cs * dk + 2.0 * cs * sn * ek + sn * sn * dk1;
    d[k + 1] = sn * sn * dk - 2.0 * cs * sn * ek + cs * cs * dk1;
    e[k] = cs * sn * (dk1 - dk) + (cs * cs - sn * sn) * ek;
    if (k + 1 < hi) {
      z = sn * e[k + 1];
      e[k + 1] *= cs;
      x = e[k];
    }
    for (int r2 = 0; r2 < n; ++r2) {
      double &zk = Z[static_cast<size_t>(r2) * n + k];
      double &zk1 = Z[static_cast<size_t>(r2) * n + k + 1];
      double const u = zk, w = zk1;
      zk = cs * u + sn * w;
      zk1 = -sn * u + cs * w;
    }
  }
}

} // namespace

bool symmetric_eigh(std::vector<double> A, int n, std::vector<double> &w, std::vector<double> &Z) {
  if (n <= 0) {
    w.clear();
    Z.clear();
    return true;
  }
  std::vector<double> d, e;
  tridiagonalise(A, n, d, e, Z);
  double const eps = std::numeric_limits<double>::epsilon();
  int hi = n - 1;
  int iterations = 0;
  while (hi > 0) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
    if (++iterations > 60 * n) return false;
    qr_step(d, e, lo, hi, Z, n);
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int i, int j) { return d[i] < d[j]; });
  w.resize(n);
  std::vector<double> Zs(static_cast<size_t>(n) * n);
  for (int c = 0; c < n; ++c) {
    w[c] = d[order[c]];
    for (int r = 0; r < n; ++r) Zs[static_cast<size_t>(r) * n + c] = Z[static_cast<size_t>(r) * n + order[c]];
  }
  Z.swap(Zs);
  return true;
}

bool hermitian_eigh(std::vector<std::complex<double>> A, int n, std::vector<double> &w,
                    std::vector<std::complex<double>> &Z) {
  using C = std::complex<double>;
  if (n <= 0) {
    w.clear();
    Z.clear();
    return true;
  }
  auto a = [&](int i, int j) -> C & { return A[static_cast<size_t>(i) * n + j]; };
  // Householder reduction A = Q T Q^H, T Hermitian tridiagonal
  std::vector<C> Q(static_cast<size_t>(n) * n, C(0.0));
  for (int i = 0; i < n; ++i) Q[static_cast<size_t>(i) * n + i] = 1.0;
  std::vector<C> v(n), p(n), q(n), qv(n);
  for (int k = 0; k + 2 < n; ++k) {
    int const len = n - k - 1;
    double norm2 = 0.0;
    for (int i = 0; i < len; ++i) norm2 += std::norm(a(k + 1 + i, k));
    C const x0 = a(k + 1, k);
    if (norm2 - std::norm(x0) == 0.0) continue; // column already reduced
    double const norm = std::sqrt(norm2), ax0 = std::abs(x0);
    C const phase = ax0 > 0.0 ? x0 / ax0 : C(1.0);
    C const alpha = -phase * norm; // H x = alpha e_1
    for (int i = 0; i < len; ++i) v[i] = a(k + 1 + i, k);
    v[0] -= alpha;
    double vn2 = 0.0;
    for (int i = 0; i < len; ++i) vn2 += std::norm(v[i]);
    double const beta = 2.0 / vn2;
    // S <- H S H on the trailing block S = A[k+1:, k+1:], H = I - beta v v^H
    for (int i = 0; i < len; ++i) {
      C s = 0.0;
      for (int j = 0; j < len; ++j) s += a(k + 1 + i, k + 1 + j) * v[j];
      p[i] = beta * s;
    }
    C vp = 0.0;
    for (int i = 0; i < len; ++i) vp += std::conj(v[i]) * p[i];
    double const K = 0.5 * beta * vp.real();
    for (int i = 0; i < len; ++i) q[i] = p[i] - K * v[i];
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j)
        a(k + 1 + i, k + 1 + j) -= v[i] * std::conj(q[j]) + q[i] * std::conj(v[j]);
    a(k + 1, k) = alpha;
    a(k, k + 1) = std::conj(alpha);
    for (int i = 1; i < len; ++i) a(k + 1 + i, k) = a(k, k + 1 + i) = 0.0;
    // Q <- Q H on columns k+1 ..
    for (int r = 0; r < n; ++r) {
      C s = 0.0;
      for (int j = 0; j < len; ++j) s += Q[static_cast<size_t>(r) * n + k + 1 + j] * v[j];
      qv[r] = beta * s;
    }
    for (int r = 0; r < n; ++r)
      for (int j = 0; j < len; ++j) Q[static_cast<size_t>(r) * n + k + 1 + j] -= qv[r] * std::conj(v[j]);
  }
  // D^H T D is real with D = diag(ph), ph_{i+1} = ph_i t_i / |t_i| for the
  // subdiagonal t_i; then the real implicit QR on (d, |t|)
  std::vector<double> d(n), e(std::max(n - 1, 0));
  std::vector<C> ph(n, C(1.0));
  for (int i = 0; i < n; ++i) d[i] = a(i, i).real();
  for (int i = 0; i + 1 < n; ++i) {
    C const t = a(i + 1, i);
    double const at = std::abs(t);
    e[i] = at;
    ph[i + 1] = at > 0.0 ? ph[i] * (t / at) : ph[i];
  }
  std::vector<double> Zr(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) Zr[static_cast<size_t>(i) * n + i] = 1.0;
  double const eps = std::numeric_limits<double>::epsilon();
  int hi = n - 1;
  int iterations = 0;
  while (hi > 0) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
    if (++iterations > 60 * n) return false;
    qr_step(d, e, lo, hi, Zr, n);
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int i, int j) { return d[i] < d[j]; });
  w.resize(n);
  // eigenvectors Q D Zr, columns in ascending order
  std::vector<C> QD(static_cast<size_t>(n) * n);
  for (int r = 0; r < n; ++r)
    for (int j = 0; j < n; ++j) QD[static_cast<size_t>(r) * n + j] = Q[static_cast<size_t>(r) * n + j] * ph[j];
  Z.assign(static_cast<size_t>(n) * n, C(0.0));
  for (int c = 0; c < n; ++c) {
    int const oc = order[c];
    w[c] = d[oc];
    for (int r = 0; r < n; ++r) {
      C s = 0.0;
      for (int j = 0; j < n; ++j) s += QD[static_cast<size_t>(r) * n + j] * Zr[static_cast<size_t>(j) * n + oc];
      Z[static_cast<size_t>(r) * n + c] = s;
    }
  }
  return true;
}

} // namespace lsamd

extern "C" int ls_amd_dense_eigh_hermitian(int n, double const *A, double *w, double *Z) {
  using namespace lsamd;
  if (n < 0 || (n > 0 && (A == nullptr || w == nullptr)))
    return fail("ls_amd_dense_eigh_hermitian: bad arguments") ? 0 : -1;
  size_t const nn = static_cast<size_t>(n) * n;
  std::vector<std::complex<double>> a(nn), zz;
  for (size_t i = 0; i < nn; ++i) a[i] = {A[2 * i], A[2 * i + 1]};
  std::vector<double> ww;
  if (!hermitian_eigh(std::move(a), n, ww, zz))
    return fail("ls_amd_dense_eigh_hermitian: QR iteration did not converge") ? 0 : -1;
  std::copy(ww.begin(), ww.end(), w);
  if (Z != nullptr)
    for (size_t i = 0; i < nn; ++i) {
      Z[2 * i] = zz[i].real();
      Z[2 * i + 1] = zz[i].imag();
    }
  return 0;
}

extern "C" int ls_amd_dense_eigh(int n, double const *A, double *w, double *Z) {
  using namespace lsamd;
  if (n < 0 || (n > 0 && (A == nullptr || w == nullptr))) return fail("ls_amd_dense_eigh: bad arguments") ? 0 : -1;
  std::vector<double> a(A, A + static_cast<size_t>(n) * n), ww, zz;
  if (!symmetric_eigh(std::move(a), n, ww, zz)) return fail("ls_amd_dense_eigh: QR iteration did not converge") ? 0 : -1;
  std::copy(ww.begin(), ww.end(), w);
  if (Z != nullptr) std::copy(zz.begin(), zz.end(), Z);
  return 0;
}
