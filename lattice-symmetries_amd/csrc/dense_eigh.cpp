// dense_eigh.cpp -- eigen-decomposition of the small dense symmetric
// matrices the Krylov solvers project onto (host side).
//
// Householder reduction to tridiagonal form with the transformations
// accumulated, then the implicit symmetric QR iteration with Wilkinson
// shifts on the tridiagonal matrix (Golub & Van Loan, Matrix Computations,
// 4th ed., algorithms 8.3.1 and 8.3.2), with eigenvectors carried along.
// About 8 n^3 flops, against roughly 10 sweeps x 6 n^3 for cyclic Jacobi.
//
// Complex Hermitian matrices (the projected matrices of complex sectors) go
// through cyclic Jacobi with complex rotations (Golub & Van Loan, 4th ed.,
// 8.5.2, after a diagonal phase that makes the pivot element real).
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
  auto z = [&](int i, int j) -> C & { return Z[static_cast<size_t>(i) * n + j]; };
  Z.assign(static_cast<size_t>(n) * n, C(0.0));
  for (int i = 0; i < n; ++i) {
    z(i, i) = 1.0;
    a(i, i) = a(i, i).real();
  }
  bool converged = false;
  for (int sweep = 0; sweep < 60 && !converged; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; ++p) {
      diag += std::norm(a(p, p));
      for (int q = p + 1; q < n; ++q) off += std::norm(a(p, q));
    }
    if (off <= 1e-32 * (diag + 2.0 * off)) {
      converged = true;
      break;
    }
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double const r = std::abs(a(p, q));
        if (r == 0.0) continue;
        // U = diag(1, e) [[c, s], [-s, c]] on (p, q), e = conj(a_pq) / |a_pq|
        C const e = std::conj(a(p, q)) / r;
        double const tau = (a(q, q).real() - a(p, p).real()) / (2.0 * r);
        double const t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        double const c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) { // A <- A U
          C const kp = a(k, p), kq = a(k, q);
          a(k, p) = c * kp - s * e * kq;
          a(k, q) = s * kp + c * e * kq;
        }
        for (int k = 0; k < n; ++k) { // A <- U^H A
          C const pk = a(p, k), qk = a(q, k);
          a(p, k) = c * pk - s * std::conj(e) * qk;
          a(q, k) = s * pk + c * std::conj(e) * qk;
        }
        a(p, q) = a(q, p) = 0.0;
        a(p, p) = a(p, p).real();
        a(q, q) = a(q, q).real();
        for (int k = 0; k < n; ++k) { // Z <- Z U
          C const kp = z(k, p), kq = z(k, q);
          z(k, p) = c * kp - s * e * kq;
          z(k, q) = s * kp + c * e * kq;
        }
      }
  }
  if (!converged) return false;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int i, int j) { return a(i, i).real() < a(j, j).real(); });
  w.resize(n);
  std::vector<C> Zs(static_cast<size_t>(n) * n);
  for (int c = 0; c < n; ++c) {
    w[c] = a(order[c], order[c]).real();
    for (int r = 0; r < n; ++r) Zs[static_cast<size_t>(r) * n + c] = z(r, order[c]);
  }
  Z.swap(Zs);
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
    return fail("ls_amd_dense_eigh_hermitian: Jacobi sweeps did not converge") ? 0 : -1;
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
