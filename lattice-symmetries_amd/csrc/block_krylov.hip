// block_krylov.hip -- thick-restart block Lanczos on FP64 row-major blocks,
// for one operator or for many independent segments (symmetry sectors) at
// once.
//
// The consumer of the product: the reference drives its product from scipy
// eigsh (python/lattice_symmetries/__init__.py:591-752) or PRIMME
// (chapel/src/Diagonalize.chpl:134-162), one sector at a time.  Here every
// segment runs its own Krylov process but they share each product launch
// and each reduction launch.
//
// Blocks are [rows][B] row-major FP64.  Per block step: one product (FP32
// or FP64, supplied by the caller), full re-orthogonalisation against every
// stored block by segmented Gram reductions (fixed-order two-pass sums) and
// block updates (a second pass only after a large norm drop), CholQR2 per
// segment, and the projected matrix
// filled from the re-orthogonalisation coefficients.  At the end of a
// cycle each segment's projected matrix is diagonalised on a host thread and
// the lowest `keep` Ritz blocks plus the residual block restart the next
// cycle.  The product runs in FP32 until the relative residual falls below
// switch_tol (or stops improving by 10 % per cycle), then the process restarts explicitly from
// the lowest Ritz block with the FP64 product.  Final eigenvalues and
// residuals are FP64 Rayleigh quotients and residual norms of the Ritz
// vectors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <thread>
#include <type_traits>
#include <vector>

#include "operator_internal.hpp"
#include "orbit_device.hpp"

namespace lsamd {
namespace {

constexpr int kThreads = 256;

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// ---------------------------------------------------------------------------
// Segmented FP64 block algebra on row-major [rows][B] blocks.  Rows are cut
// into chunks that never straddle a segment; every reduction is a fixed-order
// two-pass sum, so results do not depend on scheduling.
// ---------------------------------------------------------------------------

constexpr int kChunkRows = 2048;

// A block step re-orthogonalises once, and a second time only when the first
// pass left some direction of the block (a combination of its columns, not
// only a single column) with less than kReorthDrop of the block's largest
// squared column norm: then one pass leaves mostly rounding error (an
// exhausted Krylov space).  Otherwise
// the orthogonality lost in one pass stays near the storage precision times
// 10^3, and the reported eigenvalues and residuals are explicit FP64 Rayleigh
// quotients and residual norms in any case.
constexpr double kReorthDrop = 1e-6;

// partial[(c * nblk + i) * B*B + p*B + q] = sum over rows r of chunk c of
// A_i[r][p] * W[r][q], with A_i = A + i * a_stride, except that the last
// block is `extra` when given (A_{nblk-1} = extra).  Blocks are stored as TA
// / TW (float or double); products and sums are FP64.
template <int B, typename TA, typename TW>
__global__ __launch_bounds__(kThreads) void seg_gram_kernel(TA const *__restrict__ A, int64_t a_stride, int nblk,
                                                            TA const *__restrict__ extra,
                                                            TW const *__restrict__ W,
                                                            int64_t const *__restrict__ chunk_rows,
                                                            double *__restrict__ partial) {
  constexpr int V = B * B;
  __shared__ double s[kThreads / 64][64];
  int const c = blockIdx.x;
  int64_t const r0 = chunk_rows[2 * c], r1 = chunk_rows[2 * c + 1];
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = 0; i < nblk; ++i) {
    TA const *__restrict__ Ai =
        (extra != nullptr && i == nblk - 1) ? extra : A + static_cast<int64_t>(i) * a_stride;
    double acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kThreads) {
      double a[B], w[B];
#pragma unroll
      for (int p = 0; p < B; ++p) {
        a[p] = static_cast<double>(Ai[r * B + p]);
        w[p] = static_cast<double>(W[r * B + p]);
      }
#pragma unroll
      for (int p = 0; p < B; ++p)
#pragma unroll
        for (int q = 0; q < B; ++q) acc[p * B + q] = fma(a[p], w[q], acc[p * B + q]);
    }
    // Transpose-reduce across the wave: after the steps below lane L holds
    // the wave's sum for value index L & (V - 1); then the waves' sums are
    // added in wave order.  (Written out here and in seg_combine_kernel: a
    // helper taking the accumulator array by reference made the compiler
    // keep it in scratch memory.)
#pragma unroll
    for (int h = V / 2; h > 0; h >>= 1) {
      bool const upper = (lane & h) != 0;
#pragma unroll
      for (int v = 0; v < h; ++v) {
        double const send = upper ? acc[v] : acc[v + h];
        double const keep = upper ? acc[v + h] : acc[v];
        acc[v] = keep + __shfl_xor(send, h, 64);
      }
    }
    double tot = acc[0];
    for (int off = V; off < 64; off <<= 1) tot += __shfl_xor(tot, off, 64);
    s[wave][lane] = tot;
    __syncthreads();
    if (threadIdx.x < V) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) sum += s[w][threadIdx.x];
      partial[(static_cast<int64_t>(c) * nblk + i) * V + threadIdx.x] = sum;
    }
    __syncthreads();
  }
}

// out[(s * nblk + i) * V + v] = sum over the chunks of segment s.  The
// kReduceThreads threads form G = kReduceThreads / V groups; group g sums
// chunks c0 + g, c0 + g + G, ... in order, and the G group sums are added in
// group order, so the result does not depend on scheduling.
constexpr int kReduceThreads = 1024;

__global__ __launch_bounds__(kReduceThreads) void seg_reduce_kernel(double const *__restrict__ partial, int nblk,
                                                                    int V, int const *__restrict__ seg_chunk,
                                                                    double *__restrict__ out) {
  __shared__ double part[kReduceThreads];
  int const s = blockIdx.x, i = blockIdx.y, t = threadIdx.x;
  int const G = kReduceThreads / V, v = t % V, g = t / V;
  double sum = 0.0;
  if (g < G)
    for (int c = seg_chunk[s] + g; c < seg_chunk[s + 1]; c += G)
      sum += partial[(static_cast<int64_t>(c) * nblk + i) * V + v];
  part[t] = sum;
  __syncthreads();
  if (t < V) {
    double tot = 0.0;
    for (int h = 0; h < G; ++h) tot += part[h * V + t];
    out[(static_cast<int64_t>(s) * nblk + i) * V + t] = tot;
  }
}

// out[r][q] = beta * out[r][q] + sum_{i < nblk} sum_p A_i[r][p] C_s[i][p][q]
// for the segment s of row r, summed in FP64 and stored as TO.  Each thread
// owns whole rows, so out may alias one of the A_i (when TO is TA).  With
// GRAM, partial[c * B*B + p*B + q] also receives the Gram sum out^T out of
// the chunk's stored rows, with the rows, order and reduction of
// seg_gram_kernel, so it equals that kernel's result bit for bit.
template <int B, typename TO, typename TA, bool GRAM>
__global__ __launch_bounds__(kThreads) void seg_combine_kernel(TO *out, double beta, TA const *A,
                                                               int64_t a_stride, int nblk,
                                                               double const *__restrict__ coef,
                                                               int64_t const *__restrict__ chunk_rows,
                                                               int const *__restrict__ chunk_seg,
                                                               double *__restrict__ partial) {
  constexpr int V = B * B;
  extern __shared__ double sc[];
  __shared__ double sg[GRAM ? kThreads / 64 : 1][64];
  int const c = blockIdx.x;
  int const s = chunk_seg[c];
  int64_t const r0 = chunk_rows[2 * c], r1 = chunk_rows[2 * c + 1];
  double const *__restrict__ cs = coef + static_cast<int64_t>(s) * nblk * V;
  for (int t = threadIdx.x; t < nblk * V; t += kThreads) sc[t] = cs[t];
  __syncthreads();
  double g[GRAM ? V : 1];
#pragma unroll
  for (int v = 0; v < (GRAM ? V : 1); ++v) g[v] = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kThreads) {
    double acc[B];
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] = 0.0;
    for (int i = 0; i < nblk; ++i) {
      double a[B];
#pragma unroll
      for (int p = 0; p < B; ++p) a[p] = static_cast<double>(A[static_cast<int64_t>(i) * a_stride + r * B + p]);
#pragma unroll
      for (int p = 0; p < B; ++p)
#pragma unroll
        for (int q = 0; q < B; ++q) acc[q] = fma(a[p], sc[i * V + p * B + q], acc[q]);
    }
    if (beta != 0.0) {
#pragma unroll
      for (int q = 0; q < B; ++q) acc[q] = fma(beta, static_cast<double>(out[r * B + q]), acc[q]);
    }
#pragma unroll
    for (int q = 0; q < B; ++q) out[r * B + q] = static_cast<TO>(acc[q]);
    if constexpr (GRAM) {
      double o[B];
#pragma unroll
      for (int q = 0; q < B; ++q) o[q] = static_cast<double>(static_cast<TO>(acc[q]));
#pragma unroll
      for (int p = 0; p < B; ++p)
#pragma unroll
        for (int q = 0; q < B; ++q) g[p * B + q] = fma(o[p], o[q], g[p * B + q]);
    }
  }
  if constexpr (GRAM) { // the reduction of seg_gram_kernel
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int h = V / 2; h > 0; h >>= 1) {
      bool const upper = (lane & h) != 0;
#pragma unroll
      for (int v = 0; v < h; ++v) {
        double const send = upper ? g[v] : g[v + h];
        double const keep = upper ? g[v + h] : g[v];
        g[v] = keep + __shfl_xor(send, h, 64);
      }
    }
    double tot = g[0];
    for (int off = V; off < 64; off <<= 1) tot += __shfl_xor(tot, off, 64);
    sg[wave][lane] = tot;
    __syncthreads();
    if (threadIdx.x < V) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) sum += sg[w][threadIdx.x];
      partial[static_cast<int64_t>(c) * V + threadIdx.x] = sum;
    }
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Seeded N(0, 1) entries (splitmix64 of (seed, index), Box-Muller).
template <typename T>
__global__ __launch_bounds__(kThreads) void normal_kernel(uint64_t seed, T *__restrict__ x, int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  uint64_t const a = mix64(seed ^ mix64(2 * static_cast<uint64_t>(i)));
  uint64_t const b = mix64(seed ^ mix64(2 * static_cast<uint64_t>(i) + 1));
  double const u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
  double const u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  x[i] = static_cast<T>(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
}

// Seeded N(0, 1) entries (as normal_kernel, indexed by the element of the
// block) in the columns c >= keep[s] of the rows of segment s; a row holds
// `width` reals, column c of a field entry being real column c (real blocks)
// or c and field_cols + c (complex blocks stored [re | im]).
template <typename T>
__global__ __launch_bounds__(kThreads) void refill_kernel(uint64_t seed, T *__restrict__ x, int width,
                                                          int field_cols, int64_t const *__restrict__ chunk_rows,
                                                          int const *__restrict__ chunk_seg,
                                                          int const *__restrict__ keep) {
  int const c = blockIdx.x;
  int const k = keep[chunk_seg[c]];
  if (k >= field_cols) return;
  int64_t const e0 = chunk_rows[2 * c] * width, e1 = chunk_rows[2 * c + 1] * width;
  for (int64_t i = e0 + threadIdx.x; i < e1; i += kThreads) {
    if (static_cast<int>(i % width) % field_cols < k) continue;
    uint64_t const a = mix64(seed ^ mix64(2 * static_cast<uint64_t>(i)));
    uint64_t const b = mix64(seed ^ mix64(2 * static_cast<uint64_t>(i) + 1));
    double const u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
    double const u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
    x[i] = static_cast<T>(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
  }
}

template <typename F> bool dispatch_block(int b, F &&f) {
  switch (b) {
  case 1: return f.template operator()<1>();
  case 2: return f.template operator()<2>();
  case 4: return f.template operator()<4>();
  case 8: return f.template operator()<8>();
  default: return fail("block size must be 1, 2, 4 or 8");
  }
}

// Chunks of the rows, per segment.  Only the listed segments get chunks
// (numbered 0 .. nseg - 1 in list order); rows of the others are never read
// or written by the block algebra.
struct Segments {
  int nseg = 0;
  int64_t rows = 0;
  int chunks = 0;
  DeviceArray<int64_t> d_chunk_rows; // chunk c covers rows [chunk_rows[2c], chunk_rows[2c + 1])
  DeviceArray<int> d_chunk_seg, d_seg_chunk;
  bool init(std::vector<int64_t> const &seg_rows, std::vector<int> const &listed) {
    nseg = static_cast<int>(listed.size());
    rows = seg_rows.back();
    std::vector<int64_t> cr;
    std::vector<int> cs, sc{0};
    for (int a = 0; a < nseg; ++a) {
      int64_t const r0 = seg_rows[listed[a]], r1 = seg_rows[listed[a] + 1];
      for (int64_t r = r0; r < r1; r += kChunkRows) {
        cr.push_back(r);
        cr.push_back(std::min(r + kChunkRows, r1));
        cs.push_back(a);
      }
      sc.push_back(static_cast<int>(cs.size()));
    }
    chunks = static_cast<int>(cs.size());
    if (cr.empty()) cr.assign(2, 0);
    return d_chunk_rows.upload(cr.data(), cr.size()) && d_chunk_seg.upload(cs.data(), std::max<size_t>(1, cs.size())) &&
           d_seg_chunk.upload(sc.data(), sc.size());
  }
};

// Host-side dense helpers for the per-segment small matrices, over the field
// F of the Krylov process: double, or std::complex<double> for blocks of
// complex columns stored as real rows [re | im].
template <typename F> struct Field;
template <> struct Field<double> {
  static constexpr int kWidth = 1; // reals per entry of a block row
  static double conj(double x) { return x; }
  static double re(double x) { return x; }
};
template <> struct Field<std::complex<double>> {
  static constexpr int kWidth = 2;
  static std::complex<double> conj(std::complex<double> x) { return std::conj(x); }
  static double re(std::complex<double> x) { return x.real(); }
};

// `count` real Bw x Bw Gram matrices X^T Y of blocks with rows [re | im] ->
// the field's Bf x Bf matrices X^H Y (Bw = Bf * width).
template <typename F> void gram_to_field(std::vector<double> const &G, size_t count, int Bf, std::vector<F> &out) {
  size_t const Vf = static_cast<size_t>(Bf) * Bf;
  out.resize(count * Vf);
  if constexpr (Field<F>::kWidth == 1) {
    std::copy(G.begin(), G.begin() + static_cast<std::ptrdiff_t>(count * Vf), out.begin());
  } else {
    int const Bw = 2 * Bf;
    size_t const Vw = static_cast<size_t>(Bw) * Bw;
    for (size_t k = 0; k < count; ++k) {
      double const *g = G.data() + k * Vw;
      F *o = out.data() + k * Vf;
      for (int a = 0; a < Bf; ++a)
        for (int b = 0; b < Bf; ++b)
          o[a * Bf + b] = {g[a * Bw + b] + g[(Bf + a) * Bw + Bf + b], g[a * Bw + Bf + b] - g[(Bf + a) * Bw + b]};
    }
  }
}

// The real Bw x Bw coefficients with which the block update applies the
// field matrix M (row vector [re | im] times [[Re M, Im M], [-Im M, Re M]]).
template <typename F> void field_to_coef(F const *M, int Bf, double *out) {
  if constexpr (Field<F>::kWidth == 1) {
    std::copy(M, M + static_cast<size_t>(Bf) * Bf, out);
  } else {
    int const Bw = 2 * Bf;
    for (int a = 0; a < Bf; ++a)
      for (int b = 0; b < Bf; ++b) {
        double const r = M[a * Bf + b].real(), i = M[a * Bf + b].imag();
        out[a * Bw + b] = r;
        out[a * Bw + Bf + b] = i;
        out[(Bf + a) * Bw + b] = -i;
        out[(Bf + a) * Bw + Bf + b] = r;
      }
  }
}

// Cholesky G = R^H R of a B x B Hermitian positive definite matrix; returns
// false if a pivot is not positive (rank deficiency).  R upper, Rinv = R^-1.
// Fails when a pivot (the squared part of a column outside the span of the
// columns before it) is not above 1e-28 of the largest diagonal entry, or not
// above `floor_abs`.
template <typename F>
bool cholesky_upper(std::vector<F> const &G, int B, std::vector<F> &R, std::vector<F> &Rinv, double floor_abs) {
  std::vector<F> L(static_cast<size_t>(B) * B, F(0.0));
  double scale = 0.0;
  for (int i = 0; i < B; ++i) scale = std::max(scale, Field<F>::re(G[static_cast<size_t>(i) * B + i]));
  for (int j = 0; j < B; ++j) {
    double d = Field<F>::re(G[static_cast<size_t>(j) * B + j]);
    for (int k = 0; k < j; ++k) d -= std::norm(L[static_cast<size_t>(j) * B + k]);
    if (!(d > std::max(1e-28 * std::max(scale, 1e-300), floor_abs))) return false;
    double const ljj = std::sqrt(d);
    L[static_cast<size_t>(j) * B + j] = ljj;
    for (int i = j + 1; i < B; ++i) {
      F v = G[static_cast<size_t>(i) * B + j];
      for (int k = 0; k < j; ++k)
        v -= L[static_cast<size_t>(i) * B + k] * Field<F>::conj(L[static_cast<size_t>(j) * B + k]);
      L[static_cast<size_t>(i) * B + j] = v / ljj;
    }
  }
  R.assign(static_cast<size_t>(B) * B, F(0.0));
  for (int i = 0; i < B; ++i)
    for (int j = 0; j < B; ++j) R[static_cast<size_t>(i) * B + j] = Field<F>::conj(L[static_cast<size_t>(j) * B + i]);
  // Rinv by back substitution, column by column
  Rinv.assign(static_cast<size_t>(B) * B, F(0.0));
  for (int col = 0; col < B; ++col) {
    for (int i = B - 1; i >= 0; --i) {
      F v = (i == col) ? F(1.0) : F(0.0);
      for (int k = i + 1; k < B; ++k) v -= R[static_cast<size_t>(i) * B + k] * Rinv[static_cast<size_t>(k) * B + col];
      Rinv[static_cast<size_t>(i) * B + col] = v / R[static_cast<size_t>(i) * B + i];
    }
  }
  return true;
}

// C = A * B for B x B row-major matrices.
template <typename F> std::vector<F> matmul(std::vector<F> const &A, std::vector<F> const &Bm, int B) {
  std::vector<F> C(static_cast<size_t>(B) * B, F(0.0));
  for (int i = 0; i < B; ++i)
    for (int k = 0; k < B; ++k)
      for (int j = 0; j < B; ++j)
        C[static_cast<size_t>(i) * B + j] += A[static_cast<size_t>(i) * B + k] * Bm[static_cast<size_t>(k) * B + j];
  return C;
}

bool field_eigh(std::vector<double> const &T, int n, std::vector<double> &w, std::vector<double> &Z) {
  return symmetric_eigh(T, n, w, Z);
}
bool field_eigh(std::vector<std::complex<double>> const &T, int n, std::vector<double> &w,
                std::vector<std::complex<double>> &Z) {
  return hermitian_eigh(T, n, w, Z);
}

template <typename F> void parallel_for(int count, F &&f) {
  unsigned const hw = std::max(1u, std::thread::hardware_concurrency());
  int const nt = static_cast<int>(std::min<unsigned>(std::min<unsigned>(hw, 16u), static_cast<unsigned>(count)));
  if (nt <= 1) {
    for (int i = 0; i < count; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (int i = t; i < count; i += nt) f(i);
    });
  for (auto &th : pool) th.join();
}

struct BlockAlgebra {
  Segments seg;
  int B = 8;
  DeviceArray<double> partial, d_out, d_coef;
  DeviceArray<int> d_flags;
  std::vector<double> h_coef;
  int max_nblk = 0;
  bool init(std::vector<int64_t> const &offsets, std::vector<int> const &listed, int b, int nblk_max) {
    B = b;
    max_nblk = nblk_max;
    if (!seg.init(offsets, listed)) return false;
    size_t const V = static_cast<size_t>(B) * B;
    return partial.allocate(std::max<size_t>(1, static_cast<size_t>(seg.chunks) * nblk_max * V)) &&
           d_out.allocate(static_cast<size_t>(seg.nseg) * nblk_max * V) &&
           d_coef.allocate(static_cast<size_t>(seg.nseg) * nblk_max * V);
  }
  // out_host[(s * nblk + i) * V + p * B + q] = A_i^T W over segment s; with
  // `extra`, one more block follows the nblk strided ones: out_host then
  // holds nblk + 1 matrices per segment, the last extra^T W.
  template <typename TA, typename TW>
  bool gram(TA const *A, int64_t a_stride, int nblk, TW const *W, std::vector<double> &out_host,
            TA const *extra = nullptr) {
    int const total = nblk + (extra != nullptr ? 1 : 0);
    if (total > max_nblk) return fail("internal: too many blocks in a Gram product");
    int const V = B * B;
    if (seg.chunks > 0) {
      bool ok = dispatch_block(B, [&]<int BB>() -> bool {
        seg_gram_kernel<BB, TA, TW><<<seg.chunks, kThreads, 0, stream()>>>(A, a_stride, total, extra, W,
                                                                           seg.d_chunk_rows.ptr, partial.ptr);
        return true;
      });
      if (!ok) return false;
    }
    return reduce_to_host(total, out_host);
  }
  // Per-segment sums of the chunk partials of `nblk` Gram matrices, to the host.
  bool reduce_to_host(int nblk, std::vector<double> &out_host) {
    int const V = B * B;
    if (seg.chunks > 0) {
      seg_reduce_kernel<<<dim3(seg.nseg, nblk), kReduceThreads, 0, stream()>>>(partial.ptr, nblk, V,
                                                                                seg.d_seg_chunk.ptr, d_out.ptr);
      LSAMD_HIP(hipGetLastError());
    }
    out_host.assign(static_cast<size_t>(seg.nseg) * nblk * V, 0.0);
    if (seg.chunks > 0)
      LSAMD_HIP(hipMemcpyAsync(out_host.data(), d_out.ptr, sizeof(double) * out_host.size(), hipMemcpyDeviceToHost,
                               stream()));
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
  // Fresh seeded random entries in the field columns c >= keep[s] of every
  // segment s (field_cols columns per row).
  template <typename T> bool refill(T *X, int field_cols, std::vector<int> const &keep, uint64_t seed) {
    if (seg.chunks == 0) return true;
    if (!d_flags.upload(keep.data(), keep.size())) return false;
    refill_kernel<T><<<seg.chunks, kThreads, 0, stream()>>>(seed, X, B, field_cols, seg.d_chunk_rows.ptr,
                                                            seg.d_chunk_seg.ptr, d_flags.ptr);
    LSAMD_HIP(hipGetLastError());
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
  // out = beta * out + A * C (per segment), coefficients [s][i][p][q] on the
  // host.  With gram_out, also out^T out per segment (as gram(out, 1, out)
  // would give it, bit for bit), from the same pass.
  template <typename TO, typename TA>
  bool combine(TO *out, double beta, TA const *A, int64_t a_stride, int nblk, std::vector<double> const &coef_host,
               std::vector<double> *gram_out = nullptr) {
    if (nblk > max_nblk) return fail("internal: too many blocks in a block update");
    int const V = B * B;
    if (seg.chunks == 0) {
      if (gram_out != nullptr) gram_out->assign(static_cast<size_t>(seg.nseg) * V, 0.0);
      return true;
    }
    h_coef = coef_host;
    LSAMD_HIP(hipMemcpyAsync(d_coef.ptr, h_coef.data(), sizeof(double) * h_coef.size(), hipMemcpyHostToDevice,
                             stream()));
    size_t const smem = sizeof(double) * static_cast<size_t>(nblk) * V;
    bool ok = dispatch_block(B, [&]<int BB>() -> bool {
      if (gram_out != nullptr)
        seg_combine_kernel<BB, TO, TA, true><<<seg.chunks, kThreads, smem, stream()>>>(
            out, beta, A, a_stride, nblk, d_coef.ptr, seg.d_chunk_rows.ptr, seg.d_chunk_seg.ptr, partial.ptr);
      else
        seg_combine_kernel<BB, TO, TA, false><<<seg.chunks, kThreads, smem, stream()>>>(
            out, beta, A, a_stride, nblk, d_coef.ptr, seg.d_chunk_rows.ptr, seg.d_chunk_seg.ptr, nullptr);
      return true;
    });
    if (!ok) return false;
    LSAMD_HIP(hipGetLastError());
    if (gram_out != nullptr) return reduce_to_host(1, *gram_out); // synchronises
    // the host coefficients must outlive the copy
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
};


__global__ __launch_bounds__(kThreads) void column_kernel(double const *__restrict__ X, int b, int col,
                                                          double *__restrict__ out, int64_t rows) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < rows) out[i] = X[i * b + col];
}

} // namespace

bool block_column(double const *X, int b, int col, double *out, int64_t rows) {
  if (rows <= 0) return true;
  column_kernel<<<grid_for(rows, kThreads), kThreads, 0, stream()>>>(X, b, col, out, rows);
  LSAMD_HIP(hipGetLastError());
  return true;
}

bool check_block_krylov_options(BlockKrylovOptions const &o, int multiplicity) {
  int const B = o.block;
  if (o.complex_values) {
    if (B != 1 && B != 2 && B != 4) return fail("block size must be 1, 2 or 4 complex columns");
  } else if (B != 1 && B != 2 && B != 4 && B != 8) {
    return fail("block size must be 1, 2, 4 or 8");
  }
  if (o.nev < 1 || multiplicity * o.nev > B)
    return fail("nev (times 2 for complex sectors) must not exceed the block size");
  if (o.keep < 1 || o.keep >= o.max_blocks - 1) return fail("keep must be in [1, max_blocks - 2]");
  if (o.max_cycles < 1) return fail("max_cycles must be positive");
  return true;
}

namespace {

// The solver over the field F (see block_krylov_solve).  Host-side matrices
// are F; device blocks are real rows of Bw = Bf * width entries.
template <typename F>
bool krylov_solve(std::vector<int64_t> const &seg_rows, std::vector<char> const &active, int multiplicity,
                  BlockProduct const &product, BlockKrylovOptions const &o, BlockKrylovResult &res,
                  double *d_ritz_block) {
  auto const t_start = Clock::now();
  if (!check_block_krylov_options(o, multiplicity)) return false;
  int const Bf = o.block;                        // columns of a block (field entries)
  int const Bw = Bf * Field<F>::kWidth;          // reals per block row on the device
  int const want = multiplicity * o.nev;         // Ritz values per segment
  int const nseg_all = static_cast<int>(seg_rows.size()) - 1;
  if (nseg_all < 1) return fail("no segments");
  if (!active.empty() && static_cast<int>(active.size()) != nseg_all) return fail("internal: bad segment mask");
  // the segments solved here, renumbered 0 .. ns - 1
  std::vector<int> listed;
  for (int s = 0; s < nseg_all; ++s)
    if (active.empty() || active[s]) listed.push_back(s);
  int const ns = static_cast<int>(listed.size());
  if (ns < 1) return fail("no segments");
  int const m = o.max_blocks;
  int const Vf = Bf * Bf, Vw = Bw * Bw;
  int64_t min_rows = INT64_MAX;
  for (int s : listed) min_rows = std::min(min_rows, seg_rows[s + 1] - seg_rows[s]);
  if (min_rows < static_cast<int64_t>(m + 1) * Bf)
    return fail("a segment is too small for (max_blocks + 1) * block_size Krylov vectors (" +
                std::to_string(min_rows) + " rows)");
  BlockAlgebra alg;
  if (!alg.init(seg_rows, listed, Bw, m + 1)) return false; // m blocks + W in one Gram pass
  int64_t const rows = alg.seg.rows;
  int64_t const blk = rows * Bw; // reals per block
  // FP64 buffers; while the product is FP32 the blocks are stored as FP32 in
  // the same memory (block j at float offset j * blk).
  DeviceArray<double> Vb, W, Y;
  if (!Vb.allocate(static_cast<size_t>(blk) * m) || !W.allocate(static_cast<size_t>(blk)) ||
      !Y.allocate(static_cast<size_t>(blk) * o.keep))
    return false;
  bool const partial_rows = ns < nseg_all;
  // rows outside the solved segments are only ever read by the product,
  // which keeps them inside their own (block-diagonal) segment
  auto clear_unsolved = [&](bool with_y) -> bool {
    if (!partial_rows) return true;
    LSAMD_HIP(hipMemsetAsync(Vb.ptr, 0, sizeof(double) * blk * m, stream()));
    LSAMD_HIP(hipMemsetAsync(W.ptr, 0, sizeof(double) * blk, stream()));
    if (with_y) LSAMD_HIP(hipMemsetAsync(Y.ptr, 0, sizeof(double) * blk * o.keep, stream()));
    return true;
  };
  if (!clear_unsolved(true)) return false;

  double prod_s = 0.0;
  int64_t prods = 0;
  auto apply = [&](bool fp32, void const *x, void *y) -> bool {
    auto t0 = Clock::now();
    if (!product(fp32, x, y)) return false;
    LSAMD_HIP(hipStreamSynchronize(stream()));
    prod_s += seconds_since(t0);
    ++prods;
    return true;
  };
  std::vector<double> G, coef;
  std::vector<F> Gf, Rf(static_cast<size_t>(ns) * Vf);
  // coef (real, per segment and block) from field matrices M[s][i]
  auto set_coef = [&](std::vector<F> const &M, int nb) {
    coef.assign(static_cast<size_t>(ns) * nb * Vw, 0.0);
    for (size_t k = 0; k < static_cast<size_t>(ns) * nb; ++k)
      field_to_coef(M.data() + k * Vf, Bf, coef.data() + k * Vw);
  };
  // CholQR2 of block X into out (may be X); Rf receives the factor R per
  // segment (X = out R).  A segment whose Gram matrix is singular has run out
  // of Krylov directions (A applied to the last block falls inside the
  // stored blocks, in some or all directions).  Singular also means a
  // direction below kNoise of w_scale[s], the largest squared column norm of
  // the product block X came from: so small a direction is rounding noise of
  // the orthogonalisation, redrawn at its own size by every rounding of the
  // block (components along the stored blocks included), and normalising it
  // would carry those components into the basis.  Such a block keeps the
  // directions whose Gram eigenvalues stand above kKeep of the largest and
  // above the noise limit (normalised), gets fresh random columns in place of
  // the others, orthogonalised twice against the nprev stored blocks at
  // Vbase, and goes through CholQR2 again; R is then formed explicitly as
  // out^H X, so the Krylov relation and the residual bounds hold up to the
  // directions replaced, which lie below the noise limit.  The projected
  // matrix is filled from explicit projections and needs no change.
  uint64_t refills = 0;
  DeviceArray<double> Xorig; // X before a refill (reals of the block's type)
  auto cholqr = [&]<typename T>(T *X, T *out, T const *Vbase, int nprev, std::vector<double> const *w_scale) -> bool {
    constexpr double kKeep = std::is_same_v<T, float> ? 1e-10 : 1e-24; // relative Gram eigenvalue kept
    // squared size, against ||W||^2, below which a direction is rounding
    // noise of the orthogonalisation: sigma below 1e-4 ||W|| (FP32 storage)
    // or min(1e-8, tol / 10) ||W|| (FP64: a refill drops what it replaces
    // from the Krylov relation, so it must stay below the residuals asked for)
    double const noise64 = std::min(1e-8, 0.1 * o.tol);
    double const kNoise = std::is_same_v<T, float> ? 1e-8 : noise64 * noise64;
    auto floor_of = [&](int s) { return w_scale != nullptr ? kNoise * (*w_scale)[s] : 0.0; };
    std::vector<int> refilled(ns, 0);
    bool saved = false;
    std::vector<double> Cr;
    std::vector<F> Cfr;
    for (int attempt = 0;; ++attempt) {
      std::vector<F> Racc(static_cast<size_t>(ns) * Vf, F(0.0)), Minv(static_cast<size_t>(ns) * Vf);
      for (int s = 0; s < ns; ++s)
        for (int i = 0; i < Bf; ++i) Racc[static_cast<size_t>(s) * Vf + i * Bf + i] = 1.0;
      std::vector<int> singular(ns, 0);
      int first = -1;
      T *src = X;
      std::vector<double> Gnext; // the Gram of the first pass's output, from its update
      for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
          G.swap(Gnext);
        } else if (!alg.gram(src, blk, 1, src, G)) {
          return false;
        }
        gram_to_field(G, static_cast<size_t>(ns), Bf, Gf);
        for (int s = 0; s < ns; ++s) {
          std::vector<F> Gs(Gf.begin() + static_cast<int64_t>(s) * Vf, Gf.begin() + static_cast<int64_t>(s + 1) * Vf);
          std::vector<F> R, Rinv;
          if (!cholesky_upper(Gs, Bf, R, Rinv, pass == 0 ? floor_of(s) : 0.0)) {
            // only the first pass sees X itself; the second sees an
            // orthonormalised block and cannot be singular unless X was
            if (pass == 1)
              return fail("block Lanczos breakdown in segment " + std::to_string(listed[s]) +
                          " (the orthonormalised block is singular)");
            singular[s] = 1;
            if (first < 0) first = s;
            continue;
          }
          std::copy(Rinv.begin(), Rinv.end(), Minv.begin() + static_cast<int64_t>(s) * Vf);
          std::vector<F> Ra(Racc.begin() + static_cast<int64_t>(s) * Vf,
                            Racc.begin() + static_cast<int64_t>(s + 1) * Vf);
          std::vector<F> const Rn = matmul(R, Ra, Bf);
          std::copy(Rn.begin(), Rn.end(), Racc.begin() + static_cast<int64_t>(s) * Vf);
        }
        if (first >= 0) break; // X is untouched: repair below and start over
        set_coef(Minv, 1);
        if (!alg.combine(out, 0.0, src, blk, 1, coef, pass == 0 ? &Gnext : nullptr)) return false;
        src = out;
      }
      if (first < 0) {
        Rf = Racc;
        if (saved) { // R = out^H X for the repaired segments
          if (!alg.gram(out, blk, 1, reinterpret_cast<T const *>(Xorig.ptr), Cr)) return false;
          gram_to_field(Cr, static_cast<size_t>(ns), Bf, Cfr);
          for (int s = 0; s < ns; ++s)
            if (refilled[s])
              std::copy(Cfr.begin() + static_cast<int64_t>(s) * Vf, Cfr.begin() + static_cast<int64_t>(s + 1) * Vf,
                        Rf.begin() + static_cast<int64_t>(s) * Vf);
        }
        return true;
      }
      if (attempt == 3)
        return fail("block Lanczos breakdown in segment " + std::to_string(listed[first]) +
                    " (no independent directions left; lower max_blocks or the block size)");
      if (!saved) {
        if (!Xorig.ptr && !Xorig.allocate(static_cast<size_t>(blk))) return false;
        LSAMD_HIP(hipMemcpyAsync(Xorig.ptr, X, sizeof(T) * blk, hipMemcpyDeviceToDevice, stream()));
        saved = true;
      }
      // keep the independent directions: X <- X M with M = U_K Lambda_K^-1/2
      // (columns of the kept Gram eigenvectors, largest first, then zeros)
      std::vector<F> M(static_cast<size_t>(ns) * Vf, F(0.0));
      std::vector<int> keep(ns, Bf);
      for (int s = 0; s < ns; ++s) {
        F *Ms = M.data() + static_cast<size_t>(s) * Vf;
        if (!singular[s]) {
          for (int i = 0; i < Bf; ++i) Ms[i * Bf + i] = 1.0;
          continue;
        }
        std::vector<F> Gs(Gf.begin() + static_cast<int64_t>(s) * Vf, Gf.begin() + static_cast<int64_t>(s + 1) * Vf);
        std::vector<double> lam;
        std::vector<F> U;
        if (!field_eigh(Gs, Bf, lam, U)) return fail("block Lanczos: Gram eigen-decomposition failed");
        double const top = lam[Bf - 1];
        int k = 0;
        double const thr = std::max(kKeep * top, floor_of(s));
        for (int c = Bf - 1; c >= 0 && top > 0.0 && lam[c] > thr; --c, ++k)
          for (int r = 0; r < Bf; ++r) Ms[r * Bf + k] = U[static_cast<size_t>(r) * Bf + c] / std::sqrt(lam[c]);
        keep[s] = k;
      }
      set_coef(M, 1);
      if (!alg.combine(X, 0.0, X, blk, 1, coef)) return false;
      // fresh random columns in place of the others, orthogonalised twice
      // against the stored blocks (zero coefficients keep other segments)
      ++refills;
      if (!alg.refill(X, Bf, keep, o.seed ^ (0x5bd1e995ull * refills))) return false;
      for (int pass = 0; pass < 2 && nprev > 0; ++pass) {
        if (!alg.gram(Vbase, blk, nprev, X, Cr)) return false;
        gram_to_field(Cr, static_cast<size_t>(ns) * nprev, Bf, Cfr);
        for (int s = 0; s < ns; ++s)
          for (size_t t = 0; t < static_cast<size_t>(nprev) * Vf; ++t) {
            F &c = Cfr[static_cast<size_t>(s) * nprev * Vf + t];
            c = singular[s] ? -c : F(0.0);
          }
        set_coef(Cfr, nprev);
        if (!alg.combine(X, 1.0, Vbase, blk, nprev, coef)) return false;
      }
      for (int s = 0; s < ns; ++s) refilled[s] |= singular[s];
    }
  };

  int const dimT = m * Bf;
  std::vector<std::vector<F>> Tp(ns, std::vector<F>(static_cast<size_t>(dimT) * dimT, F(0.0)));
  std::vector<std::vector<double>> theta(ns);
  std::vector<std::vector<F>> Z(ns);
  std::vector<double> worst(ns, 0.0);
  int kept = 0;
  int cycle = 0, cycles_run = 0, fp64_cycles = 0, blocks_done = 0;
  double prev_worst = INFINITY;
  bool done = false;
  std::vector<double> C, Cw, Gw;
  std::vector<F> Cf, Csum, Mf, Gwf;
  // Rayleigh-Ritz on the leading nbk blocks of each segment's projected
  // matrix: theta, Z (nbk Bf square) and the residual bounds || Rf Z_last ||,
  // Rf being the factor of block nbk (the one after the last used).
  int zdim = dimT; // order of the last Rayleigh-Ritz problem
  auto rayleigh_ritz = [&](int nbk) -> bool {
    int const d = nbk * Bf;
    zdim = d;
    std::vector<int> eig_ok(ns, 1);
    parallel_for(ns, [&](int s) {
      if (d == dimT) {
        eig_ok[s] = field_eigh(Tp[s], dimT, theta[s], Z[s]) ? 1 : 0;
      } else {
        std::vector<F> Ts(static_cast<size_t>(d) * d);
        for (int r = 0; r < d; ++r)
          std::copy(Tp[s].begin() + static_cast<int64_t>(r) * dimT, Tp[s].begin() + static_cast<int64_t>(r) * dimT + d,
                    Ts.begin() + static_cast<int64_t>(r) * d);
        eig_ok[s] = field_eigh(Ts, d, theta[s], Z[s]) ? 1 : 0;
      }
      if (!eig_ok[s]) return;
      double w = 0.0;
      F const *Rs = Rf.data() + static_cast<size_t>(s) * Vf;
      for (int q = 0; q < want; ++q) {
        double rs = 0.0;
        for (int p = 0; p < Bf; ++p) {
          F acc = 0.0;
          for (int c = 0; c < Bf; ++c) acc += Rs[p * Bf + c] * Z[s][static_cast<size_t>((nbk - 1) * Bf + c) * d + q];
          rs += std::norm(acc);
        }
        w = std::max(w, std::sqrt(rs) / std::max(1.0, std::fabs(theta[s][q])));
      }
      worst[s] = w;
    });
    for (int s = 0; s < ns; ++s)
      if (!eig_ok[s]) return fail("projected eigenproblem did not converge in segment " + std::to_string(listed[s]));
    return true;
  };
  // coefficients of Ritz block t over nbk blocks: M[s][i] = Z_s[i-block rows, t-block columns]
  auto ritz_coef = [&](int t, int nbk) {
    Mf.assign(static_cast<size_t>(ns) * nbk * Vf, F(0.0));
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < nbk; ++i)
        for (int p = 0; p < Bf; ++p)
          for (int q = 0; q < Bf; ++q)
            Mf[(static_cast<size_t>(s) * nbk + i) * Vf + p * Bf + q] =
                Z[s][static_cast<size_t>(i * Bf + p) * zdim + t * Bf + q];
    set_coef(Mf, nbk);
  };
  // In the FP64 cycles a trial Rayleigh-Ritz after each block lets the solve
  // stop inside a cycle, where the host work is small beside a product
  // (segments x order^3 of the projected matrix at most kTrialWork).
  constexpr double kTrialWork = 4e6;
  bool const trials = static_cast<double>(ns) * dimT * dimT * dimT <= kTrialWork;

  // Restart cycles with the blocks stored as T: float runs the FP32 product,
  // double the FP64 product.  `kept` blocks (and their part of Tp) are in
  // place on entry.  Returns 1 when FP32 gives way to FP64, 0 when converged
  // or out of cycles, -1 on error; on 0 and 1 the lowest Ritz block is in Y
  // as FP64.
  auto run_cycles = [&]<typename T>(T *Vbase, T *Wt, T *Yt) -> int {
    constexpr bool fp32 = std::is_same_v<T, float>;
    auto Vj = [&](int j) { return Vbase + static_cast<int64_t>(j) * blk; };
    for (; cycle < o.max_cycles; ++cycle) {
      if (!fp32) ++fp64_cycles;
      for (int j = kept; j < m; ++j) {
        if (!apply(fp32, Vj(j), Wt)) return -1;
        ++blocks_done;
        int const nb = j + 1;
        Csum.assign(static_cast<size_t>(ns) * nb * Vf, F(0.0));
        // The second pass only when the first removed nearly all of some
        // direction of the block (the smallest eigenvalue of what it leaves
        // below kReorthDrop of ||W||^2, by Pythagoras from W^H W and V^H W),
        // where one pass leaves mostly rounding error (an exhausted Krylov
        // space, a converged direction).
        // Pass 0 also forms W^H W (the last block of its Gram pass).
        bool second = false;
        std::vector<double> wscale(ns, 0.0); // largest squared column norm of W per segment
        for (int pass = 0; pass < (second ? 2 : 1); ++pass) {
          if (pass == 0) {
            if (!alg.gram(Vbase, blk, nb, Wt, Cw, Wt)) return -1;
            C.resize(static_cast<size_t>(ns) * nb * Vw);
            Gw.resize(static_cast<size_t>(ns) * Vw);
            for (int s = 0; s < ns; ++s) {
              auto const src = Cw.begin() + static_cast<std::ptrdiff_t>(s) * (nb + 1) * Vw;
              std::copy(src, src + static_cast<std::ptrdiff_t>(nb) * Vw,
                        C.begin() + static_cast<std::ptrdiff_t>(s) * nb * Vw);
              std::copy(src + static_cast<std::ptrdiff_t>(nb) * Vw, src + static_cast<std::ptrdiff_t>(nb + 1) * Vw,
                        Gw.begin() + static_cast<std::ptrdiff_t>(s) * Vw);
            }
            gram_to_field(Gw, static_cast<size_t>(ns), Bf, Gwf);
          } else if (!alg.gram(Vbase, blk, nb, Wt, C)) {
            return -1;
          }
          gram_to_field(C, static_cast<size_t>(ns) * nb, Bf, Cf);
          if (pass == 0) {
            // what the first pass leaves, X^H X = W^H W - C^H C (Pythagoras);
            // its smallest eigenvalue against the largest squared column
            // norm of W: a direction of the block can vanish while every
            // column keeps part of its norm
            for (int s = 0; s < ns; ++s) {
              std::vector<F> Ms(Vf);
              double wmax = 0.0;
              for (int p = 0; p < Bf; ++p) {
                wmax = std::max(wmax, Field<F>::re(Gwf[static_cast<size_t>(s) * Vf + p * Bf + p]));
                for (int q = 0; q < Bf; ++q) {
                  F v = Gwf[static_cast<size_t>(s) * Vf + p * Bf + q];
                  for (int i = 0; i < nb; ++i) {
                    F const *Ci = Cf.data() + (static_cast<size_t>(s) * nb + i) * Vf;
                    for (int a = 0; a < Bf; ++a) v -= Field<F>::conj(Ci[a * Bf + p]) * Ci[a * Bf + q];
                  }
                  Ms[p * Bf + q] = v;
                }
              }
              for (int p = 0; p < Bf; ++p)
                for (int q = 0; q < p; ++q) {
                  F const h = 0.5 * (Ms[p * Bf + q] + Field<F>::conj(Ms[q * Bf + p]));
                  Ms[p * Bf + q] = h;
                  Ms[q * Bf + p] = Field<F>::conj(h);
                }
              wscale[s] = wmax;
              std::vector<double> lam;
              std::vector<F> U;
              if (!field_eigh(Ms, Bf, lam, U) || lam[0] < kReorthDrop * wmax) second = true;
            }
          }
          for (size_t t = 0; t < Cf.size(); ++t) {
            Csum[t] += Cf[t];
            Cf[t] = -Cf[t];
          }
          set_coef(Cf, nb);
          if (!alg.combine(Wt, 1.0, Vbase, blk, nb, coef)) return -1;
        }
        // column j of the projected matrix (and its mirror) from the
        // coefficients V_i^H A V_j
        for (int s = 0; s < ns; ++s) {
          std::vector<F> &Ts = Tp[s];
          for (int i = 0; i < nb; ++i) {
            F const *Ci = Csum.data() + (static_cast<size_t>(s) * nb + i) * Vf;
            for (int p = 0; p < Bf; ++p)
              for (int q = 0; q < Bf; ++q) {
                size_t const r = static_cast<size_t>(i * Bf + p), c = static_cast<size_t>(j * Bf + q);
                if (i == j) {
                  Ts[r * dimT + c] = 0.5 * (Ci[p * Bf + q] + Field<F>::conj(Ci[q * Bf + p]));
                } else {
                  Ts[r * dimT + c] = Ci[p * Bf + q];
                  Ts[c * dimT + r] = Field<F>::conj(Ci[p * Bf + q]);
                }
              }
          }
        }
        if (j + 1 < m) {
          if (!cholqr(Wt, Vj(j + 1), Vbase, nb, &wscale)) return -1;
          if (!fp32 && trials && j + 1 - kept >= 2) {
            if (!rayleigh_ritz(j + 1)) return -1;
            double wmax = 0.0;
            for (int s = 0; s < ns; ++s) wmax = std::max(wmax, worst[s]);
            if (wmax < o.tol) {
              ++cycles_run;
              res.history.push_back(wmax);
              done = true;
              ritz_coef(0, j + 1);
              if (!alg.combine(Y.ptr, 0.0, Vbase, blk, j + 1, coef)) return -1;
              ++cycle;
              return 0;
            }
          }
        } else {
          if (!cholqr(Wt, Wt, Vbase, nb, &wscale)) return -1; // residual block; Rf holds its factor
        }
      }
      // Rayleigh-Ritz per segment, residual bounds || Rf Z_last ||
      if (!rayleigh_ritz(m)) return -1;
      double wmax = 0.0;
      for (int s = 0; s < ns; ++s) wmax = std::max(wmax, worst[s]);
      ++cycles_run;
      res.history.push_back(wmax);
      done = !fp32 && wmax < o.tol;
      bool const last = cycle + 1 == o.max_cycles;
      // FP32 gives way to FP64 at the switch tolerance or when FP32 stalls
      bool const to_fp64 = fp32 && !last && (wmax < o.switch_tol || wmax > 0.9 * prev_worst);
      prev_worst = wmax;
      if (done || last || to_fp64) {
        ritz_coef(0, m);
        if (!alg.combine(Y.ptr, 0.0, Vbase, blk, m, coef)) return -1;
        ++cycle;
        return to_fp64 ? 1 : 0;
      }
      // Restart: Y_t = sum_i V_i Z_s[i-block, t-block], then the residual block.
      // (One pass per Ritz block: a pass forming several at once has B*B
      // FMAs per stored row and output and is bound by instruction issue,
      // not memory, for B = 8.)
      int const r = o.keep;
      for (int t = 0; t < r; ++t) {
        ritz_coef(t, m);
        if (!alg.combine(Yt + static_cast<int64_t>(t) * blk, 0.0, Vbase, blk, m, coef)) return -1;
      }
      LSAMD_HIP(hipMemcpyAsync(Vbase, Yt, sizeof(T) * blk * r, hipMemcpyDeviceToDevice, stream()));
      LSAMD_HIP(hipMemcpyAsync(Vj(r), Wt, sizeof(T) * blk, hipMemcpyDeviceToDevice, stream()));
      for (int s = 0; s < ns; ++s) {
        std::fill(Tp[s].begin(), Tp[s].end(), F(0.0));
        for (int q = 0; q < r * Bf; ++q) Tp[s][static_cast<size_t>(q) * dimT + q] = theta[s][q];
      }
      kept = r;
      LSAMD_HIP(hipStreamSynchronize(stream()));
    }
    return 0;
  };

  int status = 0;
  if (o.fp32) {
    float *V32 = reinterpret_cast<float *>(Vb.ptr);
    normal_kernel<float><<<grid_for(blk, kThreads), kThreads, 0, stream()>>>(o.seed, V32, blk);
    LSAMD_HIP(hipGetLastError());
    if (!cholqr(V32, V32, V32, 0, nullptr)) return false;
    status = run_cycles(V32, reinterpret_cast<float *>(W.ptr), reinterpret_cast<float *>(Y.ptr));
    if (status < 0) return false;
    if (status == 1) {
      // explicit restart from the lowest Ritz block (in Y) with the FP64
      // product; Y's unsolved rows are still zero
      LSAMD_HIP(hipStreamSynchronize(stream()));
      if (!clear_unsolved(false)) return false;
      LSAMD_HIP(hipMemcpyAsync(Vb.ptr, Y.ptr, sizeof(double) * blk, hipMemcpyDeviceToDevice, stream()));
      for (int s = 0; s < ns; ++s) std::fill(Tp[s].begin(), Tp[s].end(), F(0.0));
      kept = 0;
      if (!cholqr(Vb.ptr, Vb.ptr, Vb.ptr, 0, nullptr)) return false;
    }
  } else {
    normal_kernel<double><<<grid_for(blk, kThreads), kThreads, 0, stream()>>>(o.seed, Vb.ptr, blk);
    LSAMD_HIP(hipGetLastError());
    if (!cholqr(Vb.ptr, Vb.ptr, Vb.ptr, 0, nullptr)) return false;
    status = 1;
  }
  if (status == 1 && run_cycles(Vb.ptr, W.ptr, Y.ptr) < 0) return false;
  // Final: the lowest Ritz block (in Y), its FP64 Rayleigh quotients and residuals.
  if (!apply(false, Y.ptr, W.ptr)) return false;
  std::vector<double> Gr, Gy;
  std::vector<F> Th, Grf, Gyf;
  if (!alg.gram(Y.ptr, blk, 1, W.ptr, G)) return false; // Y^H A Y
  gram_to_field(G, static_cast<size_t>(ns), Bf, Th);
  // Y^H Y: the quotients and residuals are taken per unit column, so a Ritz
  // block that is orthonormal only to the solver's accuracy still gives
  // Rayleigh quotients accurate to the square of its residuals
  if (!alg.gram(Y.ptr, blk, 1, Y.ptr, Gy)) return false;
  gram_to_field(Gy, static_cast<size_t>(ns), Bf, Gyf);
  std::vector<F> negTh(Th.size());
  for (size_t t = 0; t < Th.size(); ++t) negTh[t] = -Th[t];
  set_coef(negTh, 1);
  if (!alg.combine(W.ptr, 1.0, Y.ptr, blk, 1, coef, &Gr)) return false; // A Y - Y (Y^H A Y), and its Gram
  gram_to_field(Gr, static_cast<size_t>(ns), Bf, Grf);
  res.evals.assign(static_cast<size_t>(nseg_all) * o.nev, NAN);
  res.residuals.assign(static_cast<size_t>(nseg_all) * o.nev, NAN);
  double rmax = 0.0;
  for (int s = 0; s < ns; ++s)
    for (int e = 0; e < o.nev; ++e) {
      double lam = 0.0, r2 = 0.0;
      for (int c = 0; c < multiplicity; ++c) {
        int const q = multiplicity * e + c;
        double const n2 = Field<F>::re(Gyf[static_cast<size_t>(s) * Vf + q * Bf + q]);
        lam += Field<F>::re(Th[static_cast<size_t>(s) * Vf + q * Bf + q]) / n2;
        r2 = std::max(r2, Field<F>::re(Grf[static_cast<size_t>(s) * Vf + q * Bf + q]) / n2);
      }
      lam /= multiplicity;
      double const rel = std::sqrt(std::max(0.0, r2)) / std::max(1.0, std::fabs(lam));
      res.evals[static_cast<size_t>(listed[s]) * o.nev + e] = lam;
      res.residuals[static_cast<size_t>(listed[s]) * o.nev + e] = rel;
      rmax = std::max(rmax, rel);
    }
  res.first_rayleigh = Field<F>::re(Th[0]) / Field<F>::re(Gyf[0]);
  res.first_residual = std::sqrt(std::max(0.0, Field<F>::re(Grf[0]) / Field<F>::re(Gyf[0])));
  if (d_ritz_block != nullptr)
    LSAMD_HIP(hipMemcpyAsync(d_ritz_block, Y.ptr, sizeof(double) * blk, hipMemcpyDeviceToDevice, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  res.cycles = cycles_run;
  // the bounds decide when to stop; convergence is also claimed only when
  // the explicit residuals agree
  res.converged = done && rmax <= std::max(10.0 * o.tol, 1e-12);
  res.fp64_cycles = fp64_cycles;
  res.blocks = blocks_done;
  res.products = prods;
  res.product_seconds = prod_s;
  res.max_residual = rmax;
  res.total_seconds = seconds_since(t_start);
  return true;
}

} // namespace

bool block_krylov_solve(std::vector<int64_t> const &seg_rows, std::vector<char> const &active, int multiplicity,
                        BlockProduct const &product, BlockKrylovOptions const &o, BlockKrylovResult &res,
                        double *d_ritz_block) {
  if (o.complex_values)
    return krylov_solve<std::complex<double>>(seg_rows, active, multiplicity, product, o, res, d_ritz_block);
  return krylov_solve<double>(seg_rows, active, multiplicity, product, o, res, d_ritz_block);
}

} // namespace lsamd
