// sectors.hip -- many symmetry sectors in one product, and a block Lanczos
// run on all of them at once.
//
// The sectors of one Hamiltonian (for a chain: every momentum times both
// spin-inversion eigenvalues) each give a complex Hermitian matrix
// M_s = A_s + i B_s.  Their complex CSR caches are concatenated into one
// block-diagonal CSR with 32-bit global columns.  Vectors carry K complex
// right-hand sides per state, laid out [state][re | im][K]; read as
// [2 * state + part][K] this is the real embedding [[A, -B], [B, A]] of M,
// whose spectrum is that of M with every eigenvalue twice.
//
// The product is one persistent launch over every sector (a fixed grid of
// workgroups looping over row tiles, optionally with each accelerator die's
// workgroups kept on one contiguous range of rows), captured once in a HIP
// graph and replayed.  The eigensolver is a thick-restart block Lanczos over
// all sectors together: one product launch per block step, FP64 Krylov
// blocks, two passes of full re-orthogonalisation by segmented (per-sector)
// Gram reductions and updates, CholQR2 per sector, and per-sector Rayleigh-
// Ritz solves on host threads.  The product runs in FP32 until the FP32
// residual floor, then in FP64.
//
// Reference counterpart: the reference builds one basis per sector and runs
// PRIMME / scipy eigsh on each in turn (chapel/src/Diagonalize.chpl:134-162,
// python/lattice_symmetries/__init__.py:591-752) with the real-valued
// product of DistributedMatrixVector.chpl:70-140; there is no batched form.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "operator_internal.hpp"
#include "orbit_device.hpp"

struct ls_amd_sector_batch {
  int count = 0;
  std::vector<int64_t> offsets; // state offset of every sector, count + 1
  lsamd::Csr csr;               // block-diagonal, complex values
  // product schedule
  int lanes = 4;
  int blocks = 2048;
  int xcd_aware = 1;
  // graph of the last product launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void const *g_x = nullptr;
  void *g_y = nullptr;
  int g_k = 0, g_form = -1, g_lanes = 0, g_blocks = 0, g_xcd = -1;
  ~ls_amd_sector_batch() {
    if (exec != nullptr) (void)hipGraphExecDestroy(exec);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
  }
};

namespace lsamd {
namespace {

constexpr int kThreads = 256;
constexpr int kDies = 8; // accelerator complex dies; workgroups are dealt round-robin

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// ---------------------------------------------------------------------------
// The product
// ---------------------------------------------------------------------------

template <typename T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

// y[i] = d_i x[i] + sum_k M_{i, cols[k]} x[cols[k]], for K complex columns.
// vals[2k], vals[2k + 1] are (re, im) of M_{i, cols[k]}.  x is T, y is TO
// (FP32 products may write FP64 output directly).
template <int K, int LPR, typename T, typename TO>
__global__ __launch_bounds__(kThreads) void sector_spmv_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                               uint32_t const *__restrict__ cols,
                                                               T const *__restrict__ vals,
                                                               T const *__restrict__ diag,
                                                               T const *__restrict__ x, TO *__restrict__ y,
                                                               int xcd_aware) {
  using T2 = typename Pair<T>::type;
  constexpr int kRows = kThreads / LPR;
  int64_t const tiles = (n + kRows - 1) / kRows;
  int64_t t, t_end, t_step;
  if (xcd_aware) {
    // gridDim.x is a multiple of kDies; die d owns tiles [tiles d / 8, tiles (d + 1) / 8)
    int const die = static_cast<int>(blockIdx.x) % kDies;
    int64_t const local = static_cast<int64_t>(blockIdx.x) / kDies;
    t = tiles * die / kDies + local;
    t_end = tiles * (die + 1) / kDies;
    t_step = static_cast<int64_t>(gridDim.x) / kDies;
  } else {
    t = blockIdx.x;
    t_end = tiles;
    t_step = gridDim.x;
  }
  int const sub = static_cast<int>(threadIdx.x) % LPR;
  int const r_in = static_cast<int>(threadIdx.x) / LPR;
  T2 const *__restrict__ v2 = reinterpret_cast<T2 const *>(vals);
  for (; t < t_end; t += t_step) {
    int64_t const row = t * kRows + r_in;
    T re[K], im[K];
#pragma unroll
    for (int q = 0; q < K; ++q) re[q] = im[q] = T(0);
    if (row < n) {
      int64_t const b = rp[row], e = rp[row + 1];
      for (int64_t k = b + sub; k < e; k += LPR) {
        T2 const v = v2[k];
        T const *__restrict__ xr = x + static_cast<size_t>(cols[k]) * (2 * K);
        T xre[K], xim[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
          xre[q] = xr[q];
          xim[q] = xr[K + q];
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
          re[q] = fma(v.x, xre[q], re[q]);
          re[q] = fma(-v.y, xim[q], re[q]);
          im[q] = fma(v.y, xre[q], im[q]);
          im[q] = fma(v.x, xim[q], im[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) {
        re[q] += __shfl_xor(re[q], off, LPR);
        im[q] += __shfl_xor(im[q], off, LPR);
      }
    }
    if (row < n && sub == 0) {
      T const d = diag[row];
      T const *__restrict__ xi = x + static_cast<size_t>(row) * (2 * K);
      TO *__restrict__ yi = y + static_cast<size_t>(row) * (2 * K);
#pragma unroll
      for (int q = 0; q < K; ++q) {
        yi[q] = static_cast<TO>(fma(d, xi[q], re[q]));
        yi[K + q] = static_cast<TO>(fma(d, xi[K + q], im[q]));
      }
    }
  }
}

template <typename F> bool dispatch_lanes(int lpr, F &&f) {
  switch (lpr) {
  case 1: return f.template operator()<1>();
  case 2: return f.template operator()<2>();
  case 4: return f.template operator()<4>();
  case 8: return f.template operator()<8>();
  case 16: return f.template operator()<16>();
  default: return fail("lanes per row must be 1, 2, 4, 8 or 16");
  }
}

template <typename F> bool dispatch_columns(int k, F &&f) {
  switch (k) {
  case 1: return f.template operator()<1>();
  case 2: return f.template operator()<2>();
  case 4: return f.template operator()<4>();
  case 8: return f.template operator()<8>();
  default: return fail("block size must be 1, 2, 4 or 8");
  }
}

// form: 0 FP64 in / out, 1 FP32 in / out, 2 FP32 in / FP64 out
bool launch_product(ls_amd_sector_batch const &b, int form, int k, void const *x, void *y, int lanes, int blocks,
                    int xcd) {
  Csr const &c = b.csr;
  int64_t const n = c.rows;
  if (n == 0) return true;
  return dispatch_columns(k, [&]<int K>() -> bool {
    return dispatch_lanes(lanes, [&]<int LPR>() -> bool {
      if (form == 0) {
        sector_spmv_kernel<K, LPR, double, double><<<blocks, kThreads, 0, stream()>>>(
            n, c.row_ptr.ptr, c.cols.ptr, c.vals64.ptr, c.diag64.ptr, static_cast<double const *>(x),
            static_cast<double *>(y), xcd);
      } else if (form == 1) {
        sector_spmv_kernel<K, LPR, float, float><<<blocks, kThreads, 0, stream()>>>(
            n, c.row_ptr.ptr, c.cols.ptr, c.vals32.ptr, c.diag32.ptr, static_cast<float const *>(x),
            static_cast<float *>(y), xcd);
      } else {
        sector_spmv_kernel<K, LPR, float, double><<<blocks, kThreads, 0, stream()>>>(
            n, c.row_ptr.ptr, c.cols.ptr, c.vals32.ptr, c.diag32.ptr, static_cast<float const *>(x),
            static_cast<double *>(y), xcd);
      }
      return true;
    });
  });
}

bool check_schedule(int lanes, int blocks, int xcd) {
  if (blocks < kDies || blocks > (1 << 20)) return fail("the product grid must hold 8 .. 2^20 workgroups");
  if (xcd && blocks % kDies != 0) return fail("a die-aware grid must be a multiple of 8 workgroups");
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16)
    return fail("lanes per row must be 1, 2, 4, 8 or 16");
  return true;
}

// The product through the cached graph: captured on first use of a given
// (form, k, x, y, schedule) and replayed afterwards.  The graph holds the one
// kernel node and no memory operations.
bool graph_product(ls_amd_sector_batch &b, int form, int k, void const *x, void *y) {
  if (!check_schedule(b.lanes, b.blocks, b.xcd_aware)) return false;
  if (b.csr.rows == 0) return true;
  bool const same = b.exec != nullptr && b.g_x == x && b.g_y == y && b.g_k == k && b.g_form == form &&
                    b.g_lanes == b.lanes && b.g_blocks == b.blocks && b.g_xcd == b.xcd_aware;
  if (!same) {
    if (b.exec != nullptr) (void)hipGraphExecDestroy(b.exec);
    if (b.graph != nullptr) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    LSAMD_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    bool const ok = launch_product(b, form, k, x, y, b.lanes, b.blocks, b.xcd_aware);
    hipGraph_t g = nullptr;
    hipError_t const e = hipStreamEndCapture(stream(), &g);
    if (!ok) {
      if (g != nullptr) (void)hipGraphDestroy(g);
      return false;
    }
    LSAMD_HIP(e);
    b.graph = g;
    LSAMD_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    b.g_x = x;
    b.g_y = y;
    b.g_k = k;
    b.g_form = form;
    b.g_lanes = b.lanes;
    b.g_blocks = b.blocks;
    b.g_xcd = b.xcd_aware;
  }
  timer_begin(kTimerSectorProduct);
  LSAMD_HIP(hipGraphLaunch(b.exec, stream()));
  timer_end(kTimerSectorProduct);
  return true;
}

__global__ __launch_bounds__(kThreads) void offset_rows_kernel(int64_t n, int64_t const *__restrict__ in,
                                                               int64_t add, int64_t *__restrict__ out) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) out[i] = in[i] + add;
}

__global__ __launch_bounds__(kThreads) void offset_cols_kernel(int64_t n, uint32_t const *__restrict__ in,
                                                               uint32_t add, uint32_t *__restrict__ out) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) out[i] = in[i] + add;
}

// ---------------------------------------------------------------------------
// Segmented FP64 block algebra on row-major [rows][B] blocks.  Rows are cut
// into chunks that never straddle a sector; every reduction is a fixed-order
// two-pass sum, so results do not depend on scheduling.
// ---------------------------------------------------------------------------

constexpr int kChunkRows = 2048;

// partial[(c * nblk + i) * B*B + p*B + q] = sum over rows r of chunk c of
// A_i[r][p] * W[r][q], with A_i = A + i * a_stride.
template <int B>
__global__ __launch_bounds__(kThreads) void seg_gram_kernel(double const *__restrict__ A, int64_t a_stride, int nblk,
                                                            double const *__restrict__ W,
                                                            int64_t const *__restrict__ chunk_rows,
                                                            double *__restrict__ partial) {
  constexpr int V = B * B;
  __shared__ double s[kThreads / 64][64];
  int const c = blockIdx.x;
  int64_t const r0 = chunk_rows[c], r1 = chunk_rows[c + 1];
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = 0; i < nblk; ++i) {
    double const *__restrict__ Ai = A + static_cast<int64_t>(i) * a_stride;
    double acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kThreads) {
      double a[B], w[B];
#pragma unroll
      for (int p = 0; p < B; ++p) {
        a[p] = Ai[r * B + p];
        w[p] = W[r * B + p];
      }
#pragma unroll
      for (int p = 0; p < B; ++p)
#pragma unroll
        for (int q = 0; q < B; ++q) acc[p * B + q] = fma(a[p], w[q], acc[p * B + q]);
    }
    // Transpose-reduce across the wave: after the steps below lane L holds
    // the wave's sum for value index L & (V - 1).
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

// out[(s * nblk + i) * V + t] = sum over the chunks of sector s, in order.
__global__ __launch_bounds__(64) void seg_reduce_kernel(double const *__restrict__ partial, int nblk, int V,
                                                        int const *__restrict__ seg_chunk,
                                                        double *__restrict__ out) {
  int const s = blockIdx.x, i = blockIdx.y, t = threadIdx.x;
  if (t >= V) return;
  double sum = 0.0;
  for (int c = seg_chunk[s]; c < seg_chunk[s + 1]; ++c) sum += partial[(static_cast<int64_t>(c) * nblk + i) * V + t];
  out[(static_cast<int64_t>(s) * nblk + i) * V + t] = sum;
}

// out[r][q] = beta * out[r][q] + sum_{i < nblk} sum_p A_i[r][p] C_s[i][p][q]
// for the sector s of row r.  Each thread owns whole rows, so out may alias
// one of the A_i.
template <int B>
__global__ __launch_bounds__(kThreads) void seg_combine_kernel(double *out, double beta, double const *A,
                                                               int64_t a_stride, int nblk,
                                                               double const *__restrict__ coef,
                                                               int64_t const *__restrict__ chunk_rows,
                                                               int const *__restrict__ chunk_seg) {
  constexpr int V = B * B;
  extern __shared__ double sc[];
  int const c = blockIdx.x;
  int const s = chunk_seg[c];
  int64_t const r0 = chunk_rows[c], r1 = chunk_rows[c + 1];
  double const *__restrict__ cs = coef + static_cast<int64_t>(s) * nblk * V;
  for (int t = threadIdx.x; t < nblk * V; t += kThreads) sc[t] = cs[t];
  __syncthreads();
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kThreads) {
    double acc[B];
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] = 0.0;
    for (int i = 0; i < nblk; ++i) {
      double a[B];
#pragma unroll
      for (int p = 0; p < B; ++p) a[p] = A[static_cast<int64_t>(i) * a_stride + r * B + p];
#pragma unroll
      for (int p = 0; p < B; ++p)
#pragma unroll
        for (int q = 0; q < B; ++q) acc[q] = fma(a[p], sc[i * V + p * B + q], acc[q]);
    }
    if (beta != 0.0) {
#pragma unroll
      for (int q = 0; q < B; ++q) acc[q] = fma(beta, out[r * B + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < B; ++q) out[r * B + q] = acc[q];
  }
}

__global__ __launch_bounds__(kThreads) void to_f32_rows_kernel(double const *__restrict__ x, float *__restrict__ y,
                                                               int64_t n) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) y[i] = static_cast<float>(x[i]);
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

template <typename F> bool dispatch_block(int b, F &&f) {
  switch (b) {
  case 1: return f.template operator()<1>();
  case 2: return f.template operator()<2>();
  case 4: return f.template operator()<4>();
  case 8: return f.template operator()<8>();
  default: return fail("block size must be 1, 2, 4 or 8");
  }
}

// Chunks of the embedded rows, per sector.
struct Segments {
  int nseg = 0;
  int64_t rows = 0;
  int chunks = 0;
  std::vector<int64_t> row_begin; // nseg + 1
  DeviceArray<int64_t> d_chunk_rows;
  DeviceArray<int> d_chunk_seg, d_seg_chunk;
  bool init(std::vector<int64_t> const &state_offsets) {
    nseg = static_cast<int>(state_offsets.size()) - 1;
    row_begin.resize(nseg + 1);
    for (int s = 0; s <= nseg; ++s) row_begin[s] = 2 * state_offsets[s];
    rows = row_begin[nseg];
    std::vector<int64_t> cr{0};
    std::vector<int> cs, sc{0};
    for (int s = 0; s < nseg; ++s) {
      for (int64_t r = row_begin[s]; r < row_begin[s + 1]; r += kChunkRows) {
        cr.push_back(std::min(r + kChunkRows, row_begin[s + 1]));
        cs.push_back(s);
      }
      sc.push_back(static_cast<int>(cs.size()));
    }
    chunks = static_cast<int>(cs.size());
    return d_chunk_rows.upload(cr.data(), cr.size()) && d_chunk_seg.upload(cs.data(), std::max<size_t>(1, cs.size())) &&
           d_seg_chunk.upload(sc.data(), sc.size());
  }
};

// Host-side dense helpers for the per-sector small matrices.
// Cholesky G = L L^T of a B x B SPD matrix; returns false if a pivot is not
// positive (rank deficiency).  R = L^T (upper), Rinv = R^{-1}.
bool cholesky_upper(std::vector<double> const &G, int B, std::vector<double> &R, std::vector<double> &Rinv) {
  std::vector<double> L(static_cast<size_t>(B) * B, 0.0);
  double scale = 0.0;
  for (int i = 0; i < B; ++i) scale = std::max(scale, G[static_cast<size_t>(i) * B + i]);
  for (int j = 0; j < B; ++j) {
    double d = G[static_cast<size_t>(j) * B + j];
    for (int k = 0; k < j; ++k) d -= L[static_cast<size_t>(j) * B + k] * L[static_cast<size_t>(j) * B + k];
    if (!(d > 1e-28 * std::max(scale, 1e-300))) return false;
    double const ljj = std::sqrt(d);
    L[static_cast<size_t>(j) * B + j] = ljj;
    for (int i = j + 1; i < B; ++i) {
      double v = G[static_cast<size_t>(i) * B + j];
      for (int k = 0; k < j; ++k) v -= L[static_cast<size_t>(i) * B + k] * L[static_cast<size_t>(j) * B + k];
      L[static_cast<size_t>(i) * B + j] = v / ljj;
    }
  }
  R.assign(static_cast<size_t>(B) * B, 0.0);
  for (int i = 0; i < B; ++i)
    for (int j = 0; j < B; ++j) R[static_cast<size_t>(i) * B + j] = L[static_cast<size_t>(j) * B + i];
  // Rinv by back substitution, column by column
  Rinv.assign(static_cast<size_t>(B) * B, 0.0);
  for (int col = 0; col < B; ++col) {
    for (int i = B - 1; i >= 0; --i) {
      double v = (i == col) ? 1.0 : 0.0;
      for (int k = i + 1; k < B; ++k) v -= R[static_cast<size_t>(i) * B + k] * Rinv[static_cast<size_t>(k) * B + col];
      Rinv[static_cast<size_t>(i) * B + col] = v / R[static_cast<size_t>(i) * B + i];
    }
  }
  return true;
}

// C = A * B for B x B row-major matrices.
std::vector<double> matmul(std::vector<double> const &A, std::vector<double> const &Bm, int B) {
  std::vector<double> C(static_cast<size_t>(B) * B, 0.0);
  for (int i = 0; i < B; ++i)
    for (int k = 0; k < B; ++k)
      for (int j = 0; j < B; ++j)
        C[static_cast<size_t>(i) * B + j] += A[static_cast<size_t>(i) * B + k] * Bm[static_cast<size_t>(k) * B + j];
  return C;
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
  std::vector<double> h_coef;
  int max_nblk = 0;
  bool init(std::vector<int64_t> const &offsets, int b, int nblk_max) {
    B = b;
    max_nblk = nblk_max;
    if (!seg.init(offsets)) return false;
    size_t const V = static_cast<size_t>(B) * B;
    return partial.allocate(std::max<size_t>(1, static_cast<size_t>(seg.chunks) * nblk_max * V)) &&
           d_out.allocate(static_cast<size_t>(seg.nseg) * nblk_max * V) &&
           d_coef.allocate(static_cast<size_t>(seg.nseg) * nblk_max * V);
  }
  // out_host[(s * nblk + i) * V + p * B + q] = A_i^T W over sector s
  bool gram(double const *A, int64_t a_stride, int nblk, double const *W, std::vector<double> &out_host) {
    if (nblk > max_nblk) return fail("internal: too many blocks in a Gram product");
    int const V = B * B;
    if (seg.chunks > 0) {
      bool ok = dispatch_block(B, [&]<int BB>() -> bool {
        seg_gram_kernel<BB><<<seg.chunks, kThreads, 0, stream()>>>(A, a_stride, nblk, W, seg.d_chunk_rows.ptr,
                                                                   partial.ptr);
        return true;
      });
      if (!ok) return false;
      seg_reduce_kernel<<<dim3(seg.nseg, nblk), 64, 0, stream()>>>(partial.ptr, nblk, V, seg.d_seg_chunk.ptr,
                                                                    d_out.ptr);
      LSAMD_HIP(hipGetLastError());
    }
    out_host.assign(static_cast<size_t>(seg.nseg) * nblk * V, 0.0);
    if (seg.chunks > 0)
      LSAMD_HIP(hipMemcpyAsync(out_host.data(), d_out.ptr, sizeof(double) * out_host.size(), hipMemcpyDeviceToHost,
                               stream()));
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
  // out = beta * out + A * C (per sector), coefficients [s][i][p][q] on the host
  bool combine(double *out, double beta, double const *A, int64_t a_stride, int nblk,
               std::vector<double> const &coef_host) {
    if (nblk > max_nblk) return fail("internal: too many blocks in a block update");
    int const V = B * B;
    if (seg.chunks == 0) return true;
    h_coef = coef_host;
    LSAMD_HIP(hipMemcpyAsync(d_coef.ptr, h_coef.data(), sizeof(double) * h_coef.size(), hipMemcpyHostToDevice,
                             stream()));
    size_t const smem = sizeof(double) * static_cast<size_t>(nblk) * V;
    bool ok = dispatch_block(B, [&]<int BB>() -> bool {
      seg_combine_kernel<BB><<<seg.chunks, kThreads, smem, stream()>>>(out, beta, A, a_stride, nblk, d_coef.ptr,
                                                                       seg.d_chunk_rows.ptr, seg.d_chunk_seg.ptr);
      return true;
    });
    if (!ok) return false;
    LSAMD_HIP(hipGetLastError());
    // the host coefficients must outlive the copy
    LSAMD_HIP(hipStreamSynchronize(stream()));
    return true;
  }
};

} // namespace

// ---------------------------------------------------------------------------
// Batch construction
// ---------------------------------------------------------------------------

bool sector_batch_build(ls_amd_sector_batch &b, OperatorCore *const *ops, int count) {
  if (count < 1) return fail("a sector batch needs at least one operator");
  b.count = count;
  b.offsets.assign(count + 1, 0);
  int64_t nnz = 0;
  for (int s = 0; s < count; ++s) {
    Csr const &c = ops[s]->csr;
    if (!c.is_complex) return fail("sector " + std::to_string(s) + ": build the complex CSR (precision 128) first");
    b.offsets[s + 1] = b.offsets[s] + c.rows;
    nnz += c.nnz;
  }
  int64_t const n = b.offsets[count];
  if (n >= (int64_t{1} << 32) - 1) return fail("sector batch columns are 32-bit: too many states");
  Csr &c = b.csr;
  c.rows = n;
  c.nnz = nnz;
  if (!c.row_ptr.allocate(static_cast<size_t>(n + 1)) || !c.cols.allocate(static_cast<size_t>(std::max<int64_t>(nnz, 1))) ||
      !c.vals64.allocate(static_cast<size_t>(2 * std::max<int64_t>(nnz, 1))) ||
      !c.vals32.allocate(static_cast<size_t>(2 * std::max<int64_t>(nnz, 1))) ||
      !c.diag64.allocate(static_cast<size_t>(std::max<int64_t>(n, 1))) ||
      !c.diag32.allocate(static_cast<size_t>(std::max<int64_t>(n, 1))))
    return false;
  int64_t nz0 = 0;
  for (int s = 0; s < count; ++s) {
    Csr const &cs = ops[s]->csr;
    int64_t const r0 = b.offsets[s], rn = cs.rows;
    if (rn > 0) {
      offset_rows_kernel<<<grid_for(rn, kThreads), kThreads, 0, stream()>>>(rn, cs.row_ptr.ptr, nz0,
                                                                            c.row_ptr.ptr + r0);
      LSAMD_HIP(hipMemcpyAsync(c.diag64.ptr + r0, cs.diag64.ptr, sizeof(double) * rn, hipMemcpyDeviceToDevice,
                               stream()));
      LSAMD_HIP(hipMemcpyAsync(c.diag32.ptr + r0, cs.diag32.ptr, sizeof(float) * rn, hipMemcpyDeviceToDevice,
                               stream()));
    }
    if (cs.nnz > 0) {
      offset_cols_kernel<<<grid_for(cs.nnz, kThreads), kThreads, 0, stream()>>>(
          cs.nnz, cs.cols.ptr, static_cast<uint32_t>(r0), c.cols.ptr + nz0);
      LSAMD_HIP(hipMemcpyAsync(c.vals64.ptr + 2 * nz0, cs.vals64.ptr, sizeof(double) * 2 * cs.nnz,
                               hipMemcpyDeviceToDevice, stream()));
      LSAMD_HIP(hipMemcpyAsync(c.vals32.ptr + 2 * nz0, cs.vals32.ptr, sizeof(float) * 2 * cs.nnz,
                               hipMemcpyDeviceToDevice, stream()));
    }
    LSAMD_HIP(hipGetLastError());
    nz0 += cs.nnz;
  }
  LSAMD_HIP(hipMemcpyAsync(c.row_ptr.ptr + n, &nz0, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  c.has32 = c.has64 = true;
  c.is_complex = true;
  // default schedule: a few workgroups per compute unit, die-aware
  hipDeviceProp_t prop;
  int dev = 0;
  LSAMD_HIP(hipGetDevice(&dev));
  LSAMD_HIP(hipGetDeviceProperties(&prop, dev));
  b.blocks = std::max(kDies, (prop.multiProcessorCount * 4 / kDies) * kDies);
  b.lanes = 4;
  b.xcd_aware = 1;
  return true;
}

bool sector_product(ls_amd_sector_batch &b, int form, int k, void const *x, void *y) {
  if (k != 1 && k != 2 && k != 4 && k != 8) return fail("block size must be 1, 2, 4 or 8");
  return graph_product(b, form, k, x, y);
}

// Times the product over lanes per row x grid size x die-aware mapping and
// keeps the fastest, for the FP32 form with k columns.
bool sector_autotune(ls_amd_sector_batch &b, int k) {
  int64_t const n = b.csr.rows;
  if (n == 0) return true;
  DeviceArray<float> x, y;
  size_t const len = static_cast<size_t>(n) * 2 * k;
  if (!x.allocate(len) || !y.allocate(len)) return false;
  normal_kernel<float><<<grid_for(static_cast<int64_t>(len), kThreads), kThreads, 0, stream()>>>(
      12345, x.ptr, static_cast<int64_t>(len));
  LSAMD_HIP(hipGetLastError());
  hipDeviceProp_t prop;
  int dev = 0;
  LSAMD_HIP(hipGetDevice(&dev));
  LSAMD_HIP(hipGetDeviceProperties(&prop, dev));
  int const cu = prop.multiProcessorCount;
  hipEvent_t e0, e1;
  LSAMD_HIP(hipEventCreate(&e0));
  LSAMD_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bl = b.lanes, bb = b.blocks, bx = b.xcd_aware;
  for (int lanes : {1, 2, 4, 8, 16})
    for (int per_cu : {2, 4, 8, 16})
      for (int xcd : {0, 1}) {
        int const blocks = std::max(kDies, (cu * per_cu / kDies) * kDies);
        if (!launch_product(b, 1, k, x.ptr, y.ptr, lanes, blocks, xcd)) return false;
        LSAMD_HIP(hipEventRecord(e0, stream()));
        for (int r = 0; r < 3; ++r)
          if (!launch_product(b, 1, k, x.ptr, y.ptr, lanes, blocks, xcd)) return false;
        LSAMD_HIP(hipEventRecord(e1, stream()));
        LSAMD_HIP(hipEventSynchronize(e1));
        float ms = 0;
        LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          bl = lanes;
          bb = blocks;
          bx = xcd;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b.lanes = bl;
  b.blocks = bb;
  b.xcd_aware = bx;
  LSAMD_HIP(hipStreamSynchronize(stream()));
  return true;
}

// ---------------------------------------------------------------------------
// Thick-restart block Lanczos over every sector at once
// ---------------------------------------------------------------------------

struct SectorEigOptions {
  int block = 8;       // columns per block (the real embedding's block)
  int nev = 4;         // eigenvalues of M wanted per sector (2 nev Ritz values of the embedding)
  int max_blocks = 12; // Krylov blocks per cycle
  int keep = 2;        // Ritz blocks kept at a restart
  int max_cycles = 40;
  double tol = 1e-7;        // relative residual for convergence (FP64 product)
  double switch_tol = 1e-5; // relative residual at which the FP32 product gives way to FP64
  uint64_t seed = 1;
};

struct SectorEigResult {
  std::vector<double> evals;     // count * nev
  std::vector<double> residuals; // count * nev: FP64 relative residual norms
  int cycles = 0, fp64_cycles = 0, blocks = 0;
  bool converged = false;
  int64_t products = 0;
  double product_seconds = 0.0, total_seconds = 0.0, max_residual = 0.0;
};

namespace {
void jacobi_eigh_local(std::vector<double> A, int m, std::vector<double> &w, std::vector<double> &Z) {
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
} // namespace

bool sector_eigensolve(ls_amd_sector_batch &bt, SectorEigOptions const &o, SectorEigResult &res) {
  auto const t_start = Clock::now();
  int const B = o.block;
  if (B != 1 && B != 2 && B != 4 && B != 8) return fail("block size must be 1, 2, 4 or 8");
  int const want = 2 * o.nev; // Ritz values of the embedding per sector
  if (o.nev < 1 || want > B) return fail("2 nev must not exceed the block size");
  if (o.keep < 1 || o.keep >= o.max_blocks - 1) return fail("keep must be in [1, max_blocks - 2]");
  int const ns = bt.count;
  int const m = o.max_blocks;
  int const V = B * B;
  int64_t min_rows = INT64_MAX;
  for (int s = 0; s < ns; ++s) min_rows = std::min(min_rows, 2 * (bt.offsets[s + 1] - bt.offsets[s]));
  if (min_rows < static_cast<int64_t>(m + 1) * B)
    return fail("a sector is too small for max_blocks * block_size Krylov vectors (" + std::to_string(min_rows) +
                " embedded rows)");
  BlockAlgebra alg;
  if (!alg.init(bt.offsets, B, m)) return false;
  int64_t const rows = alg.seg.rows;
  int64_t const blk = rows * B; // doubles per block
  DeviceArray<double> Vb, W, Y;
  DeviceArray<float> X32;
  if (!Vb.allocate(static_cast<size_t>(blk) * m) || !W.allocate(static_cast<size_t>(blk)) ||
      !Y.allocate(static_cast<size_t>(blk) * o.keep) || !X32.allocate(static_cast<size_t>(blk)))
    return false;
  auto Vj = [&](int j) { return Vb.ptr + static_cast<int64_t>(j) * blk; };

  double prod_s = 0.0;
  int64_t prods = 0;
  auto product = [&](bool fp32, double const *x, double *y) -> bool {
    auto t0 = Clock::now();
    if (fp32) {
      to_f32_rows_kernel<<<grid_for(blk, kThreads), kThreads, 0, stream()>>>(x, X32.ptr, blk);
      LSAMD_HIP(hipGetLastError());
      if (!graph_product(bt, 2, B, X32.ptr, y)) return false;
    } else {
      if (!launch_product(bt, 0, B, x, y, bt.lanes, bt.blocks, bt.xcd_aware)) return false;
      LSAMD_HIP(hipGetLastError());
    }
    LSAMD_HIP(hipStreamSynchronize(stream()));
    prod_s += seconds_since(t0);
    ++prods;
    return true;
  };
  std::vector<double> G, Rf(static_cast<size_t>(ns) * V), coef;
  // CholQR2 of block X in place (or into out); Rf receives the factor R per sector.
  auto cholqr = [&](double *X, double *out) -> bool {
    std::vector<double> Racc(static_cast<size_t>(ns) * V, 0.0);
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < B; ++i) Racc[static_cast<size_t>(s) * V + i * B + i] = 1.0;
    double *src = X;
    for (int pass = 0; pass < 2; ++pass) {
      if (!alg.gram(src, blk, 1, src, G)) return false;
      coef.assign(static_cast<size_t>(ns) * V, 0.0);
      for (int s = 0; s < ns; ++s) {
        std::vector<double> Gs(G.begin() + static_cast<int64_t>(s) * V, G.begin() + static_cast<int64_t>(s + 1) * V);
        std::vector<double> R, Rinv;
        if (!cholesky_upper(Gs, B, R, Rinv))
          return fail("block Lanczos breakdown in sector " + std::to_string(s) +
                      " (the Krylov space is exhausted; lower max_blocks or the block size)");
        std::copy(Rinv.begin(), Rinv.end(), coef.begin() + static_cast<int64_t>(s) * V);
        std::vector<double> Ra(Racc.begin() + static_cast<int64_t>(s) * V, Racc.begin() + static_cast<int64_t>(s + 1) * V);
        std::vector<double> const Rn = matmul(R, Ra, B);
        std::copy(Rn.begin(), Rn.end(), Racc.begin() + static_cast<int64_t>(s) * V);
      }
      if (!alg.combine(out, 0.0, src, blk, 1, coef)) return false;
      src = out;
    }
    Rf = Racc;
    return true;
  };

  // start block
  normal_kernel<double><<<grid_for(blk, kThreads), kThreads, 0, stream()>>>(o.seed, Vj(0), blk);
  LSAMD_HIP(hipGetLastError());
  if (!cholqr(Vj(0), Vj(0))) return false;

  int const dimT = m * B;
  std::vector<std::vector<double>> T(ns, std::vector<double>(static_cast<size_t>(dimT) * dimT, 0.0));
  std::vector<std::vector<double>> theta(ns), Z(ns);
  std::vector<double> worst(ns, 0.0);
  bool fp32 = true;
  int kept = 0;
  int cycle = 0, cycles_run = 0, fp64_cycles = 0, blocks_done = 0;
  double prev_worst = INFINITY;
  bool done = false;
  std::vector<double> C, Csum;
  for (cycle = 0; cycle < o.max_cycles; ++cycle) {
    if (!fp32) ++fp64_cycles;
    for (int j = kept; j < m; ++j) {
      if (!product(fp32, Vj(j), W.ptr)) return false;
      ++blocks_done;
      int const nb = j + 1;
      Csum.assign(static_cast<size_t>(ns) * nb * V, 0.0);
      for (int pass = 0; pass < 2; ++pass) {
        if (!alg.gram(Vb.ptr, blk, nb, W.ptr, C)) return false;
        for (size_t t = 0; t < C.size(); ++t) {
          Csum[t] += C[t];
          C[t] = -C[t];
        }
        if (!alg.combine(W.ptr, 1.0, Vb.ptr, blk, nb, C)) return false;
      }
      for (int s = 0; s < ns; ++s) {
        std::vector<double> &Ts = T[s];
        for (int i = 0; i < nb; ++i) {
          double const *Ci = Csum.data() + (static_cast<size_t>(s) * nb + i) * V;
          for (int p = 0; p < B; ++p)
            for (int q = 0; q < B; ++q) {
              size_t const r = static_cast<size_t>(i * B + p), c = static_cast<size_t>(j * B + q);
              if (i == j) {
                Ts[r * dimT + c] = 0.5 * (Ci[p * B + q] + Ci[q * B + p]);
              } else {
                Ts[r * dimT + c] = Ci[p * B + q];
                Ts[c * dimT + r] = Ci[p * B + q];
              }
            }
        }
      }
      if (j + 1 < m) {
        if (!cholqr(W.ptr, Vj(j + 1))) return false;
      } else {
        if (!cholqr(W.ptr, W.ptr)) return false; // residual block; Rf holds its factor
      }
    }
    // Rayleigh-Ritz per sector
    parallel_for(ns, [&](int s) {
      jacobi_eigh_local(T[s], dimT, theta[s], Z[s]);
      double w = 0.0;
      double const *Rs = Rf.data() + static_cast<size_t>(s) * V;
      for (int q = 0; q < want; ++q) {
        double rs = 0.0;
        for (int p = 0; p < B; ++p) {
          double acc = 0.0;
          for (int c = 0; c < B; ++c) acc += Rs[p * B + c] * Z[s][static_cast<size_t>((m - 1) * B + c) * dimT + q];
          rs += acc * acc;
        }
        w = std::max(w, std::sqrt(rs) / std::max(1.0, std::fabs(theta[s][q])));
      }
      worst[s] = w;
    });
    double wmax = 0.0;
    for (int s = 0; s < ns; ++s) wmax = std::max(wmax, worst[s]);
    ++cycles_run;
    if (!fp32 && wmax < o.tol) {
      done = true;
      break;
    }
    if (cycle + 1 == o.max_cycles) break; // keep V and Z for the final Ritz block
    // FP32 gives way to FP64 at the switch tolerance or when FP32 stalls
    bool const to_fp64 = fp32 && (wmax < o.switch_tol || wmax > 0.5 * prev_worst);
    prev_worst = wmax;
    // Restart: Y_t = sum_i V_i Z_s[i-block, t-block]
    int const r = to_fp64 ? 1 : o.keep;
    for (int t = 0; t < r; ++t) {
      coef.assign(static_cast<size_t>(ns) * m * V, 0.0);
      for (int s = 0; s < ns; ++s)
        for (int i = 0; i < m; ++i)
          for (int p = 0; p < B; ++p)
            for (int q = 0; q < B; ++q)
              coef[(static_cast<size_t>(s) * m + i) * V + p * B + q] =
                  Z[s][static_cast<size_t>(i * B + p) * dimT + t * B + q];
      if (!alg.combine(Y.ptr + static_cast<int64_t>(t) * blk, 0.0, Vb.ptr, blk, m, coef)) return false;
    }
    LSAMD_HIP(hipMemcpyAsync(Vb.ptr, Y.ptr, sizeof(double) * blk * r, hipMemcpyDeviceToDevice, stream()));
    for (int s = 0; s < ns; ++s) std::fill(T[s].begin(), T[s].end(), 0.0);
    if (to_fp64) {
      // explicit restart from the lowest Ritz block with the FP64 product
      fp32 = false;
      kept = 0;
      if (!cholqr(Vj(0), Vj(0))) return false;
    } else {
      LSAMD_HIP(hipMemcpyAsync(Vj(r), W.ptr, sizeof(double) * blk, hipMemcpyDeviceToDevice, stream()));
      for (int s = 0; s < ns; ++s)
        for (int q = 0; q < r * B; ++q) T[s][static_cast<size_t>(q) * dimT + q] = theta[s][q];
      kept = r;
    }
    LSAMD_HIP(hipStreamSynchronize(stream()));
  }
  // Final: the lowest Ritz block, its FP64 residuals, and the eigenvalues.
  coef.assign(static_cast<size_t>(ns) * m * V, 0.0);
  for (int s = 0; s < ns; ++s)
    for (int i = 0; i < m; ++i)
      for (int p = 0; p < B; ++p)
        for (int q = 0; q < B; ++q)
          coef[(static_cast<size_t>(s) * m + i) * V + p * B + q] = Z[s][static_cast<size_t>(i * B + p) * dimT + q];
  if (!alg.combine(Y.ptr, 0.0, Vb.ptr, blk, m, coef)) return false;
  if (!product(false, Y.ptr, W.ptr)) return false;
  std::vector<double> Th, Gr;
  if (!alg.gram(Y.ptr, blk, 1, W.ptr, Th)) return false; // Y^T R Y
  for (double &v : Th) v = -v;
  if (!alg.combine(W.ptr, 1.0, Y.ptr, blk, 1, Th)) return false; // R Y - Y (Y^T R Y)
  if (!alg.gram(W.ptr, blk, 1, W.ptr, Gr)) return false;
  res.evals.assign(static_cast<size_t>(ns) * o.nev, 0.0);
  res.residuals.assign(static_cast<size_t>(ns) * o.nev, 0.0);
  double rmax = 0.0;
  for (int s = 0; s < ns; ++s)
    for (int e = 0; e < o.nev; ++e) {
      double const a = -Th[static_cast<size_t>(s) * V + (2 * e) * B + 2 * e];
      double const bb = -Th[static_cast<size_t>(s) * V + (2 * e + 1) * B + 2 * e + 1];
      double const lam = 0.5 * (a + bb);
      double const r2 = std::max(Gr[static_cast<size_t>(s) * V + (2 * e) * B + 2 * e],
                                 Gr[static_cast<size_t>(s) * V + (2 * e + 1) * B + 2 * e + 1]);
      double const rel = std::sqrt(std::max(0.0, r2)) / std::max(1.0, std::fabs(lam));
      res.evals[static_cast<size_t>(s) * o.nev + e] = lam;
      res.residuals[static_cast<size_t>(s) * o.nev + e] = rel;
      rmax = std::max(rmax, rel);
    }
  res.cycles = cycles_run;
  res.converged = done;
  res.fp64_cycles = fp64_cycles;
  res.blocks = blocks_done;
  res.products = prods;
  res.product_seconds = prod_s;
  res.max_residual = rmax;
  res.total_seconds = seconds_since(t_start);
  return true;
}

} // namespace lsamd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

using namespace lsamd;

extern "C" ls_amd_sector_batch *ls_amd_sector_batch_create(ls_amd_operator *const *ops, int count) {
  if (ops == nullptr || count < 1) {
    fail("ls_amd_sector_batch_create: need at least one operator");
    return nullptr;
  }
  if (!ensure_device()) return nullptr;
  std::vector<OperatorCore *> cores(count);
  for (int s = 0; s < count; ++s) {
    if (ops[s] == nullptr) {
      fail("ls_amd_sector_batch_create: null operator");
      return nullptr;
    }
    cores[s] = &ops[s]->core;
  }
  auto *b = new ls_amd_sector_batch;
  if (!sector_batch_build(*b, cores.data(), count)) {
    delete b;
    return nullptr;
  }
  return b;
}

extern "C" void ls_amd_sector_batch_destroy(ls_amd_sector_batch *b) { delete b; }

extern "C" int ls_amd_sector_batch_count(ls_amd_sector_batch const *b) { return b != nullptr ? b->count : -1; }

extern "C" int64_t ls_amd_sector_batch_number_states(ls_amd_sector_batch const *b) {
  return b != nullptr ? b->csr.rows : -1;
}

extern "C" int64_t ls_amd_sector_batch_nnz(ls_amd_sector_batch const *b) { return b != nullptr ? b->csr.nnz : -1; }

extern "C" int ls_amd_sector_batch_offsets(ls_amd_sector_batch const *b, int64_t *out) {
  if (b == nullptr || out == nullptr) return fail("null argument") ? 0 : -1;
  std::copy(b->offsets.begin(), b->offsets.end(), out);
  return 0;
}

extern "C" int ls_amd_sector_batch_set_schedule(ls_amd_sector_batch *b, int lanes, int blocks, int die_aware) {
  if (b == nullptr) return fail("null argument") ? 0 : -1;
  if (!check_schedule(lanes, blocks, die_aware != 0)) return -1;
  b->lanes = lanes;
  b->blocks = blocks;
  b->xcd_aware = die_aware != 0 ? 1 : 0;
  return 0;
}

extern "C" int ls_amd_sector_batch_get_schedule(ls_amd_sector_batch const *b, int *lanes, int *blocks,
                                                int *die_aware) {
  if (b == nullptr || lanes == nullptr || blocks == nullptr || die_aware == nullptr)
    return fail("null argument") ? 0 : -1;
  *lanes = b->lanes;
  *blocks = b->blocks;
  *die_aware = b->xcd_aware;
  return 0;
}

extern "C" int ls_amd_sector_batch_autotune(ls_amd_sector_batch *b, int k) {
  if (b == nullptr) return fail("null argument") ? 0 : -1;
  if (k != 1 && k != 2 && k != 4 && k != 8) return fail("block size must be 1, 2, 4 or 8") ? 0 : -1;
  return sector_autotune(*b, k) ? 0 : -1;
}

extern "C" int ls_amd_sector_batch_matvec_f32(ls_amd_sector_batch *b, int k, float const *d_x, float *d_y) {
  if (b == nullptr) return fail("null argument") ? 0 : -1;
  bool const ok = sector_product(*b, 1, k, d_x, d_y);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_sector_batch_matvec_f64(ls_amd_sector_batch *b, int k, double const *d_x, double *d_y) {
  if (b == nullptr) return fail("null argument") ? 0 : -1;
  bool const ok = sector_product(*b, 0, k, d_x, d_y);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_sector_batch_eigensolve(ls_amd_sector_batch *b, int block_size, int nev, int max_blocks,
                                              int keep_blocks, int max_cycles, double tol, double switch_tol,
                                              uint64_t seed, double *evals, double *residuals,
                                              ls_amd_sector_eig_info *info) {
  if (b == nullptr || evals == nullptr) return fail("null argument") ? 0 : -1;
  SectorEigOptions o;
  o.block = block_size;
  o.nev = nev;
  o.max_blocks = max_blocks;
  o.keep = keep_blocks;
  o.max_cycles = max_cycles;
  o.tol = tol;
  o.switch_tol = switch_tol;
  o.seed = seed;
  SectorEigResult r;
  if (!sector_eigensolve(*b, o, r)) return -1;
  std::copy(r.evals.begin(), r.evals.end(), evals);
  if (residuals != nullptr) std::copy(r.residuals.begin(), r.residuals.end(), residuals);
  if (info != nullptr) {
    info->cycles = r.cycles;
    info->converged = r.converged ? 1 : 0;
    info->fp64_cycles = r.fp64_cycles;
    info->blocks = r.blocks;
    info->products = r.products;
    info->product_seconds = r.product_seconds;
    info->total_seconds = r.total_seconds;
    info->max_residual = r.max_residual;
  }
  return 0;
}
