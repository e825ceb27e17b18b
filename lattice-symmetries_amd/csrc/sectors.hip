// sectors.hip -- many symmetry sectors in one product, and a block Lanczos
// run on all of them at once.
//
// The sectors of one Hamiltonian (for a chain: every momentum times both
// spin-inversion eigenvalues) each give a complex Hermitian matrix
// M_s = A_s + i B_s.  Their complex CSR caches are concatenated into one
// block-diagonal CSR with 32-bit global columns.  Vectors carry K complex
// right-hand sides per state, laid out [state][re | im][K].
//
// The product is one persistent launch over every sector (a fixed grid of
// workgroups looping over row tiles, optionally with each accelerator die's
// workgroups kept on one contiguous range of rows), captured once in a HIP
// graph and replayed.  The eigensolver is the segmented thick-restart block
// Lanczos of block_krylov.hip in complex arithmetic, a block being K complex
// columns in the same [re | im] layout, with one segment per sector: every
// block step is one product launch over all sectors.
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
  // instantiated graphs of recent product launches, one kernel node each
  struct Graph {
    void const *x = nullptr;
    void *y = nullptr;
    int k = 0, form = -1, lanes = 0, blocks = 0, xcd = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
  };
  std::vector<Graph> graphs;
  uint64_t graph_clock = 0;
  int hermitian = -1; // every sector's M Hermitian: -1 untested, 0 no, 1 yes
  hipStream_t capture_stream = nullptr; // graph capture only
  static void release(Graph &g) {
    if (g.exec != nullptr) (void)hipGraphExecDestroy(g.exec);
    if (g.graph != nullptr) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
  }
  ~ls_amd_sector_batch() {
    for (Graph &g : graphs) release(g);
    if (capture_stream != nullptr) (void)hipStreamDestroy(capture_stream);
  }
};

namespace lsamd {
namespace {

constexpr int kThreads = 256;
constexpr int kDies = 8; // accelerator complex dies; workgroups are dealt round-robin
constexpr int kMaxGrid = 1 << 26;

// ---------------------------------------------------------------------------
// The product
// ---------------------------------------------------------------------------

template <typename T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

// y[i] = d_i x[i] + sum_k M_{i, cols[k]} x[cols[k]], for K complex columns.
// vals[2k], vals[2k + 1] are (re, im) of M_{i, cols[k]}.  x is T, y is TO
// (both float or both double as launched).
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

// form: 0 FP64 in / out, 1 FP32 in / out
bool launch_product(ls_amd_sector_batch const &b, int form, int k, void const *x, void *y, int lanes, int blocks,
                    int xcd, hipStream_t st) {
  Csr const &c = b.csr;
  int64_t const n = c.rows;
  if (n == 0) return true;
  return dispatch_columns(k, [&]<int K>() -> bool {
    return dispatch_lanes(lanes, [&]<int LPR>() -> bool {
      if (form == 0) {
        sector_spmv_kernel<K, LPR, double, double><<<blocks, kThreads, 0, st>>>(
            n, c.row_ptr.ptr, c.cols.ptr, c.vals64.ptr, c.diag64.ptr, static_cast<double const *>(x),
            static_cast<double *>(y), xcd);
      } else {
        sector_spmv_kernel<K, LPR, float, float><<<blocks, kThreads, 0, st>>>(
            n, c.row_ptr.ptr, c.cols.ptr, c.vals32.ptr, c.diag32.ptr, static_cast<float const *>(x),
            static_cast<float *>(y), xcd);
      }
      return true;
    });
  });
}

bool check_schedule(int lanes, int blocks, int xcd) {
  if (blocks < kDies || blocks > kMaxGrid) return fail("the product grid must hold 8 .. 2^26 workgroups");
  if (xcd && blocks % kDies != 0) return fail("a die-aware grid must be a multiple of 8 workgroups");
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16)
    return fail("lanes per row must be 1, 2, 4, 8 or 16");
  return true;
}

// The product through a cached graph: captured on first use of a given
// (form, k, x, y, schedule) and replayed afterwards.  Each graph holds the one
// kernel node and no memory operations; the least recently used of
// kGraphCache graphs is dropped to make room (a Krylov solve cycles through
// one input block per Krylov block).
constexpr size_t kGraphCache = 64;

bool graph_product(ls_amd_sector_batch &b, int form, int k, void const *x, void *y) {
  if (!check_schedule(b.lanes, b.blocks, b.xcd_aware)) return false;
  if (b.csr.rows == 0) return true;
  using Graph = ls_amd_sector_batch::Graph;
  Graph *hit = nullptr;
  for (Graph &g : b.graphs)
    if (g.x == x && g.y == y && g.k == k && g.form == form && g.lanes == b.lanes && g.blocks == b.blocks &&
        g.xcd == b.xcd_aware) {
      hit = &g;
      break;
    }
  if (hit == nullptr) {
    // Captured on the batch's own stream, which carries nothing else: work
    // that other threads put on the engine's stream meanwhile runs normally
    // instead of landing in the graph.  The graph itself is launched on the
    // engine's stream, in order with the caller's other work.
    if (b.capture_stream == nullptr)
      LSAMD_HIP(hipStreamCreateWithFlags(&b.capture_stream, hipStreamNonBlocking));
    LSAMD_HIP(hipStreamBeginCapture(b.capture_stream, hipStreamCaptureModeThreadLocal));
    bool const ok = launch_product(b, form, k, x, y, b.lanes, b.blocks, b.xcd_aware, b.capture_stream);
    hipGraph_t gr = nullptr;
    hipError_t const e = hipStreamEndCapture(b.capture_stream, &gr);
    if (!ok || e != hipSuccess) {
      if (gr != nullptr) (void)hipGraphDestroy(gr);
      if (!ok) return false;
      LSAMD_HIP(e);
    }
    hipGraphExec_t ex = nullptr;
    hipError_t const ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(gr);
      LSAMD_HIP(ei);
    }
    if (b.graphs.size() >= kGraphCache) {
      auto lru = std::min_element(b.graphs.begin(), b.graphs.end(),
                                  [](Graph const &p, Graph const &q) { return p.last_use < q.last_use; });
      ls_amd_sector_batch::release(*lru);
      b.graphs.erase(lru);
    }
    Graph g;
    g.x = x;
    g.y = y;
    g.k = k;
    g.form = form;
    g.lanes = b.lanes;
    g.blocks = b.blocks;
    g.xcd = b.xcd_aware;
    g.graph = gr;
    g.exec = ex;
    b.graphs.push_back(g);
    hit = &b.graphs.back();
  }
  hit->last_use = ++b.graph_clock;
  timer_begin(kTimerSectorProduct);
  LSAMD_HIP(hipGraphLaunch(hit->exec, stream()));
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
  LSAMD_HIP(hipMemsetAsync(x.ptr, 0, sizeof(float) * len, stream()));
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
    for (int per_cu : {2, 4, 8, 16, 0})
      for (int xcd : {0, 1}) {
        // per_cu 0: one workgroup per row tile (rounded up to a multiple of 8)
        int64_t const tiles = (n + kThreads / lanes - 1) / (kThreads / lanes);
        int const blocks = per_cu > 0 ? std::max(kDies, (cu * per_cu / kDies) * kDies)
                                      : static_cast<int>(std::min<int64_t>((tiles + 7) / 8 * 8, kMaxGrid));
        if (!launch_product(b, 1, k, x.ptr, y.ptr, lanes, blocks, xcd, stream())) return false;
        LSAMD_HIP(hipEventRecord(e0, stream()));
        for (int r = 0; r < 3; ++r)
          if (!launch_product(b, 1, k, x.ptr, y.ptr, lanes, blocks, xcd, stream())) return false;
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
// Block Lanczos over every sector at once
// ---------------------------------------------------------------------------

// Largest sector diagonalised densely on the host (its real embedding has
// twice as many rows; about 8 (2 d)^3 flops on one host thread).
constexpr int64_t kDenseSectorMax = 512;

// Lowest nev eigenvalues of sector s by dense diagonalisation of the real
// embedding [[A, -B], [B, A]] of its Hermitian matrix M = A + iB (every
// eigenvalue of M appears twice), read from the FP64 CSR.  Eigenvalues past
// the sector's dimension are NaN; residuals are || E z - lambda z || of the
// embedding's eigenvectors, relative to max(1, |lambda|).
bool dense_sector(ls_amd_sector_batch const &bt, int s, int nev, double *evals, double *resid) {
  int64_t const r0 = bt.offsets[s], d = bt.offsets[s + 1] - r0;
  std::fill(evals, evals + nev, NAN);
  std::fill(resid, resid + nev, NAN);
  if (d == 0) return true;
  Csr const &c = bt.csr;
  std::vector<int64_t> rp(static_cast<size_t>(d + 1));
  std::vector<double> diag(static_cast<size_t>(d));
  LSAMD_HIP(hipMemcpyAsync(rp.data(), c.row_ptr.ptr + r0, sizeof(int64_t) * (d + 1), hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipMemcpyAsync(diag.data(), c.diag64.ptr + r0, sizeof(double) * d, hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  int64_t const k0 = rp[0], nz = rp[d] - k0;
  std::vector<uint32_t> cols(static_cast<size_t>(std::max<int64_t>(nz, 1)));
  std::vector<double> vals(static_cast<size_t>(2 * std::max<int64_t>(nz, 1)));
  if (nz > 0) {
    LSAMD_HIP(hipMemcpyAsync(cols.data(), c.cols.ptr + k0, sizeof(uint32_t) * nz, hipMemcpyDeviceToHost, stream()));
    LSAMD_HIP(hipMemcpyAsync(vals.data(), c.vals64.ptr + 2 * k0, sizeof(double) * 2 * nz, hipMemcpyDeviceToHost,
                             stream()));
    LSAMD_HIP(hipStreamSynchronize(stream()));
  }
  int const n2 = static_cast<int>(2 * d);
  std::vector<double> E(static_cast<size_t>(n2) * n2, 0.0);
  auto at = [&](int64_t i, int64_t j) -> double & { return E[static_cast<size_t>(i) * n2 + j]; };
  for (int64_t i = 0; i < d; ++i) {
    at(i, i) += diag[i];
    at(d + i, d + i) += diag[i];
    for (int64_t k = rp[i] - k0; k < rp[i + 1] - k0; ++k) {
      int64_t const j = static_cast<int64_t>(cols[k]) - r0;
      if (j < 0 || j >= d) return fail("sector " + std::to_string(s) + ": a matrix element leaves the sector");
      double const a = vals[2 * k], b = vals[2 * k + 1];
      at(i, j) += a;
      at(d + i, d + j) += a;
      at(i, d + j) -= b;
      at(d + i, j) += b;
    }
  }
  std::vector<double> Es(E.size());
  for (int i = 0; i < n2; ++i)
    for (int j = 0; j < n2; ++j) Es[static_cast<size_t>(i) * n2 + j] = 0.5 * (at(i, j) + at(j, i));
  std::vector<double> w, Z;
  if (!symmetric_eigh(Es, n2, w, Z)) return fail("sector " + std::to_string(s) + ": dense eigensolver did not converge");
  for (int e = 0; e < nev && e < d; ++e) {
    double const lam = 0.5 * (w[2 * e] + w[2 * e + 1]);
    double r2 = 0.0;
    for (int i = 0; i < n2; ++i) {
      double acc = -lam * Z[static_cast<size_t>(i) * n2 + 2 * e];
      for (int j = 0; j < n2; ++j) acc += at(i, j) * Z[static_cast<size_t>(j) * n2 + 2 * e];
      r2 += acc * acc;
    }
    evals[e] = lam;
    resid[e] = std::sqrt(r2) / std::max(1.0, std::fabs(lam));
  }
  return true;
}

// Complex block Lanczos over all sectors: a block is `block` complex columns,
// stored per state as [re | im] (the product's layout), and sector s owns
// rows [offsets[s], offsets[s+1]).  Sectors with fewer states than twice the
// Krylov space, 2 (max_blocks + 1) * block, are diagonalised densely instead
// and left out of the block Lanczos process.
bool sector_eigensolve(ls_amd_sector_batch &bt, BlockKrylovOptions const &opts, BlockKrylovResult &res) {
  auto const t_start = std::chrono::steady_clock::now();
  if (bt.hermitian < 0) {
    // one complex column per state: the product's [re | im] rows interleave
    auto prod = [&](double const *x, double *y) -> bool {
      if (!launch_product(bt, 0, 1, x, y, bt.lanes, bt.blocks, bt.xcd_aware, stream())) return false;
      LSAMD_HIP(hipGetLastError());
      return true;
    };
    double rel = 0.0;
    if (!asymmetry_test(2 * bt.csr.rows, true, prod, 0x53454354ull, rel)) return false;
    bt.hermitian = rel <= 1e-9 ? 1 : 0;
  }
  if (bt.hermitian == 0)
    return fail("a sector's operator is not Hermitian: the block Lanczos solver needs <u, H v> = <H u, v>");
  BlockKrylovOptions o = opts;
  o.complex_values = true;
  int const ns = bt.count;
  int64_t const krylov_min = 2 * static_cast<int64_t>(o.max_blocks + 1) * o.block;
  std::vector<int64_t> const &seg_rows = bt.offsets;
  std::vector<char> active(ns, 1);
  int dense = 0;
  for (int s = 0; s < ns; ++s) {
    int64_t const d = bt.offsets[s + 1] - bt.offsets[s];
    if (d < krylov_min) {
      if (d > kDenseSectorMax)
        return fail("sector " + std::to_string(s) + " (" + std::to_string(d) +
                    " states) is too small for the Krylov space and too large to diagonalise densely");
      active[s] = 0;
      ++dense;
    }
  }
  if (dense < ns) {
    int const B = o.block;
    BlockProduct prod = [&](bool fp32, void const *x, void *y) -> bool {
      if (fp32) return graph_product(bt, 1, B, x, y);
      if (!launch_product(bt, 0, B, x, y, bt.lanes, bt.blocks, bt.xcd_aware, stream())) return false;
      LSAMD_HIP(hipGetLastError());
      return true;
    };
    if (!block_krylov_solve(seg_rows, active, 1, prod, o, res, nullptr)) return false;
  } else {
    // nothing for the Krylov process; the options are still checked alike
    if (!check_block_krylov_options(o, 1)) return false;
    res = BlockKrylovResult{};
    res.converged = true;
    res.evals.assign(static_cast<size_t>(ns) * o.nev, NAN);
    res.residuals.assign(static_cast<size_t>(ns) * o.nev, NAN);
  }
  for (int s = 0; s < ns; ++s) {
    if (active[s]) continue;
    double *ev = res.evals.data() + static_cast<size_t>(s) * o.nev;
    double *rr = res.residuals.data() + static_cast<size_t>(s) * o.nev;
    if (!dense_sector(bt, s, o.nev, ev, rr)) return false;
    for (int e = 0; e < o.nev; ++e)
      if (std::isfinite(rr[e])) res.max_residual = std::max(res.max_residual, rr[e]);
  }
  res.total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
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
                                              ls_amd_eig_info *info) {
  if (b == nullptr || evals == nullptr) return fail("null argument") ? 0 : -1;
  BlockKrylovOptions o;
  o.block = block_size;
  o.nev = nev;
  o.max_blocks = max_blocks;
  o.keep = keep_blocks;
  o.max_cycles = max_cycles;
  o.tol = tol;
  o.switch_tol = switch_tol;
  o.seed = seed;
  BlockKrylovResult r;
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
