// probes.hip -- libprobes.so, measurement kernels for bench.py only (the
// engine, liblsamd.so, does not contain them):
//
// lsamd_probe_gather: the engine's FP64 CSR product kernel on a synthetic
//   matrix of `rows` rows x `per_row` entries whose rows gather from windows
//   of `window` columns (window = rows: the whole vector), plain or in the
//   value-table format, with `lanes` per row and the schedule flags of
//   ls_amd_operator_set_schedule; *ms_out is the mean duration of `reps`
//   launches.  Small windows keep the gathered vector in L2, so the best time
//   bounds what any ordering of a real matrix of that shape could reach with
//   one gather per nonzero.
// lsamd_probe_copy: a plain device copy of `bytes` bytes, `reps` times, whose
//   known traffic scales the profiler's FETCH_SIZE / WRITE_SIZE counters.
// lsamd_probe_l2_model (host code): an LRU model of the L2 of each
//   accelerator die (l2_bytes, `ways`-way, line_bytes lines) fed the x
//   gathers of a CSR product in which die d processes rows
//   [n d / dies, n (d + 1) / dies) in order; returns the misses and
//   accesses of all dies.
//
// Both return 0, or -1 with a message in lsamd_probe_error().
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "spmv_kernels.hpp"

namespace {

thread_local std::string g_error;

int failed(std::string const &msg) {
  g_error = msg;
  return -1;
}

#define PROBE_HIP(call)                                                                      \
  do {                                                                                       \
    hipError_t const _e = (call);                                                            \
    if (_e != hipSuccess) return failed(std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

__global__ __launch_bounds__(256) void probe_rows_kernel(int64_t rows, int per_row, int64_t *__restrict__ rp) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i <= rows) rp[i] = i * per_row;
}

// Column k of the synthetic matrix: a splitmix64 hash of k placed in the
// window of its 65 536-row block (the whole vector when window = rows).
__global__ __launch_bounds__(256) void probe_cols_kernel(int64_t rows, int per_row, int64_t window, int packed,
                                                         uint32_t *__restrict__ cols) {
  int64_t const k = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= rows * per_row) return;
  uint64_t z = 0x9e3779b97f4a7c15ull * static_cast<uint64_t>(k + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  int64_t const row = k / per_row;
  uint64_t const span = static_cast<uint64_t>(rows - window);
  uint64_t const base = window >= rows ? 0 : (static_cast<uint64_t>(row / 65536) * 2654435761ull) % (span + 1);
  uint32_t const col = static_cast<uint32_t>(base + z % static_cast<uint64_t>(window));
  cols[k] = packed ? (col | (static_cast<uint32_t>(z >> 40) & 0x3Fu) << 24) : col;
}

__global__ __launch_bounds__(256) void copy_kernel(uint4 const *__restrict__ src, uint4 *__restrict__ dst,
                                                   int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * 256)
    dst[i] = src[i];
}

int grid_of(int64_t n) { return static_cast<int>((n + 255) / 256); }

// Device buffers and the stream of one probe call, released on every return.
struct Buffers {
  std::vector<void *> ptrs;
  hipStream_t stream = nullptr;
  ~Buffers() {
    if (stream != nullptr) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T> int get(T *&out, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) return failed("hipMalloc failed");
    ptrs.push_back(p);
    out = static_cast<T *>(p);
    return 0;
  }
};

template <int LPR, bool NT, bool DIE, bool PACKED>
void launch(int grid, hipStream_t s, int64_t n, int64_t const *rp, uint32_t const *cols, double const *vals,
            double const *table, int table_size, double const *diag, double const *x, double *y) {
  lsamd::csr_f64_kernel<LPR, NT, DIE, PACKED><<<grid, 256, 0, s>>>(n, rp, cols, vals, table, table_size, diag, x, y,
                                                                  0, 1);
}

template <int LPR>
void launch_flags(int flags, int packed, int grid, hipStream_t s, int64_t n, int64_t const *rp, uint32_t const *cols,
                  double const *vals, double const *table, double const *diag, double const *x, double *y) {
  bool const nt = flags & 1, die = flags & 2;
  auto go = [&]<bool NT, bool DIE>() {
    if (packed)
      launch<LPR, NT, DIE, true>(grid, s, n, rp, cols, nullptr, table, 64, diag, x, y);
    else
      launch<LPR, NT, DIE, false>(grid, s, n, rp, cols, vals, nullptr, 0, diag, x, y);
  };
  if (nt && die) go.template operator()<true, true>();
  else if (nt) go.template operator()<true, false>();
  else if (die) go.template operator()<false, true>();
  else go.template operator()<false, false>();
}

} // namespace

extern "C" __attribute__((visibility("default"))) char const *lsamd_probe_error(void) { return g_error.c_str(); }

extern "C" __attribute__((visibility("default"))) int lsamd_probe_gather(int64_t rows, int per_row, int64_t window, int lanes, int flags, int packed,
                                  int reps, double *ms_out) {
  if (ms_out == nullptr || rows <= 0 || per_row <= 0 || window <= 0 || window > rows || reps <= 0)
    return failed("lsamd_probe_gather: bad arguments");
  if (flags < 0 || flags > 7 || flags == 4 || flags == 5) return failed("schedule flags are 0..3, 6 or 7");
  if (lanes != 8 && lanes != 16 && lanes != 32) return failed("lanes per row must be 8, 16 or 32");
  if (rows >= (int64_t{1} << 24) && packed) return failed("the packed format needs rows < 2^24");
  if (rows >= (int64_t{1} << 32) - 1) return failed("too many rows");
  Buffers buf;
  PROBE_HIP(hipSetDevice(0));
  PROBE_HIP(hipStreamCreateWithFlags(&buf.stream, hipStreamNonBlocking));
  hipStream_t const s = buf.stream;
  int64_t const nnz = rows * per_row;
  int64_t *rp = nullptr;
  uint32_t *cols = nullptr;
  double *vals = nullptr, *diag = nullptr, *x = nullptr, *y = nullptr, *table = nullptr;
  if (buf.get(rp, rows + 1) || buf.get(cols, nnz) || buf.get(diag, rows) || buf.get(x, rows) || buf.get(y, rows) ||
      buf.get(table, 64) || (!packed && buf.get(vals, nnz)))
    return -1;
  probe_rows_kernel<<<grid_of(rows + 1), 256, 0, s>>>(rows, per_row, rp);
  probe_cols_kernel<<<grid_of(nnz), 256, 0, s>>>(rows, per_row, window, packed, cols);
  PROBE_HIP(hipGetLastError());
  PROBE_HIP(hipMemsetAsync(x, 0, sizeof(double) * rows, s));
  PROBE_HIP(hipMemsetAsync(diag, 0, sizeof(double) * rows, s));
  std::vector<double> t(64);
  for (int i = 0; i < 64; ++i) t[i] = 1.0 + i;
  PROBE_HIP(hipMemcpyAsync(table, t.data(), sizeof(double) * 64, hipMemcpyHostToDevice, s));
  if (!packed) PROBE_HIP(hipMemsetAsync(vals, 0, sizeof(double) * nnz, s));
  int device = 0, cus = 256;
  if (hipGetDevice(&device) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  int const tiles = grid_of(rows * lanes);
  int const grid = (flags & 2) ? ((flags & 4) ? (tiles + 7) / 8 * 8 : (cus * 4 / 8) * 8) : tiles;
  auto product = [&]() {
    switch (lanes) {
    case 8: launch_flags<8>(flags, packed, grid, s, rows, rp, cols, vals, table, diag, x, y); break;
    case 16: launch_flags<16>(flags, packed, grid, s, rows, rp, cols, vals, table, diag, x, y); break;
    default: launch_flags<32>(flags, packed, grid, s, rows, rp, cols, vals, table, diag, x, y); break;
    }
  };
  product();
  PROBE_HIP(hipGetLastError());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PROBE_HIP(hipEventCreate(&e0));
  PROBE_HIP(hipEventCreate(&e1));
  PROBE_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) product();
  PROBE_HIP(hipEventRecord(e1, s));
  PROBE_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  PROBE_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  PROBE_HIP(hipStreamSynchronize(s));
  *ms_out = static_cast<double>(ms) / reps;
  return 0;
}

extern "C" __attribute__((visibility("default"))) int lsamd_probe_copy(int64_t bytes, int reps, double *ms_out) {
  if (bytes <= 0 || bytes % 16 != 0 || reps <= 0 || ms_out == nullptr)
    return failed("lsamd_probe_copy: bytes must be a positive multiple of 16");
  Buffers buf;
  PROBE_HIP(hipSetDevice(0));
  PROBE_HIP(hipStreamCreateWithFlags(&buf.stream, hipStreamNonBlocking));
  hipStream_t const s = buf.stream;
  uint4 *src = nullptr, *dst = nullptr;
  int64_t const n = bytes / 16;
  if (buf.get(src, n) || buf.get(dst, n)) return -1;
  PROBE_HIP(hipMemsetAsync(src, 1, static_cast<size_t>(bytes), s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PROBE_HIP(hipEventCreate(&e0));
  PROBE_HIP(hipEventCreate(&e1));
  PROBE_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) copy_kernel<<<2048, 256, 0, s>>>(src, dst, n);
  PROBE_HIP(hipGetLastError());
  PROBE_HIP(hipEventRecord(e1, s));
  PROBE_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  PROBE_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  PROBE_HIP(hipStreamSynchronize(s));
  *ms_out = static_cast<double>(ms) / reps;
  return 0;
}

extern "C" __attribute__((visibility("default"))) int lsamd_probe_l2_model(int64_t n, int64_t const *rp,
                                                                         uint32_t const *cols, int dies,
                                                                         int64_t l2_bytes, int ways, int line_bytes,
                                                                         int64_t *misses_out, int64_t *accesses_out) {
  if (n < 0 || rp == nullptr || cols == nullptr || dies < 1 || dies > 64 || ways < 1 || line_bytes < 8 ||
      (line_bytes & (line_bytes - 1)) != 0 || l2_bytes < static_cast<int64_t>(ways) * line_bytes)
    return failed("lsamd_probe_l2_model: bad arguments");
  int64_t const sets = l2_bytes / line_bytes / ways;
  int shift = 0;
  while ((8 << shift) < line_bytes) ++shift; // x entries are 8 B
  std::vector<int64_t> misses(dies, 0), accesses(dies, 0);
  std::vector<std::thread> pool;
  for (int d = 0; d < dies; ++d)
    pool.emplace_back([&, d]() {
      std::vector<uint64_t> tag(static_cast<size_t>(sets) * ways, ~uint64_t{0}), age(static_cast<size_t>(sets) * ways, 0);
      uint64_t clock = 0;
      int64_t m = 0, a = 0;
      for (int64_t r = n * d / dies; r < n * (d + 1) / dies; ++r)
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
          uint64_t const line = static_cast<uint64_t>(cols[k]) >> shift;
          // hashed set index, as a real cache avoids power-of-two aliasing
          int64_t const set = static_cast<int64_t>(((line * 0x9e3779b97f4a7c15ull) >> 40) % static_cast<uint64_t>(sets));
          uint64_t *t = tag.data() + set * ways, *g = age.data() + set * ways;
          ++clock;
          ++a;
          int hit = -1, victim = 0;
          for (int w = 0; w < ways; ++w) {
            if (t[w] == line) {
              hit = w;
              break;
            }
            if (g[w] < g[victim]) victim = w;
          }
          if (hit >= 0) {
            g[hit] = clock;
          } else {
            ++m;
            t[victim] = line;
            g[victim] = clock;
          }
        }
      misses[d] = m;
      accesses[d] = a;
    });
  for (auto &t : pool) t.join();
  int64_t mm = 0, aa = 0;
  for (int d = 0; d < dies; ++d) {
    mm += misses[d];
    aa += accesses[d];
  }
  if (misses_out != nullptr) *misses_out = mm;
  if (accesses_out != nullptr) *accesses_out = aa;
  return 0;
}
