// spmv_kernels.hpp -- the CSR product kernels (FP64 single vector and blocks
// of right-hand sides) and their row-tile schedules, shared by the engine and
// by the measurement probes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsamd {

// Value table: a real CSR with at most kValueTableSize distinct off-diagonal
// values and fewer than 2^24 rows stores column | value index << 24 per entry.
constexpr uint32_t kPackedColumnMask = (1u << 24) - 1;
constexpr int kValueTableSize = 256;
// Wide value table: at most kWideTableSize distinct values (any number of
// rows) store the plain column and a 16-bit value index per entry; the table
// is staged in LDS (dynamic shared memory of its own size, at most 32 KB in
// FP64).
constexpr int kWideTableSize = 4096;

// Stream loads of the CSR arrays, optionally non-temporal so that the
// stream does not displace the gathered vector from L2.
template <bool NT, typename T> __device__ __forceinline__ T ld_stream(T const *p) {
  if constexpr (NT) {
    return __builtin_nontemporal_load(p);
  } else {
    return *p;
  }
}

// Row tiles of 256 / LPR rows.  Die-aware (DIE): the grid is a fixed multiple
// of 8 workgroups; workgroups are dealt to the 8 accelerator dies round-robin,
// and those of die d walk the contiguous tile range [tiles d / 8, tiles (d+1)
// / 8), so rows that gather from nearby columns share one die's L2.
// Otherwise one workgroup per tile.
template <bool DIE> __device__ __forceinline__ void tile_range(int64_t tiles, int64_t &t, int64_t &t_end,
                                                              int64_t &t_step) {
  if constexpr (DIE) {
    int const die = static_cast<int>(blockIdx.x) % 8;
    t = tiles * die / 8 + static_cast<int64_t>(blockIdx.x) / 8;
    t_end = tiles * (die + 1) / 8;
    t_step = static_cast<int64_t>(gridDim.x) / 8;
  } else {
    t = blockIdx.x;
    t_end = t < tiles ? t + 1 : t;
    t_step = 1;
  }
}

// Readers of stored entry k: (column, value) from separate arrays, from a
// packed column word and the value table, or from the column, a 16-bit
// index and the wide value table.  The tables are in LDS.
template <bool NT, typename T> struct PlainEntries {
  uint32_t const *__restrict__ cols;
  T const *__restrict__ vals;
  __device__ __forceinline__ void operator()(int64_t k, uint32_t &col, T &v) const {
    col = ld_stream<NT>(cols + k);
    v = ld_stream<NT>(vals + k);
  }
};
template <bool NT, typename T> struct PackedEntries {
  uint32_t const *__restrict__ cols;
  T const *__restrict__ table;
  __device__ __forceinline__ void operator()(int64_t k, uint32_t &col, T &v) const {
    uint32_t const p = ld_stream<NT>(cols + k);
    col = p & kPackedColumnMask;
    v = table[p >> 24];
  }
};
template <bool NT, typename T> struct WideEntries {
  uint32_t const *__restrict__ cols;
  uint16_t const *__restrict__ vidx;
  T const *__restrict__ table;
  __device__ __forceinline__ void operator()(int64_t k, uint32_t &col, T &v) const {
    col = ld_stream<NT>(cols + k);
    v = table[ld_stream<NT>(vidx + k)];
  }
};

// Stages `size` table values in LDS for the workgroup.
template <typename T> __device__ __forceinline__ T const *stage_table(T *lds, T const *__restrict__ table, int size) {
  for (int i = threadIdx.x; i < size; i += blockDim.x) lds[i] = table[i];
  __syncthreads();
  return lds;
}

// y[row] = diag[row] x[row] (or y[row], or 0) + sum of the row's entries
// v x[col], in entry order, LPR lanes to a row.
template <int LPR, bool DIE, typename Entries>
__device__ __forceinline__ void csr_f64_rows(int64_t n, int64_t const *__restrict__ rp, Entries const &ent,
                                             double const *__restrict__ diag, double const *__restrict__ x,
                                             double *__restrict__ y, int accumulate, int has_diag) {
  constexpr int kRows = 256 / LPR;
  int64_t const tiles = (n + kRows - 1) / kRows;
  int64_t t, t_end, t_step;
  tile_range<DIE>(tiles, t, t_end, t_step);
  int const sub = static_cast<int>(threadIdx.x) % LPR;
  for (; t < t_end; t += t_step) {
    int64_t const row = t * kRows + static_cast<int>(threadIdx.x) / LPR;
    double acc = 0.0;
    if (row < n) {
      int64_t const b = rp[row], e = rp[row + 1];
      int64_t k = b + sub;
      // four entries per step: their gathers are in flight together; the
      // sum keeps the entries' order
      for (; k + 3 * LPR < e; k += 4 * LPR) {
        uint32_t c0, c1, c2, c3;
        double v0, v1, v2, v3;
        ent(k, c0, v0);
        ent(k + LPR, c1, v1);
        ent(k + 2 * LPR, c2, v2);
        ent(k + 3 * LPR, c3, v3);
        double const x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc = fma(v0, x0, acc);
        acc = fma(v1, x1, acc);
        acc = fma(v2, x2, acc);
        acc = fma(v3, x3, acc);
      }
      for (; k < e; k += LPR) {
        uint32_t col;
        double v;
        ent(k, col, v);
        acc = fma(v, x[col], acc);
      }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (row < n && sub == 0) {
      double const base = has_diag ? diag[row] * x[row] : (accumulate ? y[row] : 0.0);
      y[row] = base + acc;
    }
  }
}

template <int LPR, bool NT, bool DIE, bool PACKED>
__global__ __launch_bounds__(256) void csr_f64_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                      uint32_t const *__restrict__ cols,
                                                      double const *__restrict__ vals,
                                                      double const *__restrict__ table, int table_size,
                                                      double const *__restrict__ diag,
                                                      double const *__restrict__ x,
                                                      double *__restrict__ y, int accumulate,
                                                      int has_diag) {
  if constexpr (PACKED) {
    __shared__ double lds_table[kValueTableSize];
    PackedEntries<NT, double> const ent{cols, stage_table(lds_table, table, table_size)};
    csr_f64_rows<LPR, DIE>(n, rp, ent, diag, x, y, accumulate, has_diag);
  } else {
    PlainEntries<NT, double> const ent{cols, vals};
    csr_f64_rows<LPR, DIE>(n, rp, ent, diag, x, y, accumulate, has_diag);
  }
}

template <int LPR, bool NT, bool DIE>
__global__ __launch_bounds__(256) void csr_f64_wide_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                           uint32_t const *__restrict__ cols,
                                                           uint16_t const *__restrict__ vidx,
                                                           double const *__restrict__ table, int table_size,
                                                           double const *__restrict__ diag,
                                                           double const *__restrict__ x,
                                                           double *__restrict__ y, int accumulate,
                                                           int has_diag) {
  extern __shared__ unsigned char wide_lds[]; // table_size doubles (dynamic)
  WideEntries<NT, double> const ent{cols, vidx,
                                    stage_table(reinterpret_cast<double *>(wide_lds), table, table_size)};
  csr_f64_rows<LPR, DIE>(n, rp, ent, diag, x, y, accumulate, has_diag);
}

// y[row][q] = diag[row] x[row][q] + sum of the row's entries v x[col][q] for
// K right-hand sides stored [rows][K], summed in T.
template <int K, int LPR, typename T, typename TO, bool DIE, typename Entries>
__device__ __forceinline__ void csr_block_rows(int64_t n, int64_t const *__restrict__ rp, Entries const &ent,
                                               T const *__restrict__ diag, T const *__restrict__ x,
                                               TO *__restrict__ y) {
  constexpr int kRows = 256 / LPR;
  int64_t const tiles = (n + kRows - 1) / kRows;
  int64_t t, t_end, t_step;
  tile_range<DIE>(tiles, t, t_end, t_step);
  int const sub = static_cast<int>(threadIdx.x) % LPR;
  for (; t < t_end; t += t_step) {
    int64_t const row = t * kRows + static_cast<int>(threadIdx.x) / LPR;
    T acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = T(0);
    if (row < n) {
      int64_t const b = rp[row], e = rp[row + 1];
      int64_t k = b + sub;
      // two entries per step, their gathers in flight together, summed in order
      for (; k + LPR < e; k += 2 * LPR) {
        uint32_t c0, c1;
        T v0, v1;
        ent(k, c0, v0);
        ent(k + LPR, c1, v1);
        T const *x0 = x + static_cast<size_t>(c0) * K;
        T const *x1 = x + static_cast<size_t>(c1) * K;
        T a0[K], a1[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
          a0[q] = x0[q];
          a1[q] = x1[q];
        }
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] = fma(v0, a0[q], acc[q]);
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] = fma(v1, a1[q], acc[q]);
      }
      for (; k < e; k += LPR) {
        uint32_t col;
        T v;
        ent(k, col, v);
        T const *xr = x + static_cast<size_t>(col) * K;
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] = fma(v, xr[q], acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off, LPR);
    if (row < n && sub == 0) {
      T const d = diag[row];
#pragma unroll
      for (int q = 0; q < K; ++q)
        y[static_cast<size_t>(row) * K + q] = static_cast<TO>(d * x[static_cast<size_t>(row) * K + q] + acc[q]);
    }
  }
}

template <int K, int LPR, typename T, typename TO, bool NT, bool DIE, bool PACKED>
__global__ __launch_bounds__(256) void csr_block_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                        uint32_t const *__restrict__ cols,
                                                        T const *__restrict__ vals,
                                                        T const *__restrict__ table, int table_size,
                                                        T const *__restrict__ diag,
                                                        T const *__restrict__ x,
                                                        TO *__restrict__ y) {
  if constexpr (PACKED) {
    __shared__ T lds_table[kValueTableSize];
    PackedEntries<NT, T> const ent{cols, stage_table(lds_table, table, table_size)};
    csr_block_rows<K, LPR, T, TO, DIE>(n, rp, ent, diag, x, y);
  } else {
    PlainEntries<NT, T> const ent{cols, vals};
    csr_block_rows<K, LPR, T, TO, DIE>(n, rp, ent, diag, x, y);
  }
}

template <int K, int LPR, typename T, typename TO, bool NT, bool DIE>
__global__ __launch_bounds__(256) void csr_block_wide_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                             uint32_t const *__restrict__ cols,
                                                             uint16_t const *__restrict__ vidx,
                                                             T const *__restrict__ table, int table_size,
                                                             T const *__restrict__ diag,
                                                             T const *__restrict__ x,
                                                             TO *__restrict__ y) {
  extern __shared__ unsigned char wide_lds[]; // table_size values of T (dynamic)
  WideEntries<NT, T> const ent{cols, vidx, stage_table(reinterpret_cast<T *>(wide_lds), table, table_size)};
  csr_block_rows<K, LPR, T, TO, DIE>(n, rp, ent, diag, x, y);
}

} // namespace lsamd
