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

// One stored entry: (column, value), from separate arrays or from a packed
// column word and the value table in LDS.
template <bool PACKED, bool NT, typename T>
__device__ __forceinline__ void csr_entry(uint32_t const *__restrict__ cols, T const *__restrict__ vals,
                                          T const *__restrict__ table, int64_t k, uint32_t &col, T &v) {
  if constexpr (PACKED) {
    uint32_t const p = ld_stream<NT>(cols + k);
    col = p & kPackedColumnMask;
    v = table[p >> 24];
  } else {
    col = ld_stream<NT>(cols + k);
    v = ld_stream<NT>(vals + k);
  }
}

// Stages the value table in LDS (a no-op for unpacked matrices).
template <bool PACKED, typename T>
__device__ __forceinline__ T const *stage_table(T *lds, T const *__restrict__ table, int size) {
  if constexpr (PACKED) {
    for (int i = threadIdx.x; i < size; i += blockDim.x) lds[i] = table[i];
    __syncthreads();
    return lds;
  } else {
    return nullptr;
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
  __shared__ double lds_table[PACKED ? kValueTableSize : 1];
  double const *tab = stage_table<PACKED>(lds_table, table, table_size);
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
        csr_entry<PACKED, NT>(cols, vals, tab, k, c0, v0);
        csr_entry<PACKED, NT>(cols, vals, tab, k + LPR, c1, v1);
        csr_entry<PACKED, NT>(cols, vals, tab, k + 2 * LPR, c2, v2);
        csr_entry<PACKED, NT>(cols, vals, tab, k + 3 * LPR, c3, v3);
        double const x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc = fma(v0, x0, acc);
        acc = fma(v1, x1, acc);
        acc = fma(v2, x2, acc);
        acc = fma(v3, x3, acc);
      }
      for (; k < e; k += LPR) {
        uint32_t col;
        double v;
        csr_entry<PACKED, NT>(cols, vals, tab, k, col, v);
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

template <int K, int LPR, typename T, typename TO, bool NT, bool DIE, bool PACKED>
__global__ __launch_bounds__(256) void csr_block_kernel(int64_t n, int64_t const *__restrict__ rp,
                                                        uint32_t const *__restrict__ cols,
                                                        T const *__restrict__ vals,
                                                        T const *__restrict__ table, int table_size,
                                                        T const *__restrict__ diag,
                                                        T const *__restrict__ x,
                                                        TO *__restrict__ y) {
  __shared__ T lds_table[PACKED ? kValueTableSize : 1];
  T const *tab = stage_table<PACKED>(lds_table, table, table_size);
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
        csr_entry<PACKED, NT>(cols, vals, tab, k, c0, v0);
        csr_entry<PACKED, NT>(cols, vals, tab, k + LPR, c1, v1);
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
        csr_entry<PACKED, NT>(cols, vals, tab, k, col, v);
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

} // namespace lsamd
