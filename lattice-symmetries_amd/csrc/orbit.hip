// orbit.hip -- state_info / is_representative / state index kernels.
//
// One lane per input state; element loops are wave-uniform (see
// orbit_device.hpp).  These kernels back the reference-named entry points
// ls_hs_state_info_halide_kernel, ls_hs_is_representative_halide_kernel and
// ls_hs_state_index_binary_search_kernel as well as the device-array
// variants of the additive ABI.
#include <algorithm>
#include <cstdlib>

#include "orbit_device.hpp"

namespace lsamd {

// FACT: through the translation factorisation of the group (state_info_factored,
// bit for bit the same results) instead of element by element.
template <typename W, int D, int SI, bool FACT>
__global__ __launch_bounds__(256) void state_info_kernel(GroupView g, int64_t count,
                                                         uint64_t const *__restrict__ xs,
                                                         int64_t x_stride,
                                                         uint64_t *__restrict__ reps,
                                                         int64_t rep_stride,
                                                         double *__restrict__ chars,
                                                         double *__restrict__ norms) {
  __shared__ RowTables rt;
  if constexpr (FACT) row_tables_init(g, rt);
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  StateInfo s;
  if constexpr (FACT)
    s = state_info_factored<W, D, SI>(g, xs[i * x_stride], &rt);
  else
    s = state_info_one<W, D, SI>(g, xs[i * x_stride]);
  reps[i * rep_stride] = s.rep;
  chars[2 * i] = s.c_re;
  chars[2 * i + 1] = s.c_im;
  norms[i] = s.norm;
}

template <typename W, int D, int SI>
__global__ __launch_bounds__(256) void is_representative_kernel(GroupView g, int64_t count,
                                                                uint64_t const *__restrict__ xs,
                                                                int64_t x_stride,
                                                                uint8_t *__restrict__ flags,
                                                                double *__restrict__ norms) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint8_t f;
  double n;
  is_representative_one<W, D, SI>(g, xs[i * x_stride], f, n);
  flags[i] = f;
  norms[i] = n;
}

bool launch_state_info(GroupData const &g, int64_t count, uint64_t const *d_x, int64_t x_stride,
                       uint64_t *d_reps, int64_t rep_stride, double *d_chars, double *d_norms) {
  if (count <= 0) return true;
  GroupView const v = g.view();
  return dispatch_group(g, [&]<typename W, int D, int SI>() -> bool {
    if constexpr (SI != kFermionic) {
      if (g.factored()) {
        state_info_kernel<W, D, SI, true><<<grid_for(count, 256), 256, 0, stream()>>>(
            v, count, d_x, x_stride, d_reps, rep_stride, d_chars, d_norms);
        LSAMD_HIP(hipGetLastError());
        return true;
      }
    }
    state_info_kernel<W, D, SI, false><<<grid_for(count, 256), 256, 0, stream()>>>(
        v, count, d_x, x_stride, d_reps, rep_stride, d_chars, d_norms);
    LSAMD_HIP(hipGetLastError());
    return true;
  });
}

bool launch_is_representative(GroupData const &g, int64_t count, uint64_t const *d_x,
                              int64_t x_stride, uint8_t *d_flags, double *d_norms) {
  if (count <= 0) return true;
  GroupView const v = g.view();
  return dispatch_group(g, [&]<typename W, int D, int SI>() -> bool {
    is_representative_kernel<W, D, SI><<<grid_for(count, 256), 256, 0, stream()>>>(
        v, count, d_x, x_stride, d_flags, d_norms);
    LSAMD_HIP(hipGetLastError());
    return true;
  });
}

// ---------------------------------------------------------------------------
// State index
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void state_index_kernel(IndexView idx, int64_t count,
                                                          uint64_t const *__restrict__ xs,
                                                          int64_t *__restrict__ out) {
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = index_lookup(idx, xs[i]);
}

bool launch_state_index(IndexView const &idx, int64_t count, uint64_t const *d_x, int64_t *d_out) {
  if (count <= 0) return true;
  state_index_kernel<<<grid_for(count, 256), 256, 0, stream()>>>(idx, count, d_x, d_out);
  LSAMD_HIP(hipGetLastError());
  return true;
}

// offsets[b] = first position whose state has prefix >= b (b = 0 .. 2^P).
__global__ __launch_bounds__(256) void prefix_table_kernel(uint64_t const *__restrict__ reps,
                                                           int64_t count, int shift,
                                                           int64_t table_size,
                                                           uint32_t *__restrict__ offsets) {
  int64_t const b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > table_size) return;
  int64_t lo = 0, n = count;
  while (n > 0) {
    int64_t const half = n >> 1;
    bool const less = (reps[lo + half] >> shift) < static_cast<uint64_t>(b);
    lo = less ? lo + half + 1 : lo;
    n = less ? n - half - 1 : half;
  }
  offsets[b] = static_cast<uint32_t>(lo);
}

int default_prefix_bits(int number_bits, int64_t count) {
  // LSAMD_INDEX_PREFIX_BITS (8..30) overrides; otherwise about four table
  // entries per state, at least 8 and at most 26 bits
  int p = 8;
  if (char const *e = std::getenv("LSAMD_INDEX_PREFIX_BITS")) {
    p = std::max(8, std::min(30, std::atoi(e)));
  } else {
    while (p < 26 && (int64_t{1} << p) < 4 * count) ++p;
  }
  return std::min(p, std::max(number_bits, 1));
}

bool index_build(IndexData &idx, uint64_t const *d_reps, int64_t count, int number_bits,
                 int prefix_bits) {
  if (count >= (int64_t{1} << 32) - 1) return fail("state index supports fewer than 2^32 states");
  if (number_bits <= 0 || number_bits > 64) return fail("invalid bits");
  if (prefix_bits > number_bits) prefix_bits = number_bits;
  if (prefix_bits < 1) prefix_bits = 1;
  idx.number_states = count;
  idx.number_bits = number_bits;
  idx.prefix_bits = prefix_bits;
  idx.shift = number_bits - prefix_bits;
  idx.d_reps = d_reps;
  int64_t const table_size = int64_t{1} << prefix_bits;
  if (!idx.d_offsets.allocate(static_cast<size_t>(table_size + 1))) return false;
  prefix_table_kernel<<<grid_for(table_size + 1, 256), 256, 0, stream()>>>(
      d_reps, count, idx.shift, table_size, idx.d_offsets.ptr);
  LSAMD_HIP(hipGetLastError());
  LSAMD_HIP(hipStreamSynchronize(stream()));
  return true;
}

// ---------------------------------------------------------------------------
// Combinadics (Basis.hs:487-545)
// ---------------------------------------------------------------------------
namespace {
struct BinomialTable {
  uint64_t c[65][65];
  BinomialTable() {
    for (int n = 0; n <= 64; ++n)
      for (int k = 0; k <= 64; ++k) c[n][k] = 0;
    for (int n = 0; n <= 64; ++n) {
      c[n][0] = 1;
      for (int k = 1; k <= n; ++k) c[n][k] = c[n - 1][k - 1] + (k <= n - 1 ? c[n - 1][k] : 0);
    }
  }
};
BinomialTable const g_binomials;
DeviceArray<uint64_t> g_binomials_device;
std::mutex g_binomials_mutex;
} // namespace

uint64_t binomial_host(int n, int k) {
  if (n <= 0 || k > n || k < 0 || n > 64) return 0;
  return g_binomials.c[n][k];
}

uint64_t const *binomials_device() {
  std::lock_guard<std::mutex> lock(g_binomials_mutex);
  if (g_binomials_device.ptr == nullptr) {
    uint64_t flat[65 * 65];
    for (int n = 0; n <= 64; ++n)
      for (int k = 0; k <= 64; ++k) flat[n * 65 + k] = binomial_host(n, k);
    if (!g_binomials_device.upload(flat, 65 * 65)) return nullptr;
  }
  return g_binomials_device.ptr;
}

uint64_t fixed_hamming_state_to_index(uint64_t x) {
  uint64_t i = 0;
  int k = 1;
  while (x != 0) {
    int const c = __builtin_ctzll(x);
    x &= x - 1;
    i += binomial_host(c, k);
    ++k;
  }
  return i;
}

uint64_t fixed_hamming_index_to_state(uint64_t index, int h) {
  uint64_t state = 0;
  int c = 63;
  for (int i = h; i > 0; --i) {
    while (c >= i && binomial_host(c, i) > index) --c;
    if (c < i - 1) c = i - 1;
    state |= uint64_t{1} << c;
    index -= binomial_host(c, i);
    --c;
  }
  return state;
}

} // namespace lsamd
