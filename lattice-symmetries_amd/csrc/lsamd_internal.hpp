// lsamd_internal.hpp -- shared declarations of the MI355X engine.
//
// Host objects (group, index, basis, operator), error plumbing, and the
// device views passed to kernels by value.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <functional>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "lsamd.h"

namespace lsamd {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

// Records `msg` as the last error (ls_amd_last_error).  Returns false so
// callers can `return fail(...)`.  The reference-named entry points pass the
// message on to ls_hs_error, the settable handler of reference.c:17-36.
bool fail(std::string const &msg);
void set_last_error(std::string const &msg);

#define LSAMD_HIP(call)                                                        \
  do {                                                                         \
    hipError_t const _e = (call);                                              \
    if (_e != hipSuccess) {                                                    \
      return ::lsamd::fail(std::string("HIP error: ") + hipGetErrorString(_e) + \
                           " at " + __FILE__ + ":" + std::to_string(__LINE__));  \
    }                                                                          \
  } while (0)

// Device set-up; returns false when no usable device.
bool ensure_device();
hipStream_t stream();

// Copies between pageable host memory and device memory through pinned
// staging buffers (staging.cpp).  copy_to_device: on return the data is in
// device memory, ordered before later work on stream().  copy_to_host: waits
// for the work queued on stream(), returns when h holds the data.
bool copy_to_device(void *d, void const *h, size_t bytes);
bool copy_to_host(void *h, void const *d, size_t bytes);
// Copies `bytes` of d to h while they are produced: produce(first, count)
// enqueues on stream() the work that writes bytes [first, first + count) of
// d, for consecutive chunks (multiples of `unit` bytes), and each chunk goes
// to the host as soon as its work has run.  Without the staging buffers all
// chunks are produced first and copied in one piece.
bool produce_to_host(void *h, void const *d, size_t bytes, size_t unit,
                     std::function<bool(size_t, size_t)> const &produce);

// Kernel timing by family (see ls_amd_last_kernel_ms).
enum TimerFamily {
  kTimerCsrF64 = 0,
  kTimerCsrF32 = 1,
  kTimerOtfF64 = 2,
  kTimerEnumerate = 3,
  kTimerCsrBuild = 4,
  kTimerCsrF64Block = 5,
  kTimerSectorProduct = 6,
  kTimerCount = 7
};
void timer_begin(int family);
void timer_end(int family);

// RAII device buffer.
template <typename T> struct DeviceArray {
  T *ptr = nullptr;
  size_t count = 0;
  DeviceArray() = default;
  DeviceArray(DeviceArray const &) = delete;
  DeviceArray &operator=(DeviceArray const &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : ptr(o.ptr), count(o.count) {
    o.ptr = nullptr;
    o.count = 0;
  }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) {
      reset();
      ptr = o.ptr;
      count = o.count;
      o.ptr = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~DeviceArray() { reset(); }
  bool allocate(size_t n) {
    reset();
    if (n == 0) return true;
    if (hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)) != hipSuccess) {
      ptr = nullptr;
      return fail("hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    }
    count = n;
    return true;
  }
  bool upload(T const *host, size_t n) {
    if (!allocate(n)) return false;
    return copy_to_device(ptr, host, n * sizeof(T));
  }
  void reset() {
    if (ptr != nullptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
};

// ---------------------------------------------------------------------------
// Device views
// ---------------------------------------------------------------------------

// Permutation group on the device.  masks are element-major:
// masks[e * depth + k] is stage k of the Benes network of element e.
struct GroupView {
  uint64_t const *masks;
  double const *eig_re;
  double const *eig_im;
  uint64_t const *shifts; // used only by the runtime-depth kernels
  int size;               // number of group elements (permutations)
  int depth;              // Benes stages
  uint64_t flip;          // spin-inversion mask 2^n - 1
  // Fermionic bases: image[e * nbits + b] is the position that orbital b
  // takes under element e (null for spins).
  uint8_t const *image;
  int nbits;
  // Stabiliser sums <= norm_eps count as zero when norm_eps > 0 (0: the
  // reference's rule, norm > 0 and sqrt of the raw sum).
  double norm_eps;
  // Translation factorisation (fact_nq = 0: none): every element is t o q
  // for a translation t of the fact_lx x fact_ly torus (site r * lx + c) and
  // a point element q; fact_q[q] is q's element index,
  // fact_elem[(q * ly + b) * lx + a] the index of t_{a,b} o q and fact_key
  // its rank in the state_info reduction order.  fact_m1 / fact_m2 are the
  // bits with column < lx - 1 / column lx - 1.
  int fact_lx, fact_ly, fact_nq;
  int32_t const *fact_q;
  int32_t const *fact_elem;
  int32_t const *fact_key;
  uint64_t fact_m1, fact_m2;
};

// Sorted representatives plus a prefix table over their top bits:
// states with (x >> shift) == b live in reps[offsets[b] .. offsets[b+1]).
struct IndexView {
  uint64_t const *reps;
  uint32_t const *offsets;
  int64_t count;
  int shift;
  int prefix_bits;
};

// Non-branching terms, one 64-bit word each.
struct TermsView {
  double const *v; // interleaved re, im
  uint64_t const *m;
  uint64_t const *r;
  uint64_t const *x;
  uint64_t const *s;
  int count;
};

// ---------------------------------------------------------------------------
// Host objects
// ---------------------------------------------------------------------------

constexpr uint32_t kGroupMagic = 0x4c534147u; // "LSAG"
constexpr uint32_t kIndexMagic = 0x4c534149u; // "LSAI"

// A process-wide serial number for host objects, so that an object created
// at the address of a destroyed one is still told apart from it.
inline uint64_t next_object_serial() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}

// Private data behind ls_internal_create_halide_kernel_data.
struct GroupData {
  uint32_t magic = kGroupMagic;
  uint64_t serial = next_object_serial();
  int number_bits = 0;
  int size = 0;
  int depth = 0;
  int spin_inversion = 0;
  // 0: spins; 1: spinless fermions; 2: spinful fermions (the site
  // permutation acts on both species).  Fermionic groups carry the sign of
  // the relabelled creation operators (no spin inversion).
  int fermion_species = 0;
  double norm_eps = 0.0;
  bool canonical_shifts = true; // shifts are 1,2,..,N/2,..,2,1
  uint64_t flip_mask = 0;
  std::vector<uint64_t> h_masks; // element-major
  std::vector<uint64_t> h_shifts;
  std::vector<double> h_eig_re, h_eig_im;
  DeviceArray<uint64_t> d_masks;
  DeviceArray<uint64_t> d_shifts;
  DeviceArray<double> d_eig_re, d_eig_im;
  DeviceArray<uint8_t> d_image;
  // Translation factorisation (found by group_detect_translations; fact_lx
  // = 0 when the group is not a torus translation group times point
  // elements with real characters +-1).
  int fact_lx = 0, fact_ly = 0;
  std::vector<int32_t> h_fact_q, h_fact_elem, h_fact_key;
  DeviceArray<int32_t> d_fact_q, d_fact_elem, d_fact_key;
  uint64_t fact_m1 = 0, fact_m2 = 0;
  bool factored() const { return fact_lx > 0 && fermion_species == 0; }
  GroupView view() const {
    return GroupView{d_masks.ptr, d_eig_re.ptr, d_eig_im.ptr, d_shifts.ptr, size, depth,
                     flip_mask, d_image.ptr, number_bits, norm_eps,
                     fact_lx, fact_ly, static_cast<int>(h_fact_q.size()), d_fact_q.ptr, d_fact_elem.ptr,
                     d_fact_key.ptr, fact_m1, fact_m2};
  }
};

// Looks for the factorisation G = {t o q}: t runs over the translations of
// an lx x ly torus, q over the elements fixing site 0, and every character
// has real part exactly +-1 (then stabiliser sums are exact in any order).
// Leaves fact_lx = 0 when the group has no such form.
bool group_detect_translations(GroupData &g);

// Marks a group as fermionic (species 1 or 2) and uploads the orbital
// images; spin inversion must be off.
bool group_make_fermionic(GroupData &g, int species);
// Element order for the enumeration filter (most rejecting first), from a
// seeded sample of candidates of the given Hamming weight (-1: any).
std::vector<int> group_rejection_order(GroupData const &g, int hamming, int number_bits);

// Private data behind ls_hs_create_state_index_binary_search_kernel_data.
struct IndexData {
  uint32_t magic = kIndexMagic;
  uint64_t serial = next_object_serial();
  int64_t number_states = 0;
  uint64_t const *h_reps = nullptr; // borrowed host list (may be null)
  int number_bits = 0;
  int prefix_bits = 0;
  int shift = 0;
  DeviceArray<uint64_t> d_reps_owned;
  uint64_t const *d_reps = nullptr; // owned copy or borrowed device array
  DeviceArray<uint32_t> d_offsets;
  IndexView view() const {
    return IndexView{d_reps, d_offsets.ptr, number_states, shift, prefix_bits};
  }
};

GroupData *group_data_create(ls_hs_permutation_group const *g, int spin_inversion);
// Builds the device prefix table for sorted device states.
bool index_build(IndexData &idx, uint64_t const *d_reps, int64_t count,
                 int number_bits, int prefix_bits);
// Default prefix width for `count` states of `number_bits` bits.
int default_prefix_bits(int number_bits, int64_t count);

// Device enumeration of representatives (enumerate.hip).
struct EnumerationSpec {
  int number_bits = 0;
  int hamming_weight = -1;     // -1: every integer in [lower, upper]
  int spin_inversion = 0;
  GroupData const *group = nullptr; // null: no permutation symmetries
  // Spinful fermions with per-species counts: lower number_sites bits are
  // the up-spins (weight number_up), upper bits the down-spins.  With a
  // group the product states are filtered through the orbit test.
  bool spinful_per_sector = false;
  int number_sites = 0;
  int number_up = -1, number_down = -1;
};
// Whether enumerate_states may use the factorised filter for groups with a
// translation factorisation (on by default; the result is the same).
extern std::atomic<bool> g_factored_enumeration;
// Frees the enumeration scratch kept between basis builds.
void release_enumeration_scratch();
// Writes the ascending representatives into `out`; returns false on error.
bool enumerate_states(EnumerationSpec const &spec, uint64_t lower,
                      uint64_t upper, DeviceArray<uint64_t> &out,
                      int64_t &count, int64_t &candidates);

// Orbit kernels on device arrays (orbit.hip).
bool launch_state_info(GroupData const &g, int64_t count, uint64_t const *d_x,
                       int64_t x_stride, uint64_t *d_reps, int64_t rep_stride,
                       double *d_chars, double *d_norms);
bool launch_is_representative(GroupData const &g, int64_t count,
                              uint64_t const *d_x, int64_t x_stride,
                              uint8_t *d_flags, double *d_norms);
bool launch_state_index(IndexView const &idx, int64_t count,
                        uint64_t const *d_x, int64_t *d_out);

// Binomial coefficients C(n, k) for 0 <= n, k <= 64, C(n, k) = 0 when
// n <= 0 or k > n (Basis.hs:487-509).
uint64_t const *binomials_device();
uint64_t binomial_host(int n, int k);
uint64_t fixed_hamming_state_to_index(uint64_t x);
uint64_t fixed_hamming_index_to_state(uint64_t index, int h);

} // namespace lsamd

// Opaque handles of the additive API.
struct ls_amd_basis {
  int number_sites = 0;
  int hamming_weight = -1;
  int spin_inversion = 0;
  int fermion_species = 0;        // 0 spins, 1 spinless, 2 spinful fermions
  int number_up = -1, number_down = -1; // spinful per-species counts
  int number_bits = 0;
  std::unique_ptr<lsamd::GroupData> group; // null: no permutations
  lsamd::DeviceArray<uint64_t> d_states;
  int64_t number_states = -1;
  int64_t number_candidates = 0;
  lsamd::IndexData index;
  lsamd::DeviceArray<double> d_norms; // state_info norm of each representative
  bool built() const { return number_states >= 0; }
};
