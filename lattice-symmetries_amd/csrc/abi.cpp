// abi.cpp -- the C ABI of liblsamd.so.
//
// Part 1: the reference's entry points (kernels/reference.c,
// kernels/indexing.c, kernels/kernels.c and the Chapel table functions),
// host pointers in and out, work on the GPU.
// Part 2: additive handles (ls_amd_basis, ls_amd_operator) over device data.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>

#include "operator_internal.hpp"

using namespace lsamd;

namespace {

// ---------------------------------------------------------------------------
// Error handler (reference.c:11-36)
// ---------------------------------------------------------------------------
using handler_type = void (*)(char const *);

void default_handler(char const *message) {
  std::fprintf(stderr, "[Error]   %s\n[Error]   Aborting ...\n", message);
  std::abort();
}

std::atomic<handler_type> g_handler{&default_handler};

// Report the engine's last error through the handler.
void raise_last() { ls_hs_error(ls_amd_last_error()); }

// ---------------------------------------------------------------------------
// Engine table (reference.c:214-225)
// ---------------------------------------------------------------------------
ls_chpl_kernels g_chpl_kernels = {nullptr, nullptr, nullptr, nullptr};
std::mutex g_table_mutex;

// Host array -> device array.
template <typename T> bool to_device(DeviceArray<T> &d, T const *h, size_t n) { return d.upload(h, n); }

template <typename T> bool to_host(T *h, T const *d, size_t n) { return copy_to_host(h, d, n * sizeof(T)); }

// Gathers a strided host array into a contiguous vector.
std::vector<uint64_t> gather_strided(uint64_t const *p, ptrdiff_t n, ptrdiff_t stride) {
  std::vector<uint64_t> v(static_cast<size_t>(n));
  for (ptrdiff_t i = 0; i < n; ++i) v[static_cast<size_t>(i)] = p[i * stride];
  return v;
}

GroupData *as_group(void const *p) {
  auto *g = static_cast<GroupData *>(const_cast<void *>(p));
  return (g != nullptr && g->magic == kGroupMagic) ? g : nullptr;
}

IndexData *as_index(void const *p) {
  auto *d = static_cast<IndexData *>(const_cast<void *>(p));
  return (d != nullptr && d->magic == kIndexMagic) ? d : nullptr;
}

int number_bits_of(ls_hs_basis const *b) {
  return (b->particle_type == LS_HS_SPINFUL_FERMION ? 2 : 1) * b->number_sites;
}

// ---------------------------------------------------------------------------
// Operators used through ls_chpl_matrix_vector_product are cached per
// ls_hs_operator.  An entry is valid only for the exact data it was built
// from: the basis, its representatives (address, count and a sample of
// their contents), the serial numbers of the engine's group and index data
// behind the basis kernels, and the terms' contents.  Destroying a basis's
// group or index data evicts every entry built on it, and the device memory
// the entries hold is capped (least recently used entries go first).  A
// product holds its entry through a shared_ptr, so eviction during the
// product cannot free what it uses.
// ---------------------------------------------------------------------------
struct CacheSignature {
  ls_hs_basis const *basis = nullptr;
  void const *reps = nullptr;
  uint64_t count = 0;
  uint64_t reps_sample = 0;
  uint64_t group_serial = 0;
  uint64_t index_serial = 0;
  uint64_t terms = 0;
  bool operator==(CacheSignature const &o) const {
    return basis == o.basis && reps == o.reps && count == o.count && reps_sample == o.reps_sample &&
           group_serial == o.group_serial && index_serial == o.index_serial && terms == o.terms;
  }
};

struct CachedOperator {
  OperatorCore core;
  CacheSignature sig;
  GroupData const *group = nullptr; // engine data this entry reads (for eviction)
  IndexData const *index = nullptr;
  DeviceArray<uint64_t> owned_reps;
  IndexData owned_index;
  bool use_csr = false;
  size_t device_bytes = 0;
  uint64_t last_use = 0;
  // x and y of ls_chpl_matrix_vector_product, kept between calls; a call
  // that finds them in use by another thread allocates its own
  std::mutex vectors_busy;
  DeviceArray<double> dx, dy;
};
std::map<ls_hs_operator const *, std::shared_ptr<CachedOperator>> g_op_cache;
std::mutex g_op_mutex;
uint64_t g_op_clock = 0;

uint64_t fnv_bytes(uint64_t h, void const *p, size_t bytes) {
  auto const *c = static_cast<unsigned char const *>(p);
  for (size_t i = 0; i < bytes; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

uint64_t terms_hash(ls_hs_nonbranching_terms const *t) {
  if (t == nullptr) return 0x1234;
  uint64_t h = 1469598103934665603ull ^ static_cast<uint64_t>(t->number_terms);
  h = fnv_bytes(h, t->v, sizeof(ls_hs_scalar) * t->number_terms);
  for (uint64_t const *a : {t->m, t->r, t->x, t->s}) h = fnv_bytes(h, a, 8 * t->number_terms);
  return h;
}

// Count plus 64 evenly spaced words of the host representatives.
uint64_t reps_sample_hash(uint64_t const *reps, uint64_t count) {
  uint64_t h = fnv_bytes(1469598103934665603ull, &count, sizeof(count));
  if (reps == nullptr || count == 0) return h;
  for (uint64_t q = 0; q < 64; ++q) {
    uint64_t const i = count == 1 ? 0 : q * (count - 1) / 63;
    h = fnv_bytes(h, &reps[i], sizeof(uint64_t));
  }
  return h;
}

CacheSignature signature_of(ls_hs_operator const *op) {
  ls_hs_basis const *basis = op->basis;
  CacheSignature s;
  s.basis = basis;
  s.reps = basis->representatives.elts;
  s.count = basis->representatives.num_elts;
  s.reps_sample = reps_sample_hash(static_cast<uint64_t const *>(s.reps), s.count);
  if (basis->kernels != nullptr) {
    if (GroupData const *g = as_group(basis->kernels->state_info_data)) s.group_serial = g->serial;
    if (IndexData const *d = as_index(basis->kernels->state_index_data)) s.index_serial = d->serial;
  }
  s.terms = terms_hash(op->diag_terms) * 31 + terms_hash(op->off_diag_terms) * 131;
  return s;
}

size_t csr_bytes(Csr const &c) {
  size_t b = static_cast<size_t>(c.rows + 1) * 8 + static_cast<size_t>(c.nnz) * 4;
  if (c.wide) b += static_cast<size_t>(c.nnz) * 2;
  else if (!c.packed) b += static_cast<size_t>(c.nnz) * (c.is_complex ? 16 : 8);
  return b + static_cast<size_t>(c.rows) * 8;
}

// Device memory the cached entries may hold together: LSAMD_CACHE_BYTES, or
// half of the device's memory.
size_t cache_capacity() {
  if (char const *e = std::getenv("LSAMD_CACHE_BYTES")) return static_cast<size_t>(std::strtoull(e, nullptr, 10));
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  return total_b / 2;
}

// Evicts least recently used entries (never `keep`) until the total fits.
void enforce_cache_capacity(CachedOperator const *keep) {
  size_t const cap = cache_capacity();
  while (true) {
    size_t total = 0;
    auto victim = g_op_cache.end();
    for (auto it = g_op_cache.begin(); it != g_op_cache.end(); ++it) {
      total += it->second->device_bytes;
      if (it->second.get() != keep && (victim == g_op_cache.end() || it->second->last_use < victim->second->last_use))
        victim = it;
    }
    if (total <= cap || victim == g_op_cache.end()) return;
    g_op_cache.erase(victim);
  }
}

// Every entry that reads `group` or `index` (engine data about to be freed).
void evict_entries_on(void const *group, void const *index) {
  std::lock_guard<std::mutex> lock(g_op_mutex);
  for (auto it = g_op_cache.begin(); it != g_op_cache.end();) {
    CachedOperator const &c = *it->second;
    bool const hit = (group != nullptr && c.group == group) || (index != nullptr && c.index == index);
    it = hit ? g_op_cache.erase(it) : std::next(it);
  }
}

// Fills the basis side of an OperatorCore from an ls_hs_basis whose
// representatives are built.  `holder` keeps device copies made here.
bool core_from_basis(OperatorCore &core, ls_hs_basis const *basis, CachedOperator &holder) {
  if (basis == nullptr) return fail("operator has no basis");
  if (basis->representatives.elts == nullptr && basis->representatives.num_elts > 0)
    return fail("basis is not built");
  int const nbits = number_bits_of(basis);
  if (nbits > 64) return fail("bases with more than 64 bits are not yet implemented");
  core.number_bits = nbits;
  core.spin_inversion = basis->spin_inversion;
  core.number_states = static_cast<int64_t>(basis->representatives.num_elts);
  GroupData *g = basis->kernels != nullptr ? as_group(basis->kernels->state_info_data) : nullptr;
  if (!basis->requires_projection) {
    core.kind = kNoProjection;
  } else if (g != nullptr && g->size > 0) {
    core.kind = kPermutations;
    core.group = g;
    holder.group = g;
  } else if (basis->spin_inversion != 0) {
    core.kind = kInversionOnly;
  } else {
    return fail("basis requires projection but has no state_info kernel data from this engine");
  }
  core.identity_index = basis->state_index_is_identity;
  IndexData *idx = basis->kernels != nullptr ? as_index(basis->kernels->state_index_data) : nullptr;
  uint64_t const *host_reps = static_cast<uint64_t const *>(basis->representatives.elts);
  if (idx != nullptr && idx->h_reps == host_reps && idx->number_states == core.number_states) {
    core.index = idx->view();
    holder.index = idx;
  } else {
    if (!holder.owned_reps.upload(host_reps, static_cast<size_t>(core.number_states))) return false;
    if (!index_build(holder.owned_index, holder.owned_reps.ptr, core.number_states, nbits,
                     default_prefix_bits(nbits, core.number_states)))
      return false;
    core.index = holder.owned_index.view();
  }
  return true;
}

std::shared_ptr<CachedOperator> cached_operator(ls_hs_operator const *op) {
  CacheSignature const sig = signature_of(op);
  std::lock_guard<std::mutex> lock(g_op_mutex);
  auto it = g_op_cache.find(op);
  if (it != g_op_cache.end()) {
    if (it->second->sig == sig) {
      it->second->last_use = ++g_op_clock;
      return it->second;
    }
    g_op_cache.erase(it);
  }
  auto c = std::make_shared<CachedOperator>();
  c->sig = sig;
  if (!core_from_basis(c->core, op->basis, *c)) return nullptr;
  if (!c->core.diag.load(op->diag_terms) || !c->core.off.load(op->off_diag_terms)) return nullptr;
  // CSR when it fits comfortably in device memory (LSAMD_MATVEC=otf|csr overrides).
  char const *mode = std::getenv("LSAMD_MATVEC");
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  double const estimate = static_cast<double>(c->core.number_states) *
                          std::max(1, c->core.off.count) * 12.0;
  c->use_csr = mode != nullptr ? std::strcmp(mode, "csr") == 0 : estimate < 0.4 * static_cast<double>(free_b);
  if (c->use_csr) {
    if (!operator_build_csr(c->core, 64)) return nullptr;
    c->device_bytes = csr_bytes(c->core.csr);
  }
  c->device_bytes += static_cast<size_t>(c->core.number_states) * 8 * (c->owned_reps.ptr != nullptr ? 2 : 1);
  c->device_bytes += static_cast<size_t>(c->core.number_states) * 16; // x and y
  c->last_use = ++g_op_clock;
  g_op_cache[op] = c;
  enforce_cache_capacity(c.get());
  return c;
}

void free_external(void *p) { std::free(p); }

} // namespace

// ===========================================================================
// Part 1: reference entry points
// ===========================================================================

extern "C" void ls_hs_fatal_error(char const *func, int line, char const *message) {
  std::fprintf(stderr, "[Error]   [%s#%i] %s\n[Error]   Aborting ...\n", func, line, message);
  std::abort();
}

extern "C" void ls_hs_set_exception_handler(void (*handler)(char const *message)) {
  g_handler.store(handler != nullptr ? handler : &default_handler);
}

extern "C" void ls_hs_error(char const *message) {
  handler_type h = g_handler.load();
  if (h == nullptr) ls_hs_fatal_error(__func__, __LINE__, "error handler is NULL");
  h(message);
}

extern "C" void ls_hs_internal_destroy_external_array(chpl_external_array *arr) {
  if (arr == nullptr) ls_hs_fatal_error(__func__, __LINE__, "trying to destroy a NULL chpl_external_array");
  if (arr->freer != nullptr) reinterpret_cast<void (*)(void *)>(arr->freer)(arr->elts);
}

extern "C" int ls_hs_internal_read_refcount(int const *refcount) {
  return __atomic_load_n(refcount, __ATOMIC_SEQ_CST);
}
extern "C" void ls_hs_internal_write_refcount(int *refcount, int value) {
  __atomic_store_n(refcount, value, __ATOMIC_SEQ_CST);
}
extern "C" int ls_hs_internal_inc_refcount(int *refcount) {
  return __atomic_fetch_add(refcount, 1, __ATOMIC_SEQ_CST);
}
extern "C" int ls_hs_internal_dec_refcount(int *refcount) {
  return __atomic_fetch_sub(refcount, 1, __ATOMIC_SEQ_CST);
}

extern "C" void ls_internal_operator_apply_diag_x1(ls_hs_operator const *op, ptrdiff_t batch_size,
                                                   uint64_t const *alphas, double *ys,
                                                   double const *xs) {
  if (batch_size <= 0) return;
  if (op->diag_terms == nullptr || op->diag_terms->number_terms == 0) {
    std::memset(ys, 0, sizeof(double) * static_cast<size_t>(batch_size));
    return;
  }
  Terms t;
  DeviceArray<uint64_t> a;
  DeviceArray<double> x, y;
  bool ok = ensure_device() && t.load(op->diag_terms) && to_device(a, alphas, batch_size) &&
            y.allocate(batch_size) && (xs == nullptr || to_device(x, xs, batch_size)) &&
            apply_diag_device(t, batch_size, a.ptr, y.ptr, xs != nullptr ? x.ptr : nullptr) &&
            to_host(ys, y.ptr, batch_size);
  if (!ok) raise_last();
}

extern "C" void ls_internal_operator_apply_off_diag_x1(ls_hs_operator const *op, ptrdiff_t batch_size,
                                                       uint64_t const *alphas, uint64_t *betas,
                                                       ls_hs_scalar *coeffs, ptrdiff_t *offsets,
                                                       double const *xs) {
  if (op->off_diag_terms == nullptr || op->off_diag_terms->number_terms == 0) {
    std::memset(offsets, 0, sizeof(ptrdiff_t) * static_cast<size_t>(batch_size + 1));
    return;
  }
  if (batch_size <= 0) {
    offsets[0] = 0;
    return;
  }
  Terms t;
  DeviceArray<uint64_t> a, b;
  DeviceArray<double> x, c;
  DeviceArray<int64_t> o;
  int64_t total = 0;
  bool ok = ensure_device() && t.load(op->off_diag_terms) && to_device(a, alphas, batch_size) &&
            (xs == nullptr || to_device(x, xs, batch_size)) &&
            apply_off_diag_device(t, batch_size, a.ptr, xs != nullptr ? x.ptr : nullptr, b, c, o, total) &&
            to_host(reinterpret_cast<int64_t *>(offsets), o.ptr, batch_size + 1) &&
            to_host(betas, b.ptr, total) && to_host(reinterpret_cast<double *>(coeffs), c.ptr, 2 * total);
  if (!ok) raise_last();
}

extern "C" void ls_hs_state_index(ls_hs_basis const *basis, ptrdiff_t batch_size, uint64_t const *spins,
                                  ptrdiff_t spins_stride, ptrdiff_t *indices, ptrdiff_t indices_stride) {
  if (basis->kernels->state_index_kernel == nullptr)
    ls_hs_fatal_error(__func__, __LINE__, "state_index_kernel is NULL");
  basis->kernels->state_index_kernel(batch_size, spins, spins_stride, indices, indices_stride,
                                     basis->kernels->state_index_data);
}

extern "C" void ls_hs_is_representative(ls_hs_basis const *basis, ptrdiff_t batch_size,
                                        uint64_t const *alphas, ptrdiff_t alphas_stride,
                                        uint8_t *are_representatives, double *norms) {
  if (basis->kernels->is_representative_kernel == nullptr)
    ls_hs_fatal_error(__func__, __LINE__,
                      "is_representative_kernel is NULL, perhaps this basis requires no projection?");
  basis->kernels->is_representative_kernel(batch_size, alphas, alphas_stride, are_representatives, norms,
                                           basis->kernels->is_representative_data);
}

extern "C" void ls_hs_state_info(ls_hs_basis const *basis, ptrdiff_t batch_size, uint64_t const *alphas,
                                 ptrdiff_t alphas_stride, uint64_t *betas, ptrdiff_t betas_stride,
                                 ls_hs_scalar *characters, double *norms) {
  if (basis->kernels->state_info_kernel == nullptr)
    ls_hs_fatal_error(__func__, __LINE__,
                      "state_info_kernel is NULL, perhaps this basis requires no projection?");
  basis->kernels->state_info_kernel(batch_size, alphas, alphas_stride, betas, betas_stride, characters, norms,
                                    basis->kernels->state_info_data);
}

extern "C" void ls_hs_build_representatives(ls_hs_basis *basis, uint64_t lower, uint64_t upper) {
  ls_chpl_kernels const *k = ls_hs_internal_get_chpl_kernels();
  if (k->enumerate_states == nullptr)
    ls_hs_fatal_error(__func__, __LINE__,
                      "enumerate_states kernel is NULL, Chapel code was supposed to initialize it");
  if (basis->representatives.num_elts > 0) return; // already built
  k->enumerate_states(basis, lower, upper, &basis->representatives);
  int const default_cache_bits = 22;
  basis->kernels->state_index_data = ls_hs_create_state_index_binary_search_kernel_data(
      &basis->representatives, number_bits_of(basis), default_cache_bits);
  basis->kernels->state_index_kernel = &ls_hs_state_index_binary_search_kernel;
}

extern "C" void ls_hs_unchecked_set_representatives(ls_hs_basis *basis, chpl_external_array const *states,
                                                    int cache_bits) {
  if (basis->representatives.num_elts != 0)
    ls_hs_fatal_error(__func__, __LINE__, "representatives have already been set");
  if (basis->kernels == nullptr) ls_hs_fatal_error(__func__, __LINE__, "basis->kernels not set");
  if (basis->kernels->state_index_kernel != nullptr)
    ls_hs_fatal_error(__func__, __LINE__, "state_index_kernel has already been set");
  basis->representatives = *states;
  basis->kernels->state_index_data =
      ls_hs_create_state_index_binary_search_kernel_data(&basis->representatives, number_bits_of(basis), cache_bits);
  basis->kernels->state_index_kernel = &ls_hs_state_index_binary_search_kernel;
}

extern "C" ls_chpl_kernels const *ls_hs_internal_get_chpl_kernels(void) { return &g_chpl_kernels; }

extern "C" void ls_hs_internal_set_chpl_kernels(ls_chpl_kernels const *kernels) {
  std::lock_guard<std::mutex> lock(g_table_mutex);
  g_chpl_kernels = *kernels;
}

// --- indexing.c -------------------------------------------------------------

extern "C" ls_hs_state_index_binary_search_data *
ls_hs_create_state_index_binary_search_kernel_data(chpl_external_array const *representatives, int number_bits,
                                                   int prefix_bits) {
  auto d = std::make_unique<IndexData>();
  int64_t const n = static_cast<int64_t>(representatives->num_elts);
  d->h_reps = static_cast<uint64_t const *>(representatives->elts);
  if (prefix_bits > number_bits) prefix_bits = number_bits; // indexing.c:104-106
  if (!ensure_device() || !d->d_reps_owned.upload(d->h_reps, static_cast<size_t>(n)) ||
      !index_build(*d, d->d_reps_owned.ptr, n, number_bits, std::max(prefix_bits, 1))) {
    raise_last();
    return nullptr;
  }
  return reinterpret_cast<ls_hs_state_index_binary_search_data *>(d.release());
}

extern "C" void ls_hs_destroy_state_index_binary_search_kernel_data(ls_hs_state_index_binary_search_data *cache) {
  IndexData *d = as_index(cache);
  if (d == nullptr) return;
  evict_entries_on(nullptr, d);
  delete d;
}

extern "C" void ls_hs_state_index_binary_search_kernel(ptrdiff_t batch_size, uint64_t const *spins,
                                                       ptrdiff_t spins_stride, ptrdiff_t *indices,
                                                       ptrdiff_t indices_stride, void const *private_kernel_data) {
  IndexData const *d = as_index(private_kernel_data);
  if (indices_stride != 1) ls_hs_fatal_error(__func__, __LINE__, "expected indices_stride==1");
  if (spins_stride != 1) ls_hs_fatal_error(__func__, __LINE__, "expected spins_stride==1");
  if (d == nullptr) ls_hs_fatal_error(__func__, __LINE__, "state index data was not created by this engine");
  if (batch_size <= 0) return;
  DeviceArray<uint64_t> x;
  DeviceArray<int64_t> out;
  bool ok = to_device(x, spins, batch_size) && out.allocate(batch_size) &&
            launch_state_index(d->view(), batch_size, x.ptr, out.ptr) &&
            to_host(reinterpret_cast<int64_t *>(indices), out.ptr, batch_size);
  if (!ok) raise_last();
}

// --- kernels.c ----------------------------------------------------------------

extern "C" ls_internal_halide_kernel_data *ls_internal_create_halide_kernel_data(ls_hs_permutation_group const *g,
                                                                                 int spin_inversion) {
  if (spin_inversion < -1 || spin_inversion > 1) ls_hs_fatal_error(__func__, __LINE__, "invalid spin_inversion");
  GroupData *d = group_data_create(g, spin_inversion);
  if (d == nullptr) raise_last();
  return reinterpret_cast<ls_internal_halide_kernel_data *>(d);
}

extern "C" void ls_internal_destroy_halide_kernel_data(ls_internal_halide_kernel_data *p) {
  GroupData *g = as_group(p);
  if (g == nullptr) return;
  evict_entries_on(g, nullptr);
  delete g;
}

extern "C" void ls_hs_is_representative_halide_kernel(ptrdiff_t batch_size, uint64_t const *alphas,
                                                      ptrdiff_t alphas_stride, uint8_t *are_representatives,
                                                      double *norms, void const *private_data) {
  GroupData const *g = as_group(private_data);
  if (g == nullptr) ls_hs_fatal_error(__func__, __LINE__, "kernel data was not created by this engine");
  if (batch_size <= 0) return;
  std::vector<uint64_t> const h = gather_strided(alphas, batch_size, alphas_stride);
  DeviceArray<uint64_t> x;
  DeviceArray<uint8_t> f;
  DeviceArray<double> nrm;
  bool ok = to_device(x, h.data(), h.size()) && f.allocate(batch_size) && nrm.allocate(batch_size) &&
            launch_is_representative(*g, batch_size, x.ptr, 1, f.ptr, nrm.ptr) &&
            to_host(are_representatives, f.ptr, batch_size) && to_host(norms, nrm.ptr, batch_size);
  if (!ok) raise_last();
}

extern "C" void ls_hs_state_info_halide_kernel(ptrdiff_t batch_size, uint64_t const *alphas, ptrdiff_t alphas_stride,
                                               uint64_t *betas, ptrdiff_t betas_stride, ls_hs_scalar *characters,
                                               double *norms, void const *private_data) {
  GroupData const *g = as_group(private_data);
  if (g == nullptr) ls_hs_fatal_error(__func__, __LINE__, "kernel data was not created by this engine");
  if (batch_size <= 0) return;
  std::vector<uint64_t> const h = gather_strided(alphas, batch_size, alphas_stride);
  DeviceArray<uint64_t> x, r;
  DeviceArray<double> c, nrm;
  std::vector<uint64_t> reps(static_cast<size_t>(batch_size));
  bool ok = to_device(x, h.data(), h.size()) && r.allocate(batch_size) && c.allocate(2 * batch_size) &&
            nrm.allocate(batch_size) && launch_state_info(*g, batch_size, x.ptr, 1, r.ptr, 1, c.ptr, nrm.ptr) &&
            to_host(reps.data(), r.ptr, batch_size) &&
            to_host(reinterpret_cast<double *>(characters), c.ptr, 2 * batch_size) &&
            to_host(norms, nrm.ptr, batch_size);
  if (!ok) {
    raise_last();
    return;
  }
  for (ptrdiff_t i = 0; i < batch_size; ++i) betas[i * betas_stride] = reps[static_cast<size_t>(i)];
}

// --- Chapel table functions ------------------------------------------------

extern "C" void ls_chpl_enumerate_representatives(ls_hs_basis const *p, uint64_t lower, uint64_t upper,
                                                  chpl_external_array *dest) {
  EnumerationSpec spec;
  spec.number_bits = number_bits_of(p);
  spec.number_sites = p->number_sites;
  spec.spin_inversion = p->spin_inversion;
  if (p->particle_type == LS_HS_SPIN) {
    spec.hamming_weight = p->number_up; // SpinBasis: number_up is the Hamming weight (Basis.hs:669-671)
    GroupData const *g = p->kernels != nullptr ? as_group(p->kernels->is_representative_data) : nullptr;
    if (g != nullptr && g->size > 0) spec.group = g;
  } else if (p->particle_type == LS_HS_SPINFUL_FERMION) {
    if (p->number_up >= 0) {
      spec.spinful_per_sector = true;
      spec.number_up = p->number_up;
      spec.number_down = p->number_particles - p->number_up;
    } else {
      spec.hamming_weight = p->number_particles;
    }
  } else {
    spec.hamming_weight = p->number_particles;
  }
  DeviceArray<uint64_t> out;
  int64_t count = 0, candidates = 0;
  if (!enumerate_states(spec, lower, upper, out, count, candidates)) {
    raise_last();
    return;
  }
  auto *h = static_cast<uint64_t *>(std::malloc(sizeof(uint64_t) * static_cast<size_t>(count > 0 ? count : 1)));
  if (h == nullptr || !to_host(h, out.ptr, static_cast<size_t>(count))) {
    std::free(h);
    raise_last();
    return;
  }
  dest->elts = h;
  dest->num_elts = static_cast<uint64_t>(count);
  dest->freer = reinterpret_cast<void *>(&free_external);
}

extern "C" void ls_chpl_operator_apply_diag(ls_hs_operator *matrixPtr, int64_t count, uint64_t *alphas,
                                            chpl_external_array *coeffs, int64_t /*numTasks*/) {
  if (number_bits_of(matrixPtr->basis) > 64) {
    ls_hs_error("bases with more than 64 bits are not yet implemented");
    return;
  }
  if (matrixPtr->basis->requires_projection) {
    ls_hs_error("bases that require projection are not yet supported");
    return;
  }
  auto *cs = static_cast<double *>(std::malloc(sizeof(double) * static_cast<size_t>(count > 0 ? count : 1)));
  ls_internal_operator_apply_diag_x1(matrixPtr, count, alphas, cs, nullptr);
  coeffs->elts = cs;
  coeffs->num_elts = static_cast<uint64_t>(count);
  coeffs->freer = reinterpret_cast<void *>(&free_external);
}

extern "C" void ls_chpl_operator_apply_off_diag(ls_hs_operator *matrixPtr, int64_t count, uint64_t *alphas,
                                                chpl_external_array *betas, chpl_external_array *coeffs,
                                                chpl_external_array *offsets, int64_t /*numTasks*/) {
  if (number_bits_of(matrixPtr->basis) > 64) {
    ls_hs_error("bases with more than 64 bits are not yet implemented");
    return;
  }
  int const T = matrixPtr->off_diag_terms != nullptr ? matrixPtr->off_diag_terms->number_terms : 0;
  size_t const cap = static_cast<size_t>(count) * static_cast<size_t>(T);
  auto *o = static_cast<ptrdiff_t *>(std::calloc(static_cast<size_t>(count + 1), sizeof(ptrdiff_t)));
  if (T != 0) {
    auto *b = static_cast<uint64_t *>(std::calloc(cap > 0 ? cap : 1, sizeof(uint64_t)));
    auto *c = static_cast<ls_hs_scalar *>(std::calloc(cap > 0 ? cap : 1, sizeof(ls_hs_scalar)));
    ls_internal_operator_apply_off_diag_x1(matrixPtr, count, alphas, b, c, o, nullptr);
    *betas = chpl_external_array{b, cap, reinterpret_cast<void *>(&free_external)};
    *coeffs = chpl_external_array{c, cap, reinterpret_cast<void *>(&free_external)};
  } else {
    *betas = chpl_external_array{nullptr, 0, nullptr};
    *coeffs = chpl_external_array{nullptr, 0, nullptr};
  }
  *offsets = chpl_external_array{o, static_cast<uint64_t>(count + 1), reinterpret_cast<void *>(&free_external)};
}

extern "C" void ls_chpl_matrix_vector_product(ls_hs_operator *matrixPtr, int numVectors, double const *xPtr,
                                              double *yPtr) {
  if (number_bits_of(matrixPtr->basis) > 64) {
    ls_hs_error("bases with more than 64 bits are not yet implemented");
    return;
  }
  if (numVectors != 1) {
    ls_hs_error("applying the Operator to more than 1 vector is not yet implemented");
    return;
  }
  if (matrixPtr->basis->representatives.elts == nullptr) {
    ls_hs_error("basis is not built");
    return;
  }
  if (!ensure_device()) {
    raise_last();
    return;
  }
  std::shared_ptr<CachedOperator> c = cached_operator(matrixPtr);
  if (c == nullptr) {
    raise_last();
    return;
  }
  OperatorCore &core = c->core;
  int64_t const n = core.number_states;
  bool const has_diag = core.diag.count > 0;
  bool const has_off = core.off.count > 0;
  if (!has_diag && !has_off) return; // DistributedMatrixVector.chpl:1054-1057
  // x (and y when the product accumulates into it) go through the pinned
  // staging copies; the device vectors stay with the cached operator
  std::unique_lock<std::mutex> vl(c->vectors_busy, std::try_to_lock);
  DeviceArray<double> own_x, own_y;
  DeviceArray<double> &x = vl.owns_lock() ? c->dx : own_x;
  DeviceArray<double> &y = vl.owns_lock() ? c->dy : own_y;
  size_t const bytes = sizeof(double) * static_cast<size_t>(n);
  bool ok = (x.count == static_cast<size_t>(n) || x.allocate(n)) &&
            (y.count == static_cast<size_t>(n) || y.allocate(n)) && copy_to_device(x.ptr, xPtr, bytes) &&
            (has_diag || copy_to_device(y.ptr, yPtr, bytes));
  if (ok && c->use_csr) {
    // y = diag x (or the caller's y), then the rows in chunks, each chunk of
    // y copied to the host while the next one is computed
    auto rows = [&](size_t first, size_t count) -> bool {
      return operator_csr_f64_rows(core, x.ptr, y.ptr, static_cast<int64_t>(first / sizeof(double)),
                                   static_cast<int64_t>((first + count) / sizeof(double)));
    };
    ok = (!has_diag || operator_csr_f64_diag(core, x.ptr, y.ptr)) &&
         produce_to_host(yPtr, y.ptr, bytes, sizeof(double), rows);
  } else if (ok) {
    ok = operator_matvec_otf(core, x.ptr, y.ptr, !has_diag, false) && copy_to_host(yPtr, y.ptr, bytes);
  }
  if (!ok) raise_last();
}

extern "C" void ls_chpl_init(void) {
  (void)ensure_device();
  ls_chpl_kernels const k = {&ls_chpl_enumerate_representatives, &ls_chpl_operator_apply_off_diag,
                             &ls_chpl_operator_apply_diag, &ls_chpl_matrix_vector_product};
  ls_hs_internal_set_chpl_kernels(&k);
}

extern "C" void ls_chpl_finalize(void) {
  {
    std::lock_guard<std::mutex> lock(g_op_mutex);
    g_op_cache.clear();
  }
  release_enumeration_scratch();
}

extern "C" void ls_amd_forget_operator(ls_hs_operator const *op) {
  std::lock_guard<std::mutex> lock(g_op_mutex);
  g_op_cache.erase(op);
}

extern "C" int64_t ls_amd_operator_cache_entries(void) {
  std::lock_guard<std::mutex> lock(g_op_mutex);
  return static_cast<int64_t>(g_op_cache.size());
}

// ===========================================================================
// Part 2: additive handles
// ===========================================================================

extern "C" ls_amd_basis *ls_amd_basis_create(int number_sites, int hamming_weight, int spin_inversion,
                                             ls_hs_permutation_group const *group) {
  if (number_sites <= 0 || number_sites > 64) {
    fail("number of sites must be in [1, 64]");
    return nullptr;
  }
  if (spin_inversion < -1 || spin_inversion > 1) {
    fail("invalid spin_inversion");
    return nullptr;
  }
  if (hamming_weight > number_sites) {
    fail("invalid Hamming weight");
    return nullptr;
  }
  if (!ensure_device()) return nullptr;
  auto b = std::make_unique<ls_amd_basis>();
  b->number_sites = number_sites;
  b->number_bits = number_sites;
  b->hamming_weight = hamming_weight;
  b->spin_inversion = spin_inversion;
  if (group != nullptr && group->number_masks > 0) {
    if (group->number_bits != number_sites) {
      fail("group acts on a different number of sites");
      return nullptr;
    }
    b->group.reset(group_data_create(group, spin_inversion));
    if (!b->group) return nullptr;
  }
  return b.release();
}

extern "C" ls_amd_basis *ls_amd_fermion_basis_create(int number_sites, int number_up, int number_down,
                                                    ls_hs_permutation_group const *group) {
  bool const spinful = number_down >= 0;
  int const nbits = spinful ? 2 * number_sites : number_sites;
  if (number_sites <= 0 || nbits > 64) {
    fail("fermion basis needs 1..64 bits");
    return nullptr;
  }
  if (number_up > number_sites || number_down > number_sites || (spinful && number_up < 0)) {
    fail("invalid particle numbers");
    return nullptr;
  }
  if (!ensure_device()) return nullptr;
  auto b = std::make_unique<ls_amd_basis>();
  b->number_sites = number_sites;
  b->number_bits = nbits;
  b->fermion_species = spinful ? 2 : 1;
  b->number_up = number_up;
  b->number_down = number_down;
  b->hamming_weight = spinful ? -1 : number_up;
  b->spin_inversion = 0;
  if (group != nullptr && group->number_masks > 0) {
    if (group->number_bits != nbits) {
      fail("group must act on all " + std::to_string(nbits) + " bits");
      return nullptr;
    }
    b->group.reset(group_data_create(group, 0));
    if (!b->group || !group_make_fermionic(*b->group, b->fermion_species)) return nullptr;
  }
  return b.release();
}

extern "C" void ls_amd_basis_destroy(ls_amd_basis *b) { delete b; }

extern "C" int ls_amd_basis_set_norm_tolerance(ls_amd_basis *b, double eps) {
  if (!(eps >= 0.0)) return fail("tolerance must be >= 0") ? 0 : -1;
  if (!b->group) return fail("basis has no permutation symmetries") ? 0 : -1;
  b->group->norm_eps = eps;
  b->d_norms.reset();
  return 0;
}

// Basis.hs:734-774
extern "C" uint64_t ls_amd_basis_min_state(ls_amd_basis const *b) {
  auto ones = [](int h) { return h <= 0 ? uint64_t{0} : (h >= 64 ? ~uint64_t{0} : ((uint64_t{1} << h) - 1)); };
  if (b->fermion_species == 2) return (ones(b->number_down) << b->number_sites) | ones(b->number_up);
  if (b->hamming_weight >= 0) return ones(b->hamming_weight);
  return 0;
}

extern "C" uint64_t ls_amd_basis_max_state(ls_amd_basis const *b) {
  if (b->fermion_species == 2) {
    int const n = b->number_sites;
    auto top = [n](int h) { return h <= 0 ? uint64_t{0} : (((uint64_t{1} << h) - 1) << (n - h)); };
    return (top(b->number_down) << n) | top(b->number_up);
  }
  int const n = b->number_bits, h = b->hamming_weight;
  if (h >= 0) {
    uint64_t const ones = h == 0 ? 0 : ((h == 64) ? ~uint64_t{0} : ((uint64_t{1} << h) - 1));
    return b->spin_inversion == 0 ? (ones << (n - h)) : (ones << (n - h - 1));
  }
  return n == 64 ? ~uint64_t{0} : ((uint64_t{1} << n) - 1);
}

static bool basis_finish(ls_amd_basis *b) {
  b->d_norms.reset();
  int const nbits = b->number_bits;
  return index_build(b->index, b->d_states.ptr, b->number_states, nbits,
                     default_prefix_bits(nbits, b->number_states));
}

extern "C" int64_t ls_amd_basis_build(ls_amd_basis *b, uint64_t lower, uint64_t upper) {
  EnumerationSpec spec;
  spec.number_bits = b->number_bits;
  spec.number_sites = b->number_sites;
  spec.hamming_weight = b->hamming_weight;
  spec.spin_inversion = b->spin_inversion;
  spec.group = b->group.get();
  if (b->fermion_species == 2) {
    spec.spinful_per_sector = true;
    spec.number_up = b->number_up;
    spec.number_down = b->number_down;
  }
  int64_t count = 0, candidates = 0;
  if (!enumerate_states(spec, lower, upper, b->d_states, count, candidates)) return -1;
  b->number_states = count;
  b->number_candidates = candidates;
  if (!basis_finish(b)) return -1;
  return count;
}

extern "C" int64_t ls_amd_basis_number_states(ls_amd_basis const *b) { return b->number_states; }
extern "C" int64_t ls_amd_basis_number_candidates(ls_amd_basis const *b) { return b->number_candidates; }
extern "C" uint64_t const *ls_amd_basis_states_device(ls_amd_basis const *b) { return b->d_states.ptr; }

extern "C" int ls_amd_basis_copy_states(ls_amd_basis const *b, uint64_t *host_out) {
  if (!b->built()) return fail("basis is not built") ? 0 : -1;
  return to_host(host_out, b->d_states.ptr, static_cast<size_t>(b->number_states)) ? 0 : -1;
}

extern "C" int ls_amd_basis_set_states(ls_amd_basis *b, uint64_t const *host_states, int64_t count) {
  if (!b->d_states.upload(host_states, static_cast<size_t>(count))) return -1;
  b->number_states = count;
  b->number_candidates = 0;
  return basis_finish(b) ? 0 : -1;
}

extern "C" int ls_amd_basis_state_info_device(ls_amd_basis const *b, int64_t count, uint64_t const *d_alphas,
                                              uint64_t *d_betas, double *d_characters, double *d_norms) {
  if (!b->group) return fail("basis has no permutation symmetries") ? 0 : -1;
  bool const ok = launch_state_info(*b->group, count, d_alphas, 1, d_betas, 1, d_characters, d_norms);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_basis_is_representative_device(ls_amd_basis const *b, int64_t count,
                                                     uint64_t const *d_alphas, uint8_t *d_flags, double *d_norms) {
  if (!b->group) return fail("basis has no permutation symmetries") ? 0 : -1;
  bool const ok = launch_is_representative(*b->group, count, d_alphas, 1, d_flags, d_norms);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_basis_index_device(ls_amd_basis const *b, int64_t count, uint64_t const *d_states,
                                         int64_t *d_indices) {
  if (!b->built()) return fail("basis is not built") ? 0 : -1;
  bool const ok = launch_state_index(b->index.view(), count, d_states, d_indices);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" ls_amd_operator *ls_amd_operator_create(ls_amd_basis const *b, ls_hs_nonbranching_terms const *diag,
                                                   ls_hs_nonbranching_terms const *off_diag) {
  if (!b->built()) {
    fail("basis is not built");
    return nullptr;
  }
  auto op = std::make_unique<ls_amd_operator>();
  op->basis = b;
  OperatorCore &c = op->core;
  c.number_bits = b->number_bits;
  c.spin_inversion = b->spin_inversion;
  c.number_states = b->number_states;
  if (b->group) {
    c.kind = kPermutations;
    c.group = b->group.get();
  } else if (b->spin_inversion != 0) {
    c.kind = kInversionOnly;
  } else {
    c.kind = kNoProjection;
  }
  c.identity_index = (!b->group && b->spin_inversion == 0 && b->hamming_weight < 0 && b->fermion_species != 2);
  c.index = b->index.view();
  if (!c.diag.load(diag) || !c.off.load(off_diag)) return nullptr;
  return op.release();
}

extern "C" void ls_amd_operator_destroy(ls_amd_operator *op) { delete op; }

extern "C" int ls_amd_matvec_f64(ls_amd_operator *op, double const *d_x, double *d_y, int accumulate) {
  bool const ok = operator_matvec_otf(op->core, d_x, d_y, accumulate != 0, false);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_matvec_c128(ls_amd_operator *op, double const *d_x, double *d_y) {
  bool const ok = operator_matvec_otf(op->core, d_x, d_y, false, true);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int64_t ls_amd_operator_build_csr(ls_amd_operator *op, int precision) {
  if (!operator_build_csr(op->core, precision)) return -1;
  return op->core.csr.nnz;
}

extern "C" int64_t ls_amd_operator_csr_nnz(ls_amd_operator const *op) { return op->core.csr.nnz; }

extern "C" int ls_amd_operator_copy_csr(ls_amd_operator const *op, int64_t *row_ptr, uint32_t *cols,
                                        double *values, double *diag) {
  Csr const &c = op->core.csr;
  if (!c.has64) return fail("FP64 CSR has not been built") ? 0 : -1;
  size_t const width = c.is_complex ? 2 : 1;
  if (c.wide) {
    std::vector<double> table(static_cast<size_t>(c.table_size));
    std::vector<uint16_t> idx(static_cast<size_t>(c.nnz));
    bool const ok = to_host(row_ptr, c.row_ptr.ptr, static_cast<size_t>(c.rows + 1)) &&
                    to_host(cols, c.cols.ptr, static_cast<size_t>(c.nnz)) &&
                    to_host(idx.data(), c.vidx.ptr, idx.size()) &&
                    to_host(table.data(), c.table64.ptr, table.size()) &&
                    to_host(diag, c.diag64.ptr, static_cast<size_t>(c.rows));
    if (!ok) return -1;
    for (int64_t k = 0; k < c.nnz; ++k) values[k] = table[idx[static_cast<size_t>(k)]];
    return 0;
  }
  if (c.packed) {
    std::vector<double> table(static_cast<size_t>(c.table_size));
    bool const ok = to_host(row_ptr, c.row_ptr.ptr, static_cast<size_t>(c.rows + 1)) &&
                    to_host(cols, c.cols.ptr, static_cast<size_t>(c.nnz)) &&
                    to_host(table.data(), c.table64.ptr, table.size()) &&
                    to_host(diag, c.diag64.ptr, static_cast<size_t>(c.rows));
    if (!ok) return -1;
    for (int64_t k = 0; k < c.nnz; ++k) {
      values[k] = table[cols[k] >> 24];
      cols[k] &= kPackedColumnMask;
    }
    return 0;
  }
  bool const ok = to_host(row_ptr, c.row_ptr.ptr, static_cast<size_t>(c.rows + 1)) &&
                  to_host(cols, c.cols.ptr, static_cast<size_t>(c.nnz)) &&
                  to_host(values, c.vals64.ptr, width * static_cast<size_t>(c.nnz)) &&
                  to_host(diag, c.diag64.ptr, static_cast<size_t>(c.rows));
  return ok ? 0 : -1;
}

extern "C" int ls_amd_csr_matvec_f64(ls_amd_operator *op, double const *d_x, double *d_y) {
  bool const ok = operator_csr_f64(op->core, d_x, d_y, false);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_csr_matvec_f32(ls_amd_operator *op, int k, float const *d_x, float *d_y) {
  bool const ok = operator_csr_f32_block(op->core, k, d_x, d_y);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_csr_matvec_f64_block(ls_amd_operator *op, int k, double const *d_x, double *d_y) {
  bool const ok = operator_csr_f64_block(op->core, k, d_x, d_y);
  return ok && hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_operator_autotune(ls_amd_operator *op) { return operator_autotune(op->core); }

extern "C" int ls_amd_operator_is_hermitian(ls_amd_operator *op, int mode) {
  if (op == nullptr) return fail("null argument") ? 0 : -1;
  if (mode < 0 || mode > 3) return fail("mode must be 0, 1, 2 or 3") ? 0 : -1;
  Csr const &c = op->core.csr;
  if ((mode == 0 && !c.has64) || (mode == 1 && !c.has32)) return fail("build the CSR in that precision first") ? 0 : -1;
  if (mode < 2 && c.is_complex) return fail("the CSR holds complex values: use mode 3") ? 0 : -1;
  bool h = false;
  if (!operator_is_hermitian(op->core, mode, h)) return -1;
  return h ? 1 : 0;
}

extern "C" int ls_amd_operator_set_csr_packing(ls_amd_operator *op, int enable) {
  if (op == nullptr) return fail("null argument") ? 0 : -1;
  op->core.allow_packing = enable != 0;
  return 0;
}

extern "C" int ls_amd_operator_set_factored(ls_amd_operator *op, int enable) {
  if (op == nullptr) return fail("null argument") ? 0 : -1;
  op->core.allow_factored = enable != 0;
  return 0;
}

extern "C" int ls_amd_operator_uses_factored(ls_amd_operator const *op) {
  if (op == nullptr) return fail("null argument") ? 0 : -1;
  OperatorCore const &c = op->core;
  return (c.kind == kPermutations && c.allow_factored && c.group->factored()) ? 1 : 0;
}

extern "C" int ls_amd_operator_csr_info(ls_amd_operator const *op, int64_t *out) {
  if (op == nullptr || out == nullptr) return fail("null argument") ? 0 : -1;
  Csr const &c = op->core.csr;
  out[0] = c.nnz;
  out[1] = c.packed ? 1 : (c.wide ? 2 : 0);
  out[2] = c.table_size;
  out[3] = c.is_complex ? 1 : 0;
  return 0;
}

extern "C" int ls_amd_operator_schedule(ls_amd_operator const *op, int *out) {
  if (op == nullptr || out == nullptr) return fail("null argument") ? 0 : -1;
  out[0] = op->core.lanes_f64;
  out[1] = op->core.flags_f64;
  out[2] = op->core.lanes_f32;
  out[3] = op->core.flags_f32;
  out[4] = op->core.persistent_blocks;
  return 0;
}

extern "C" int ls_amd_set_factored_enumeration(int enable) {
  g_factored_enumeration.store(enable != 0);
  return 0;
}

extern "C" int ls_amd_operator_set_schedule(ls_amd_operator *op, int lanes_f64, int flags_f64, int lanes_f32,
                                            int flags_f32) {
  if (op == nullptr) return fail("null argument") ? 0 : -1;
  for (int l : {lanes_f64, lanes_f32})
    if (l < 1 || l > 64 || (l & (l - 1)) != 0) return fail("lanes per row must be a power of two <= 64") ? 0 : -1;
  for (int f : {flags_f64, flags_f32})
    if (f < 0 || f > 7 || (f == 4 || f == 5)) return fail("schedule flags are 0..3, 6 or 7") ? 0 : -1;
  op->core.lanes_f64 = lanes_f64;
  op->core.flags_f64 = flags_f64;
  op->core.lanes_f32 = lanes_f32;
  op->core.flags_f32 = flags_f32;
  op->core.schedule_set = true;
  return 0;
}
