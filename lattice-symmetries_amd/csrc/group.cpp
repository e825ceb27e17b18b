// group.cpp -- permutation groups, characters and Benes networks (host).
//
// Restates the set-up that the reference does in Haskell
// (haskell/src/LatticeSymmetries/Group.hs, Benes.hs) so that the engine can
// build its own ls_hs_permutation_group, and uploads a group to the device
// for ls_internal_create_halide_kernel_data (kernels/kernels.c:112-189).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <set>
#include <utility>

#include "lsamd_internal.hpp"

namespace lsamd {
namespace {

using Perm = std::vector<int32_t>;

// Phase as a reduced fraction num/den in [0, 1).
struct Phase {
  int64_t num = 0, den = 1;
  bool operator<(Phase const &o) const {
    return static_cast<__int128>(num) * o.den < static_cast<__int128>(o.num) * den;
  }
  bool operator==(Phase const &o) const { return num == o.num && den == o.den; }
};

Phase make_phase(int64_t num, int64_t den) {
  int64_t g = std::gcd(num, den);
  if (g == 0) g = 1;
  num /= g;
  den /= g;
  if (den < 0) {
    num = -num;
    den = -den;
  }
  num %= den; // phases live in [0, 1) (Group.hs:90-93)
  if (num < 0) num += den;
  if (num == 0) den = 1;
  return Phase{num, den};
}

Phase add_phase(Phase a, Phase b) {
  return make_phase(a.num * b.den + b.num * a.den, a.den * b.den);
}

// (a <> b)[i] = b[a[i]]: apply a's relabelling first (Benes.hs:100-101).
Perm compose(Perm const &a, Perm const &b) {
  Perm c(a.size());
  for (size_t i = 0; i < a.size(); ++i) c[i] = b[static_cast<size_t>(a[i])];
  return c;
}

bool is_identity(Perm const &p) {
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i] != static_cast<int32_t>(i)) return false;
  return true;
}

int order_of(Perm const &p) {
  Perm cur = p;
  int k = 1;
  while (!is_identity(cur)) {
    cur = compose(p, cur);
    ++k;
  }
  return k;
}

// ---------------------------------------------------------------------------
// Benes network routing.
//
// For N = 2^L positions, element q must travel to position dst[q].  The
// network is: a delta-swap stage with shift 1 on the input side, two
// independent networks of size N/2 on the even and the odd positions, and a
// delta-swap stage with shift 1 on the output side.  Inputs 2k, 2k+1 must be
// split between the halves, and so must the sources of outputs 2k, 2k+1; the
// classic looping argument 2-colours these constraints cycle by cycle.
// Sub-network stages with shift d act on global positions with shift 2d.
// ---------------------------------------------------------------------------
void route(std::vector<int> const &src_of, // src_of[i]: input that lands at output i
           int L, std::vector<std::vector<uint64_t>> &stage_masks, int stage_lo,
           int pos_mul, int pos_add) {
  int const N = 1 << L;
  int const depth = 2 * L - 1;
  if (L == 0) return;
  if (L == 1) {
    if (src_of[0] == 1)
      stage_masks[stage_lo][0] |= uint64_t{1} << pos_add; // swap positions pos_add, pos_add+pos_mul
    return;
  }
  std::vector<int> dst(N);
  for (int i = 0; i < N; ++i) dst[src_of[i]] = i;
  std::vector<int> colour(N, -1); // per input element: 0 even half, 1 odd half
  for (int start = 0; start < N; ++start) {
    if (colour[start] != -1) continue;
    int q = start;
    int c = 0;
    while (colour[q] == -1) {
      colour[q] = c;
      int const partner = q ^ 1; // input-pair partner takes the other half
      colour[partner] = 1 - c;
      // the output paired with partner's output must come from half c
      int const o = dst[partner] ^ 1;
      q = src_of[o];
      // q must take colour c (different from partner's 1 - c)
    }
  }
  int const first = stage_lo, last = stage_lo + depth - 1;
  std::vector<int> sub_src[2];
  sub_src[0].resize(N / 2);
  sub_src[1].resize(N / 2);
  for (int k = 0; k < N / 2; ++k) {
    // input stage: swap when the even input belongs to the odd half
    if (colour[2 * k] == 1) stage_masks[first][0] |= uint64_t{1} << (pos_add + pos_mul * 2 * k);
    // output stage: swap when output 2k is fed by the odd half
    int const s0 = src_of[2 * k], s1 = src_of[2 * k + 1];
    if (colour[s0] == 1) stage_masks[last][0] |= uint64_t{1} << (pos_add + pos_mul * 2 * k);
    int const even_src = colour[s0] == 0 ? s0 : s1;
    int const odd_src = colour[s0] == 0 ? s1 : s0;
    sub_src[0][k] = even_src >> 1;
    sub_src[1][k] = odd_src >> 1;
  }
  route(sub_src[0], L - 1, stage_masks, stage_lo + 1, pos_mul * 2, pos_add);
  route(sub_src[1], L - 1, stage_masks, stage_lo + 1, pos_mul * 2, pos_add + pos_mul);
}

// Stage masks and shifts for one permutation p of n sites (bit i of g(x)
// is bit p[i] of x), padded with fixed points to N = 2^L >= n
// (Benes.hs:308-314).
void benes_network(Perm const &p, std::vector<uint64_t> &masks, std::vector<uint64_t> &shifts) {
  int const n = static_cast<int>(p.size());
  int L = 0;
  while ((1 << L) < n) ++L;
  int const N = 1 << L;
  int const depth = L == 0 ? 0 : 2 * L - 1;
  std::vector<int> src_of(N);
  for (int i = 0; i < N; ++i) src_of[i] = i < n ? p[i] : i;
  std::vector<std::vector<uint64_t>> stage(depth, std::vector<uint64_t>(1, 0));
  route(src_of, L, stage, 0, 1, 0);
  masks.resize(depth);
  shifts.resize(depth);
  for (int k = 0; k < depth; ++k) {
    masks[k] = stage[k][0];
    shifts[k] = static_cast<uint64_t>(k < L ? (1 << k) : (1 << (depth - 1 - k)));
  }
}

uint64_t apply_network(std::vector<uint64_t> const &masks, std::vector<uint64_t> const &shifts,
                       uint64_t x) {
  for (size_t k = 0; k < masks.size(); ++k) {
    unsigned const d = static_cast<unsigned>(shifts[k]);
    uint64_t const y = ((x >> d) ^ x) & masks[k];
    x = (x ^ y) ^ (y << d);
  }
  return x;
}

uint64_t permute_direct(Perm const &p, uint64_t x) {
  uint64_t y = 0;
  for (size_t i = 0; i < p.size(); ++i) y |= ((x >> p[i]) & 1u) << i;
  return y;
}

struct PermGroupOwner {
  std::vector<uint64_t> masks, shifts;
  std::vector<double> re, im;
  std::vector<int32_t> perms; // [size][n]
};

} // namespace

GroupData *group_data_create(ls_hs_permutation_group const *g, int spin_inversion) {
  if (g == nullptr) {
    fail("ls_internal_create_halide_kernel_data: group is NULL");
    return nullptr;
  }
  if (spin_inversion < -1 || spin_inversion > 1) {
    fail("invalid spin_inversion"); // kernels.c:185-187
    return nullptr;
  }
  if (g->number_bits < 0 || g->number_bits > 64) {
    fail("too many bits");
    return nullptr;
  }
  if (!ensure_device()) return nullptr;
  auto d = std::make_unique<GroupData>();
  d->number_bits = g->number_bits;
  d->size = g->number_masks;
  d->depth = g->number_shifts;
  d->spin_inversion = spin_inversion;
  d->flip_mask = g->number_bits == 0 ? 0 : (~uint64_t{0}) >> (64 - g->number_bits);
  int const G = d->size, D = d->depth;
  d->h_masks.resize(static_cast<size_t>(G) * (D > 0 ? D : 1));
  for (int e = 0; e < G; ++e)
    for (int k = 0; k < D; ++k)
      d->h_masks[static_cast<size_t>(e) * D + k] = g->masks[static_cast<size_t>(k) * G + e];
  d->h_shifts.assign(g->shifts, g->shifts + D);
  d->h_eig_re.assign(g->eigvals_re, g->eigvals_re + G);
  d->h_eig_im.assign(g->eigvals_im, g->eigvals_im + G);
  // canonical shifts 1, 2, ..., 2^(L-1), ..., 2, 1 (D = 2L - 1)
  d->canonical_shifts = (D % 2 == 1);
  if (d->canonical_shifts) {
    int const L = (D + 1) / 2;
    for (int k = 0; k < D && d->canonical_shifts; ++k) {
      int const want = k < L ? (1 << k) : (1 << (D - 1 - k));
      if (d->h_shifts[k] != static_cast<uint64_t>(want)) d->canonical_shifts = false;
    }
  }
  // 32-bit words are used for depth <= 9: masks must then fit in 32 bits.
  if (d->canonical_shifts && D <= 9) {
    for (uint64_t m : d->h_masks)
      if (m >> 32) d->canonical_shifts = false;
  }
  if (!d->d_masks.upload(d->h_masks.data(), d->h_masks.size())) return nullptr;
  std::vector<uint64_t> sh = d->h_shifts;
  if (sh.empty()) sh.push_back(0);
  if (!d->d_shifts.upload(sh.data(), sh.size())) return nullptr;
  std::vector<double> re = d->h_eig_re, im = d->h_eig_im;
  if (re.empty()) {
    re.push_back(1.0);
    im.push_back(0.0);
  }
  if (!d->d_eig_re.upload(re.data(), re.size())) return nullptr;
  if (!d->d_eig_im.upload(im.data(), im.size())) return nullptr;
  if (!group_detect_translations(*d)) return nullptr;
  return d.release();
}

bool group_detect_translations(GroupData &g) {
  g.fact_lx = g.fact_ly = 0;
  g.h_fact_q.clear();
  g.h_fact_elem.clear();
  g.h_fact_key.clear();
  int const n = g.number_bits, G = g.size, D = g.depth;
  if (n < 2 || G < 2) return true;
  for (int e = 0; e < G; ++e)
    if (g.h_eig_re[e] != 1.0 && g.h_eig_re[e] != -1.0) return true;
  // permutations: bit i of g(x) is bit p[i] of x
  std::vector<Perm> perms(G, Perm(n));
  std::map<Perm, int> index;
  for (int e = 0; e < G; ++e) {
    std::vector<uint64_t> m(g.h_masks.begin() + static_cast<size_t>(e) * D,
                            g.h_masks.begin() + static_cast<size_t>(e + 1) * D);
    for (int b = 0; b < n; ++b) {
      uint64_t const y = apply_network(m, g.h_shifts, uint64_t{1} << b);
      if (y == 0 || (y & (y - 1)) != 0) return true;
      perms[e][__builtin_ctzll(y)] = b;
    }
    index.emplace(perms[e], e);
  }
  if (static_cast<int>(index.size()) != G) return true;
  auto translation = [n](int lx, int ly, int a, int b) {
    Perm p(n);
    for (int r = 0; r < ly; ++r)
      for (int c = 0; c < lx; ++c) p[r * lx + c] = ((r + b) % ly) * lx + (c + a) % lx;
    return p;
  };
  for (int lx = n; lx >= 1; --lx) {
    if (n % lx != 0) continue;
    int const ly = n / lx;
    if (lx * ly < 2) continue;
    bool all = true;
    for (int b = 0; b < ly && all; ++b)
      for (int a = 0; a < lx && all; ++a) all = index.count(translation(lx, ly, a, b)) > 0;
    if (!all) continue;
    std::vector<int32_t> q;
    for (int e = 0; e < G; ++e)
      if (perms[e][0] == 0) q.push_back(e);
    if (static_cast<int64_t>(q.size()) * lx * ly != G) continue;
    std::vector<int32_t> elem(static_cast<size_t>(G));
    std::vector<char> hit(G, 0);
    bool ok = true;
    for (size_t k = 0; k < q.size() && ok; ++k)
      for (int b = 0; b < ly && ok; ++b)
        for (int a = 0; a < lx && ok; ++a) {
          Perm const t = translation(lx, ly, a, b);
          Perm p(n);
          for (int i = 0; i < n; ++i) p[i] = perms[q[k]][t[i]]; // t(q(x))
          auto it = index.find(p);
          if (it == index.end() || hit[it->second]) {
            ok = false;
            break;
          }
          hit[it->second] = 1;
          elem[(k * ly + b) * lx + a] = it->second;
        }
    if (!ok) continue;
    uint64_t m1 = 0, m2 = 0;
    for (int r = 0; r < ly; ++r)
      for (int c = 0; c < lx; ++c) (c < lx - 1 ? m1 : m2) |= uint64_t{1} << (r * lx + c);
    g.fact_lx = lx;
    g.fact_ly = ly;
    g.fact_m1 = m1;
    g.fact_m2 = m2;
    // rank of element e (unflipped; its flipped copy ranks one higher) in the
    // eight-lane state_info reduction: lane e mod 8, then position e / 8 in
    // the lane, then the |G| mod 8 tail elements in order
    int const lanes = 8, chunks = G / lanes;
    std::vector<int32_t> key(static_cast<size_t>(G));
    for (size_t k = 0; k < elem.size(); ++k) {
      int const e = elem[k];
      key[k] = e < chunks * lanes ? (e % lanes) * 2 * chunks + (e / lanes) * 2
                                  : lanes * 2 * chunks + (e - chunks * lanes) * 2;
    }
    g.h_fact_q = q;
    g.h_fact_elem = elem;
    g.h_fact_key = key;
    return g.d_fact_q.upload(q.data(), q.size()) && g.d_fact_elem.upload(elem.data(), elem.size()) &&
           g.d_fact_key.upload(key.data(), key.size());
  }
  return true;
}

// Order in which the enumeration filter tries the group elements: the
// predicate "g(x) >= x for every g" does not depend on the order, so the
// elements that reject most candidates go first.  Greedy over a seeded
// sample of candidates: each step takes the element that rejects the most
// sample states not yet rejected; the rest follow by rejection count.
std::vector<int> group_rejection_order(GroupData const &g, int hamming, int number_bits) {
  int const G = g.size, depth = g.depth;
  std::vector<int> order(G);
  for (int e = 0; e < G; ++e) order[e] = e;
  if (G <= 2 || number_bits <= 0 || number_bits > 64) return order;
  int const samples = 1024;
  std::vector<uint64_t> xs(samples);
  uint64_t z = 0x243f6a8885a308d3ull;
  auto next = [&]() {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ull;
    r = (r ^ (r >> 27)) * 0x94d049bb133111ebull;
    return r ^ (r >> 31);
  };
  uint64_t const all = number_bits == 64 ? ~uint64_t{0} : ((uint64_t{1} << number_bits) - 1);
  for (auto &x : xs) {
    if (hamming >= 0 && hamming <= number_bits) {
      x = 0;
      int placed = 0;
      while (placed < hamming) {
        int const b = static_cast<int>(next() % static_cast<uint64_t>(number_bits));
        if (!((x >> b) & 1)) {
          x |= uint64_t{1} << b;
          ++placed;
        }
      }
    } else {
      x = next() & all;
    }
  }
  std::vector<std::vector<uint8_t>> rej(G, std::vector<uint8_t>(samples));
  std::vector<int> count(G, 0);
  for (int e = 0; e < G; ++e) {
    std::vector<uint64_t> m(g.h_masks.begin() + static_cast<size_t>(e) * depth,
                            g.h_masks.begin() + static_cast<size_t>(e + 1) * depth);
    for (int s = 0; s < samples; ++s) {
      uint64_t const y = apply_network(m, g.h_shifts, xs[s]);
      bool r = y < xs[s];
      if (g.spin_inversion != 0) r = r || ((y ^ g.flip_mask) < xs[s]);
      rej[e][s] = r ? 1 : 0;
      count[e] += r ? 1 : 0;
    }
  }
  std::vector<uint8_t> alive(samples, 1), used(G, 0);
  std::vector<int> out;
  for (int step = 0; step < std::min(G, 32); ++step) {
    int best = -1, best_gain = 0;
    for (int e = 0; e < G; ++e) {
      if (used[e]) continue;
      int gain = 0;
      for (int s = 0; s < samples; ++s) gain += alive[s] & rej[e][s];
      if (gain > best_gain) {
        best_gain = gain;
        best = e;
      }
    }
    if (best < 0) break;
    used[best] = 1;
    out.push_back(best);
    for (int s = 0; s < samples; ++s)
      if (rej[best][s]) alive[s] = 0;
  }
  std::vector<int> rest;
  for (int e = 0; e < G; ++e)
    if (!used[e]) rest.push_back(e);
  std::stable_sort(rest.begin(), rest.end(), [&](int a, int b) { return count[a] > count[b]; });
  out.insert(out.end(), rest.begin(), rest.end());
  return out;
}

bool group_make_fermionic(GroupData &g, int species) {
  if (species != 1 && species != 2) return fail("fermion species must be 1 or 2");
  if (g.spin_inversion != 0) return fail("fermionic bases have no spin inversion");
  if (species == 2 && g.number_bits % 2 != 0) return fail("spinful fermions need an even number of bits");
  int const n = g.number_bits, G = g.size, D = g.depth;
  std::vector<uint8_t> image(static_cast<size_t>(G) * n);
  for (int e = 0; e < G; ++e) {
    std::vector<uint64_t> m(g.h_masks.begin() + static_cast<size_t>(e) * D,
                            g.h_masks.begin() + static_cast<size_t>(e + 1) * D);
    for (int b = 0; b < n; ++b) {
      uint64_t const y = apply_network(m, g.h_shifts, uint64_t{1} << b);
      if (y == 0 || (y & (y - 1)) != 0) return fail("group element is not a permutation of the bits");
      image[static_cast<size_t>(e) * n + b] = static_cast<uint8_t>(__builtin_ctzll(y));
    }
    if (species == 2) // the same site permutation on both species
      for (int b = 0; b < n / 2; ++b)
        if (image[static_cast<size_t>(e) * n + b + n / 2] != image[static_cast<size_t>(e) * n + b] + n / 2)
          return fail("spinful group elements must act identically on both species");
  }
  if (!g.d_image.upload(image.data(), image.size())) return false;
  g.fermion_species = species;
  return true;
}

} // namespace lsamd

using namespace lsamd;

extern "C" ls_hs_permutation_group *
ls_amd_permutation_group_create(int n, int number_generators, int32_t const *permutations,
                                int32_t const *sectors) {
  if (n <= 0 || n > 64) {
    fail("number of sites must be in [1, 64]");
    return nullptr;
  }
  if (number_generators < 0 || (number_generators > 0 && (permutations == nullptr || sectors == nullptr))) {
    fail("invalid generators");
    return nullptr;
  }
  // Generators (Group.hs:78-85: sector in [0, periodicity)).
  std::vector<std::pair<Perm, Phase>> gens;
  for (int k = 0; k < number_generators; ++k) {
    Perm p(permutations + static_cast<size_t>(k) * n, permutations + static_cast<size_t>(k + 1) * n);
    std::vector<char> seen(n, 0);
    for (int v : p) {
      if (v < 0 || v >= n || seen[v]) {
        fail("invalid permutation");
        return nullptr;
      }
      seen[v] = 1;
    }
    int const per = order_of(p);
    if (sectors[k] < 0 || sectors[k] >= per) {
      fail("invalid sector " + std::to_string(sectors[k]) + "; permutation has periodicity " +
           std::to_string(per));
      return nullptr;
    }
    gens.emplace_back(p, make_phase(sectors[k], per));
  }
  // Closure by breadth-first products h <> g (Group.hs:174-183), kept as an
  // ordered set of (permutation, phase) so that the result is ascending.
  std::set<std::pair<Perm, Phase>> interior;
  std::set<std::pair<Perm, Phase>> boundary;
  if (!gens.empty()) {
    Perm id(n);
    std::iota(id.begin(), id.end(), 0);
    boundary.insert({id, Phase{0, 1}});
  }
  while (!boundary.empty()) {
    interior.insert(boundary.begin(), boundary.end());
    std::set<std::pair<Perm, Phase>> next;
    for (auto const &h : boundary)
      for (auto const &g : gens) {
        std::pair<Perm, Phase> s{compose(h.first, g.first), add_phase(h.second, g.second)};
        if (interior.count(s) == 0) next.insert(std::move(s));
      }
    boundary = std::move(next);
    if (interior.size() > (1u << 20)) {
      fail("group too large");
      return nullptr;
    }
  }
  // Consistency (Group.hs:124-131): one phase per permutation and integral
  // sector for every product.
  std::map<Perm, Phase> phase_of;
  for (auto const &s : interior) {
    auto it = phase_of.find(s.first);
    if (it != phase_of.end()) {
      fail("incompatible symmetries");
      return nullptr;
    }
    phase_of.emplace(s.first, s.second);
  }
  for (auto const &a : interior)
    for (auto const &b : interior) {
      Perm const c = compose(a.first, b.first);
      Phase const ph = add_phase(a.second, b.second);
      auto it = phase_of.find(c);
      if (it == phase_of.end() || !(it->second == ph) ||
          (static_cast<int64_t>(order_of(c)) * ph.num) % ph.den != 0) {
        fail("incompatible symmetries");
        return nullptr;
      }
    }
  auto *owner = new PermGroupOwner();
  int const G = static_cast<int>(interior.size());
  int L = 0;
  while ((1 << L) < n) ++L;
  int const depth = L == 0 ? 0 : 2 * L - 1;
  owner->masks.assign(static_cast<size_t>(depth) * G, 0);
  owner->shifts.assign(depth, 0);
  owner->re.resize(G);
  owner->im.resize(G);
  owner->perms.resize(static_cast<size_t>(G) * n);
  int e = 0;
  for (auto const &s : interior) {
    std::vector<uint64_t> m, sh;
    benes_network(s.first, m, sh);
    // self-check of the network against the permutation on basis vectors
    for (int i = 0; i < n; ++i) {
      uint64_t const x = uint64_t{1} << i;
      if (apply_network(m, sh, x) != permute_direct(s.first, x)) {
        delete owner;
        fail("internal error: Benes network does not realise the permutation");
        return nullptr;
      }
    }
    for (int k = 0; k < depth; ++k) owner->masks[static_cast<size_t>(k) * G + e] = m[k];
    owner->shifts = sh;
    // Group.hs:115-116
    double const phi = static_cast<double>(s.second.num) / static_cast<double>(s.second.den);
    double const arg = -(2.0 * M_PI * phi);
    owner->re[e] = std::cos(arg);
    owner->im[e] = std::sin(arg);
    std::copy(s.first.begin(), s.first.end(), owner->perms.begin() + static_cast<size_t>(e) * n);
    ++e;
  }
  auto *g = new ls_hs_permutation_group{};
  g->refcount = 1;
  g->number_bits = n;
  g->number_shifts = depth;
  g->number_masks = G;
  g->masks = owner->masks.data();
  g->shifts = owner->shifts.data();
  g->eigvals_re = owner->re.data();
  g->eigvals_im = owner->im.data();
  g->haskell_payload = owner;
  return g;
}

extern "C" void ls_amd_permutation_group_destroy(ls_hs_permutation_group *g) {
  if (g == nullptr) return;
  delete static_cast<PermGroupOwner *>(g->haskell_payload);
  delete g;
}

extern "C" int ls_amd_permutation_group_elements(ls_hs_permutation_group const *g, int32_t *out) {
  if (g == nullptr || out == nullptr || g->haskell_payload == nullptr) return -1;
  auto const *owner = static_cast<PermGroupOwner const *>(g->haskell_payload);
  std::copy(owner->perms.begin(), owner->perms.end(), out);
  return 0;
}
