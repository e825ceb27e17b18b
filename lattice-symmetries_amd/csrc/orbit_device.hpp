// orbit_device.hpp -- device functions for group orbits on gfx950.
//
// A state is a machine word W (uint32_t when the Benes network spans at most
// 32 bits, else uint64_t).  One lane owns one state and walks the group
// elements in their stored (ascending) order; element indices are
// wave-uniform, so masks and characters are read through scalar loads.
//
// The reductions reproduce the order of the reference's Halide kernels
// (kernels/generator.cpp) literally, so representatives, selected characters
// and the floating-point norms agree bit for bit with a sequential
// restatement of those kernels.
#pragma once

#include "lsamd_internal.hpp"

namespace lsamd {

// Shift of stage k in a Benes network of depth D built for N = 2^L bits:
// 1, 2, ..., N/2, N/4, ..., 1 (Benes.hs:295-323).
template <int D> __host__ __device__ constexpr int benes_shift(int k) {
  return k < (D + 1) / 2 ? (1 << k) : (1 << (D - 1 - k));
}

// One delta swap (generator.cpp:5-8): exchange bits i and i+d for every i
// set in m.
template <typename W>
__device__ __forceinline__ W delta_swap(W x, W m, int d) {
  W const y = ((x >> d) ^ x) & m;
  return (x ^ y) ^ (y << d);
}

// Apply the Benes network of group element e.  D > 0: compile-time
// canonical shifts; D == 0: depth and shifts read from the view.
template <typename W, int D>
__device__ __forceinline__ W group_apply(GroupView const &g, int e, W x) {
  if constexpr (D > 0) {
    uint64_t const *m = g.masks + static_cast<size_t>(e) * D;
#pragma unroll
    for (int k = 0; k < D; ++k) x = delta_swap<W>(x, static_cast<W>(m[k]), benes_shift<D>(k));
    return x;
  } else {
    int const depth = g.depth;
    uint64_t const *m = g.masks + static_cast<size_t>(e) * depth;
    for (int k = 0; k < depth; ++k)
      x = delta_swap<W>(x, static_cast<W>(m[k]), static_cast<int>(g.shifts[k]));
    return x;
  }
}

// Same, with the masks of element e given as a pointer (e.g. into LDS).
template <typename W, int D>
__device__ __forceinline__ W benes_apply_ptr(W const *m, W x,
                                             uint64_t const *shifts, int depth) {
  if constexpr (D > 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) x = delta_swap<W>(x, m[k], benes_shift<D>(k));
    return x;
  } else {
    for (int k = 0; k < depth; ++k)
      x = delta_swap<W>(x, m[k], static_cast<int>(shifts[k]));
    return x;
  }
}

// ---------------------------------------------------------------------------
// state_info (generator.cpp:10-166)
// ---------------------------------------------------------------------------
//
// Eight accumulators (the Halide chunk 2 * natural_vector_size(uint64) of the
// avx2 target) each see elements e = 8c + j in order; each keeps the smallest
// image and the first element reaching it, and sums Re(chi) over elements
// fixing x.  Accumulators are merged left to right (earlier lane wins ties),
// then the remaining |G| mod 8 elements are folded in one by one.
// The selected element is kept as an index (2e + flip) and its character is
// read at the end; sel == -1 means no image is smaller than x (chi = 1).

constexpr int kStateInfoLanes = 8;
constexpr int kIsRepresentativeLanes = 4;

// SI selects the symmetry flavour of the orbit kernels: 0 no spin
// inversion, +-1 spin inversion with that character, 2 fermionic signs (no
// inversion).
constexpr int kFermionic = 2;

// Sign of group element e on state x for fermions: the state is the ordered
// product of creation operators (bit order, the Jordan-Wigner order of the
// sign masks); e relabels orbital b as image[b], and the sign is the parity
// of the inversions among the relabelled occupied orbitals.
__device__ __forceinline__ double fermion_sign(GroupView const &g, int e, uint64_t x) {
  uint8_t const *img = g.image + static_cast<size_t>(e) * g.nbits;
  uint64_t seen = 0;
  int parity = 0;
  while (x != 0) {
    int const b = __builtin_ctzll(x);
    x &= x - 1;
    int const i = img[b];
    parity ^= __popcll(seen >> i) & 1; // earlier orbitals now placed above i
    seen |= uint64_t{1} << i;
  }
  return parity ? -1.0 : 1.0;
}

template <typename W>
__device__ __forceinline__ void si_update(W &r, int &sel, double &n, W x, W y,
                                          int code, double re) {
  if (y < r) {
    r = y;
    sel = code;
  }
  if (y == x) n = n + re;
}

struct StateInfo {
  uint64_t rep;
  double c_re, c_im;
  double norm;
};

template <typename W, int D, int SI>
__device__ __forceinline__ StateInfo state_info_one(GroupView const &g,
                                                    uint64_t x64) {
  constexpr int L = kStateInfoLanes;
  W const x = static_cast<W>(x64);
  W const flip = static_cast<W>(g.flip);
  W r[L];
  int sel[L];
  double n[L];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    r[j] = x;
    sel[j] = -1;
    n[j] = 0.0;
  }
  int const G = g.size;
  int const nch = G / L;
  constexpr bool kFermi = SI == kFermionic;
  for (int c = 0; c < nch; ++c) {
#pragma unroll
    for (int j = 0; j < L; ++j) {
      int const e = c * L + j;
      W const y = group_apply<W, D>(g, e, x);
      double re = g.eig_re[e];
      if constexpr (kFermi) {
        if (y == x) re = re * fermion_sign(g, e, x64);
        si_update<W>(r[j], sel[j], n[j], x, y, 2 * e, re);
      } else {
        si_update<W>(r[j], sel[j], n[j], x, y, 2 * e, re);
        if constexpr (SI != 0)
          si_update<W>(r[j], sel[j], n[j], x, static_cast<W>(y ^ flip), 2 * e + 1,
                       SI == 1 ? re : -re);
      }
    }
  }
  W rr = r[0];
  int ss = sel[0];
  double nn = n[0];
#pragma unroll
  for (int j = 1; j < L; ++j) {
    if (r[j] < rr) {
      rr = r[j];
      ss = sel[j];
    }
    nn = nn + n[j];
  }
  for (int e = nch * L; e < G; ++e) {
    W const y = group_apply<W, D>(g, e, x);
    double re = g.eig_re[e];
    if constexpr (kFermi) {
      if (y == x) re = re * fermion_sign(g, e, x64);
      si_update<W>(rr, ss, nn, x, y, 2 * e, re);
    } else {
      si_update<W>(rr, ss, nn, x, y, 2 * e, re);
      if constexpr (SI != 0)
        si_update<W>(rr, ss, nn, x, static_cast<W>(y ^ flip), 2 * e + 1,
                     SI == 1 ? re : -re);
    }
  }
  StateInfo out;
  out.rep = static_cast<uint64_t>(rr);
  if (ss < 0) {
    out.c_re = 1.0;
    out.c_im = 0.0;
  } else {
    int const e = ss >> 1;
    double cre = g.eig_re[e], cim = g.eig_im[e];
    if (SI == -1 && (ss & 1)) {
      cre = -cre;
      cim = -cim;
    }
    if constexpr (kFermi) {
      double const sg = fermion_sign(g, e, x64);
      cre = cre * sg;
      cim = cim * sg;
    }
    out.c_re = cre;
    out.c_im = cim;
  }
  double const denom = (SI == 0 || kFermi) ? static_cast<double>(G) : static_cast<double>(2 * G);
  out.norm = (g.norm_eps > 0.0 && !(nn > g.norm_eps)) ? 0.0 : sqrt(nn / denom);
  return out;
}

// state_info through the translation factorisation of the group (fact_nq >
// 0, real parts of the characters +-1, no fermionic signs): for every point
// element q, y = q(x) by its Benes network, then the translations t_{a,b}(y)
// of the torus by column and row rotations (a few word operations each
// instead of a network per element).  The orbit minimum does not depend on
// the visiting order, and the stabiliser sum of +-1 terms is exact in any
// order.  The character is that of the minimiser state_info_one selects:
// fact_key ranks each element (and its spin-flipped copy, key + 1) by the
// order in which the eight-lane reduction reaches it -- lane e mod 8 first,
// then position within the lane, the |G| mod 8 tail elements last -- and
// among equal images the lowest rank wins, as the strict comparisons of
// that reduction decide.  rep, character and norm equal state_info_one's
// bit for bit.
//
// With row tables (rows of at most kRowTableBits sites), the translations of
// q(x) are not all formed: the least image has the least top row, and the
// top row of t_(a, b)(y) is row (b - 1) mod ly of y rotated by a, so the
// least rotation of each row (a table lookup) gives the least top row and
// the few (a, b) that reach it; only those images are formed and compared
// (none when that row exceeds the top row of the current minimum).  Images
// equal to x need a row in the rotation class of x's top row, found the
// same way.  Every image equal to the minimum, or to x, is still visited,
// so the result is the same.
constexpr int kRowTableBits = 8;

struct RowTables {
  uint8_t minrot[1 << kRowTableBits]; // least rotation of a row value
  uint8_t amin[1 << kRowTableBits];   // rotations a that reach it (bit a)
  uint64_t lo[kRowTableBits], hi[kRowTableBits];
};

__device__ __forceinline__ bool row_tables_apply(GroupView const &g) {
  return g.fact_nq > 0 && g.fact_lx >= 1 && g.fact_lx <= kRowTableBits && g.fact_ly <= kRowTableBits;
}

// Fills the tables (every thread of the workgroup must call it).
__device__ inline void row_tables_init(GroupView const &g, RowTables &t) {
  if (row_tables_apply(g)) {
    int const lx = g.fact_lx;
    uint32_t const rm = (1u << lx) - 1;
    for (uint32_t v = threadIdx.x; v <= rm; v += blockDim.x) {
      uint32_t best = 0xFFu, mask = 0;
      for (int a = 0; a < lx; ++a) {
        uint32_t const r = ((v >> a) | (v << (lx - a))) & rm; // rotation by a
        if (r < best) {
          best = r;
          mask = 1u << a;
        } else if (r == best) {
          mask |= 1u << a;
        }
      }
      t.minrot[v] = static_cast<uint8_t>(best);
      t.amin[v] = static_cast<uint8_t>(mask);
    }
    if (static_cast<int>(threadIdx.x) < lx) {
      int const a = static_cast<int>(threadIdx.x);
      uint64_t lo = 0, hi = 0;
      for (int r = 0; r < g.fact_ly; ++r)
        for (int c = 0; c < lx; ++c) (c < lx - a ? lo : hi) |= uint64_t{1} << (r * lx + c);
      t.lo[a] = lo;
      t.hi[a] = hi;
    }
  }
  __syncthreads();
}

template <typename W, int D, int SI>
__device__ __forceinline__ StateInfo state_info_factored(GroupView const &g, uint64_t x64,
                                                        RowTables const *rt = nullptr) {
  W const x = static_cast<W>(x64);
  W const flip = static_cast<W>(g.flip);
  W const m1 = static_cast<W>(g.fact_m1), m2 = static_cast<W>(g.fact_m2);
  int const lx = g.fact_lx, ly = g.fact_ly, nq = g.fact_nq;
  int const up = g.nbits - lx; // row rotation: the bottom row moves to the top
  // The candidate (q, s, b, a) has position k = (q * ly + b) * lx + a in the
  // tables and code 2 k + s; the minimiser is kept as a code (-1: x itself)
  // and the tables are read only on the rare ties and stabiliser hits, which
  // the branches skip when no lane of the wave has one.
  W best = x;
  int best_code = -1;
  double nn = 0.0;
  if (rt != nullptr && row_tables_apply(g)) {
    uint32_t const rm = (1u << lx) - 1;
    int const top = (ly - 1) * lx;
    uint32_t const xt = static_cast<uint32_t>(x >> top) & rm;
    uint32_t const xcls = rt->minrot[xt];
    // t_(0, b)(y): bit i is bit (i + b lx) mod n of y; t_(a, 0): bit (r, c) is
    // bit (r, (c + a) mod lx)
    auto rows_by = [&](W y, int b) -> W {
      return b == 0 ? y : static_cast<W>((y >> (b * lx)) | ((y << (g.nbits - b * lx)) & flip));
    };
    auto columns_by = [&](W y, int a) -> W {
      return static_cast<W>(((y >> a) & static_cast<W>(rt->lo[a])) | ((y << (lx - a)) & static_cast<W>(rt->hi[a])));
    };
    for (int q = 0; q < nq; ++q) {
      W const y = group_apply<W, D>(g, g.fact_q[q], x);
#pragma unroll
      for (int s = 0; s < (SI != 0 ? 2 : 1); ++s) {
        W const ys = s == 1 ? static_cast<W>(y ^ flip) : y;
        // least rotation of every row; rows reaching the least one (cmask)
        // and rows in the rotation class of x's top row (smask)
        uint32_t mr[kRowTableBits];
        uint32_t m = 0xFFu, smask = 0, cmask = 0;
#pragma unroll
        for (int j = 0; j < kRowTableBits; ++j) {
          mr[j] = j < ly ? rt->minrot[static_cast<uint32_t>(ys >> (j * lx)) & rm] : 0xFFu;
          m = mr[j] < m ? mr[j] : m;
          smask |= (j < ly && mr[j] == xcls ? 1u : 0u) << j;
        }
#pragma unroll
        for (int j = 0; j < kRowTableBits; ++j) cmask |= (j < ly && mr[j] == m ? 1u : 0u) << j;
        if (m > (static_cast<uint32_t>(best >> top) & rm)) cmask = 0;
        for (; cmask != 0; cmask &= cmask - 1) {
          int const j = __ffs(cmask) - 1;
          uint32_t const vj = static_cast<uint32_t>(ys >> (j * lx)) & rm;
          int const b = j + 1 == ly ? 0 : j + 1;
          W const yb = rows_by(ys, b);
          for (uint32_t am = rt->amin[vj]; am != 0; am &= am - 1) {
            int const a = __ffs(am) - 1;
            W const w = columns_by(yb, a);
            int const k = (q * ly + b) * lx + a;
            int const code = 2 * k + s;
            if (w < best) {
              best = w;
              best_code = code;
            } else if (w == best && best_code >= 0 && w != x) {
              int const rank_new = g.fact_key[k] + s;
              int const rank_old = g.fact_key[best_code >> 1] + (best_code & 1);
              if (rank_new < rank_old) best_code = code;
            }
          }
        }
        for (; smask != 0; smask &= smask - 1) {
          int const j = __ffs(smask) - 1;
          uint32_t const vj = static_cast<uint32_t>(ys >> (j * lx)) & rm;
          int const b = j + 1 == ly ? 0 : j + 1;
          W const yb = rows_by(ys, b);
          for (int a = 0; a < lx; ++a) {
            if ((((vj >> a) | (vj << (lx - a))) & rm) != xt) continue;
            if (columns_by(yb, a) == x) {
              double const re = g.eig_re[g.fact_elem[(q * ly + b) * lx + a]];
              nn = nn + ((s == 1 && SI == -1) ? -re : re);
            }
          }
        }
      }
    }
  } else {
  for (int q = 0; q < nq; ++q) {
    W const y = group_apply<W, D>(g, g.fact_q[q], x);
#pragma unroll
    for (int s = 0; s < (SI != 0 ? 2 : 1); ++s) {
      W row = s == 1 ? static_cast<W>(y ^ flip) : y;
      for (int b = 0; b < ly; ++b) {
        W w = row;
        int const k0 = (q * ly + b) * lx;
        for (int a = 0; a < lx; ++a) {
          int const code = 2 * (k0 + a) + s;
          if (w < best) {
            best = w;
            best_code = code;
          } else if (w == best && best_code >= 0 && w != x) {
            // tie below x: the lower rank in the reference's reduction wins
            int const rank_new = g.fact_key[k0 + a] + s;
            int const rank_old = g.fact_key[best_code >> 1] + (best_code & 1);
            if (rank_new < rank_old) best_code = code;
          }
          if (w == x) {
            double const re = g.eig_re[g.fact_elem[k0 + a]];
            nn = nn + ((s == 1 && SI == -1) ? -re : re);
          }
          w = static_cast<W>(((w >> 1) & m1) | ((w << (lx - 1)) & m2));
        }
        if (ly > 1) row = static_cast<W>((row >> lx) | ((row << up) & flip));
      }
    }
  }
  }
  StateInfo out;
  out.rep = static_cast<uint64_t>(best);
  if (best_code < 0) {
    out.c_re = 1.0;
    out.c_im = 0.0;
  } else {
    int const e = g.fact_elem[best_code >> 1];
    double const cre = g.eig_re[e], cim = g.eig_im[e];
    bool const neg = SI == -1 && (best_code & 1);
    out.c_re = neg ? -cre : cre;
    out.c_im = neg ? -cim : cim;
  }
  int const G = g.size;
  double const denom = SI == 0 ? static_cast<double>(G) : static_cast<double>(2 * G);
  out.norm = (g.norm_eps > 0.0 && !(nn > g.norm_eps)) ? 0.0 : sqrt(nn / denom);
  return out;
}

// ---------------------------------------------------------------------------
// is_representative (generator.cpp:168-272)
// ---------------------------------------------------------------------------
//
// Chunks of four elements (natural_vector_size(uint64) on avx2): within a
// chunk the terms (y == x) * Re(chi) are summed from 0 in order and the chunk
// sum is added to the total; the flag is the AND of y >= x (and y^flip >= x).
// Chunks stop after the first one that clears the flag; the remaining
// |G| mod 4 elements are always folded in.  is_rep = norm > 0 ? flag : 0.

template <typename W, int D, int SI>
__device__ __forceinline__ void is_rep_term(GroupView const &g, int e, W x,
                                            double &t, bool &f) {
  W const y = group_apply<W, D>(g, e, x);
  double re = g.eig_re[e];
  if constexpr (SI == kFermionic) {
    if (y == x) re = re * fermion_sign(g, e, static_cast<uint64_t>(x));
    t = static_cast<double>(y == x) * re;
    f = y >= x;
  } else if constexpr (SI == 0) {
    t = static_cast<double>(y == x) * re;
    f = y >= x;
  } else {
    W const yf = static_cast<W>(y ^ static_cast<W>(g.flip));
    double const a = static_cast<double>(y == x);
    double const b = static_cast<double>(yf == x);
    t = (SI == 1 ? (a + b) : (a - b)) * re;
    f = (y >= x) && (yf >= x);
  }
}

template <typename W, int D, int SI>
__device__ __forceinline__ void is_representative_one(GroupView const &g,
                                                      uint64_t x64,
                                                      uint8_t &flag_out,
                                                      double &norm_out) {
  constexpr int L = kIsRepresentativeLanes;
  W const x = static_cast<W>(x64);
  int const G = g.size;
  int const nch = G / L;
  double total = 0.0;
  bool flag = true;
  for (int c = 0; c < nch; ++c) {
    if (!flag) break;
    double lane = 0.0;
    bool lf = true;
#pragma unroll
    for (int j = 0; j < L; ++j) {
      double t;
      bool f;
      is_rep_term<W, D, SI>(g, c * L + j, x, t, f);
      lane = lane + t;
      lf = lf && f;
    }
    total = total + lane;
    flag = flag && lf;
  }
  for (int e = nch * L; e < G; ++e) {
    double t;
    bool f;
    is_rep_term<W, D, SI>(g, e, x, t, f);
    total = total + t;
    flag = flag && f;
  }
  norm_out = total;
  flag_out = (total > g.norm_eps) ? static_cast<uint8_t>(flag) : static_cast<uint8_t>(0);
}

// ---------------------------------------------------------------------------
// State index lookup over the prefix table (indexing.c:273-325 semantics)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t index_lookup(IndexView const &idx, uint64_t x) {
  uint64_t const b = x >> idx.shift;
  if (b >= (uint64_t{1} << idx.prefix_bits)) return -1;
  int64_t lo = idx.offsets[b];
  int64_t n = static_cast<int64_t>(idx.offsets[b + 1]) - lo;
  // branch-free lower bound in [lo, lo + n)
  while (n > 0) {
    int64_t const half = n >> 1;
    uint64_t const v = idx.reps[lo + half];
    bool const less = v < x;
    lo = less ? lo + half + 1 : lo;
    n = less ? n - half - 1 : half;
  }
  uint64_t const hi_bucket_end = idx.offsets[b + 1];
  return (static_cast<uint64_t>(lo) < hi_bucket_end && idx.reps[lo] == x) ? lo : -1;
}

// ---------------------------------------------------------------------------
// Dispatch on (word type, Benes depth, spin inversion)
// ---------------------------------------------------------------------------
//
// f is a generic callable invoked as f.template operator()<W, D, SI>().
// Canonical networks of depth 7 and 9 (9..32 sites) run on 32-bit words,
// depth 11 (33..64 sites) on 64-bit words; anything else uses the runtime
// depth path.
template <typename F> bool dispatch_group(GroupData const &g, F &&f) {
  auto by_si = [&]<typename W, int D>() -> bool {
    if (g.fermion_species != 0) return f.template operator()<W, D, kFermionic>();
    switch (g.spin_inversion) {
    case 0: return f.template operator()<W, D, 0>();
    case 1: return f.template operator()<W, D, 1>();
    case -1: return f.template operator()<W, D, -1>();
    default: return fail("invalid spin_inversion");
    }
  };
  if (g.canonical_shifts) {
    switch (g.depth) {
    case 7: return by_si.template operator()<uint32_t, 7>();
    case 9: return by_si.template operator()<uint32_t, 9>();
    case 11: return by_si.template operator()<uint64_t, 11>();
    default: break;
    }
  }
  return by_si.template operator()<uint64_t, 0>();
}

inline int grid_for(int64_t n, int block) {
  int64_t const b = (n + block - 1) / block;
  return static_cast<int>(b < 1 ? 1 : b);
}

} // namespace lsamd
