// enumerate.hip -- representatives of a basis on the GPU.
//
// Restates chapel/src/StatesEnumeration.chpl (_enumerateStatesProjected
// :242-268, _enumerateStatesUnprojected :269-289, the spinful per-sector
// product :290-326) as three passes:
//
//  1. count:  every lane owns a contiguous range of candidates by
//     combinadics rank (StatesEnumeration.chpl:130-152 splits the same way),
//     unranks its first state and steps with Gosper's rule (:29-32).  The
//     lane tests one group element per loop iteration and moves on to its
//     next candidate as soon as the current one is rejected (an image below
//     x) or accepted (all elements seen), so no lane idles while its
//     neighbours finish long orbits.  Group masks and Re(chi) are staged in
//     LDS.  Accepted candidates set a bit in a per-lane bitmask.
//  2. scan:   exclusive prefix sum of the per-lane counts.
//  3. write:  every lane replays its range and writes accepted states at
//     its offset, which keeps the global order ascending.
//
// Acceptance is flag && norm > 0 with the norm summed in the chunk order of
// generator.cpp:224-253 (chunks of four, chunk sums added left to right,
// tail elements one by one), so sectors whose characters are not +-1 keep
// exactly the states the restated reference keeps.
#include <algorithm>
#include <atomic>

#include "orbit_device.hpp"

namespace lsamd {

constexpr int kAllIntegers = -1; // hamming: every integer in the range
constexpr int kSpinful = -2;     // hamming: product of two fixed-weight species

struct EnumParams {
  uint64_t rank0;   // rank of the first candidate (or the state itself)
  int64_t total;    // number of candidates
  int64_t per_lane; // candidates per lane (acceptance bits: ceil(per_lane / 32) words per lane)
  int64_t lanes;
  int hamming; // >= 0 fixed weight, kAllIntegers, kSpinful
  uint64_t const *binom;
  // kSpinful: x = (xB << sites) | xA with xA of weight up in [minA, maxA]
  // and xB of weight down; candidate r is (rB, rA) = divmod(r, countA).
  int sites, up, down;
  uint64_t rankA0, countA, rankB0, minA, maxA;
  // candidates from an explicit ascending list instead (list != nullptr)
  uint64_t const *list;
  // acceptance filter: group elements tried per candidate; survive_only
  // accepts a candidate that passes `limit` elements without the norm test
  int limit, survive_only;
};

__host__ __device__ constexpr int64_t words_per_lane(int64_t per_lane) { return (per_lane + 31) / 32; }

__device__ __forceinline__ uint64_t enum_first(EnumParams const &p, int64_t first);

__device__ __forceinline__ uint64_t unrank_weight(uint64_t const *binom, uint64_t r, int h) {
  uint64_t state = 0;
  int c = 63;
  for (int i = h; i > 0; --i) {
    while (c >= i && binom[c * 65 + i] > r) --c;
    state |= uint64_t{1} << c;
    r -= binom[c * 65 + i];
    --c;
  }
  return state;
}

__device__ __forceinline__ uint64_t gosper_next(uint64_t v) {
  uint64_t const t = v | (v - 1);
  return (t + 1) | (((~t & (t + 1)) - 1) >> (__builtin_ctzll(v) + 1));
}

__device__ __forceinline__ uint64_t enum_unrank(EnumParams const &p, uint64_t r) {
  if (p.hamming == kAllIntegers) return r;
  if (p.hamming == kSpinful) {
    uint64_t const xa = unrank_weight(p.binom, p.rankA0 + r % p.countA, p.up);
    uint64_t const xb = unrank_weight(p.binom, p.rankB0 + r / p.countA, p.down);
    return (xb << p.sites) | xa;
  }
  return unrank_weight(p.binom, r, p.hamming);
}

__device__ __forceinline__ uint64_t enum_first(EnumParams const &p, int64_t first) {
  if (p.list != nullptr) return p.list[first];
  return enum_unrank(p, p.rank0 + static_cast<uint64_t>(first));
}

// Candidate k of a lane's range given candidate k - 1 (v).
__device__ __forceinline__ uint64_t enum_step(EnumParams const &p, int64_t first, int64_t k, uint64_t v);

__device__ __forceinline__ uint64_t enum_next(EnumParams const &p, uint64_t v) {
  if (p.hamming == kAllIntegers) return v + 1;
  if (p.hamming == kSpinful) {
    uint64_t const amask = (uint64_t{1} << p.sites) - 1;
    uint64_t const xa = v & amask, xb = v >> p.sites;
    if (xa == p.maxA) return (gosper_next(xb) << p.sites) | p.minA;
    return (xb << p.sites) | gosper_next(xa);
  }
  return gosper_next(v);
}

// Stabiliser entries a lane keeps while it scans the group out of order; a
// candidate with more of them is re-scanned in the reference order.
constexpr int kMaxStabiliser = 4;

// Acceptance filter of the projected enumeration (is_representative with
// norm > 0, StatesEnumeration.chpl:242-268).  One lane walks its candidates;
// each loop iteration tests one group element of the lane's current
// candidate, so lanes refill independently.  Elements are tried in `order`
// (most rejecting first): the flag "g(x) >= x for every g" does not depend
// on the order.  The norm sum must follow the reference's reduction order
// (chunks of kIsRepresentativeLanes, then single elements): the lane records
// the stabiliser contributions (element, term) and, once every element has
// passed, replays them in element order -- the skipped terms are exact zeros,
// so the sum is the reference's bit for bit.  More than kMaxStabiliser
// contributions send the candidate through a second scan in the reference
// order.
__device__ __forceinline__ uint64_t enum_step(EnumParams const &p, int64_t first, int64_t k, uint64_t v) {
  if (p.list != nullptr) return p.list[first + k];
  return enum_next(p, v);
}

template <typename W, int D, int SI, bool LDS>
__global__ __launch_bounds__(256) void enum_count_kernel(EnumParams p, GroupView g, int const *__restrict__ order,
                                                         uint32_t *__restrict__ bits,
                                                         uint32_t *__restrict__ counts) {
  extern __shared__ uint64_t smem[];
  int const G = g.size;
  int const depth = D > 0 ? D : g.depth;
  W *s_masks = reinterpret_cast<W *>(smem);
  double *s_re = reinterpret_cast<double *>(
      smem + (static_cast<size_t>(G) * depth * sizeof(W) + 7) / 8);
  int *s_order = reinterpret_cast<int *>(s_re + G);
  if constexpr (LDS) {
    // masks and characters staged in scan order, so the scan indexes LDS
    // directly; s_order maps a scan position back to the element
    for (int i = threadIdx.x; i < G * depth; i += blockDim.x) {
      int const pos = i / depth, k = i - pos * depth;
      s_masks[i] = static_cast<W>(g.masks[static_cast<size_t>(order[pos]) * depth + k]);
    }
    for (int i = threadIdx.x; i < G; i += blockDim.x) {
      s_re[i] = g.eig_re[order[i]];
      s_order[i] = order[i];
    }
    __syncthreads();
  }
  int64_t const lane = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t const first = lane * p.per_lane;
  int64_t remaining = 0;
  if (lane < p.lanes && first < p.total) remaining = min(p.per_lane, p.total - first);
  W x = remaining > 0 ? static_cast<W>(enum_first(p, first)) : W{0};
  uint32_t *my_bits = bits + lane * words_per_lane(p.per_lane);
  W const flip = static_cast<W>(g.flip);
  int const nch4 = (G / kIsRepresentativeLanes) * kIsRepresentativeLanes;
  int const limit = p.limit;
  int e = 0;
  int64_t pos = 0;
  uint32_t word = 0, count = 0;
  double total = 0.0, chunk = 0.0;
  bool in_order = false; // second scan, reference order
  int nrec = 0;
  bool overflow = false;
  int rec_e[kMaxStabiliser];
  double rec_t[kMaxStabiliser];

  while (__any(remaining > 0)) {
    if (remaining > 0) {
      // element of this iteration: scan position e (LDS holds the scan
      // order), or element e itself in the in-order re-scan
      W y;
      double re;
      if (LDS && !in_order) {
        y = benes_apply_ptr<W, D>(s_masks + static_cast<size_t>(e) * depth, x, g.shifts, depth);
        re = s_re[e];
      } else {
        int const el = in_order ? e : order[e];
        y = group_apply<W, D>(g, el, x);
        re = g.eig_re[el];
      }
      bool ok = y >= x;
      bool stab = y == x;
      double t = static_cast<double>(y == x) * re;
      if constexpr (SI == kFermionic) {
        if (y == x) {
          int const el = in_order ? e : (LDS ? s_order[e] : order[e]);
          t = re * fermion_sign(g, el, static_cast<uint64_t>(x));
        }
      } else if constexpr (SI != 0) {
        W const yf = static_cast<W>(y ^ flip);
        ok = ok && (yf >= x);
        stab = stab || (yf == x);
        double const a = static_cast<double>(y == x), b = static_cast<double>(yf == x);
        t = (SI == 1 ? (a + b) : (a - b)) * re;
      }
      bool done = false, accept = false;
      if (!ok) {
        done = true;
      } else if (in_order) {
        if (e < nch4) {
          chunk = chunk + t;
          if ((e & 3) == 3) {
            total = total + chunk;
            chunk = 0.0;
          }
        } else {
          total = total + t;
        }
        ++e;
        if (e == G) {
          done = true;
          accept = total > g.norm_eps;
        }
      } else {
        if (stab) {
          if (nrec < kMaxStabiliser) {
            int const el = LDS ? s_order[e] : order[e];
#pragma unroll
            for (int r = 0; r < kMaxStabiliser; ++r)
              if (r == nrec) {
                rec_e[r] = el;
                rec_t[r] = t;
              }
            ++nrec;
          } else {
            overflow = true;
          }
        }
        ++e;
        if (p.survive_only && e == limit) {
          done = true; // passed the first `limit` elements: a survivor
          accept = true;
        } else if (e == G) {
          if (overflow) {
            in_order = true; // every element passed; recompute the norm in order
            e = 0;
            total = 0.0;
            chunk = 0.0;
          } else {
            // replay the recorded terms in element order with the
            // reference's chunking
#pragma unroll
            for (int a = 1; a < kMaxStabiliser; ++a)
#pragma unroll
              for (int b = kMaxStabiliser - 1; b >= a; --b)
                if (b < nrec && rec_e[b - 1] > rec_e[b]) {
                  int const te = rec_e[b - 1];
                  rec_e[b - 1] = rec_e[b];
                  rec_e[b] = te;
                  double const tt = rec_t[b - 1];
                  rec_t[b - 1] = rec_t[b];
                  rec_t[b] = tt;
                }
            total = 0.0;
            chunk = 0.0;
            int cur = -1;
#pragma unroll
            for (int r = 0; r < kMaxStabiliser; ++r) {
              if (r < nrec) {
                int const er = rec_e[r];
                if (er < nch4) {
                  int const c = er >> 2;
                  if (c != cur) {
                    total = total + chunk;
                    chunk = 0.0;
                    cur = c;
                  }
                  chunk = chunk + rec_t[r];
                } else {
                  total = total + chunk;
                  chunk = 0.0;
                  cur = -1;
                  total = total + rec_t[r];
                }
              }
            }
            total = total + chunk;
            done = true;
            accept = total > g.norm_eps;
          }
        }
      }
      if (done) {
        if (accept) {
          word |= 1u << (pos & 31);
          ++count;
        }
        ++pos;
        --remaining;
        if ((pos & 31) == 0 || remaining == 0) {
          my_bits[(pos - 1) >> 5] = word;
          word = 0;
        }
        if (remaining > 0) x = static_cast<W>(enum_step(p, first, pos, static_cast<uint64_t>(x)));
        e = 0;
        total = 0.0;
        chunk = 0.0;
        in_order = false;
        nrec = 0;
        overflow = false;
      }
    }
  }
  if (lane < p.lanes) counts[lane] = count;
}

// ---------------------------------------------------------------------------
// Factorised filter for groups G = {t o q} of torus translations t and point
// elements q with real characters +-1 (GroupView::fact_*): the flag "every
// image >= x" does not depend on the order of the elements, and the
// stabiliser sum of +-1 terms is exact in any order, so the accepted states
// are those of the filter above.
// ---------------------------------------------------------------------------

template <typename W> struct TorusRotation {
  W m1, m2, flip;
  int lx, up;
  // translation by one column: bit r lx + c of the result is bit
  // r lx + (c + 1) mod lx of w
  __device__ __forceinline__ W column(W w) const {
    return static_cast<W>(((w >> 1) & m1) | ((w << (lx - 1)) & m2));
  }
  // translation by one row: bit i of the result is bit (i + lx) mod n of w
  __device__ __forceinline__ W row(W w) const { return static_cast<W>((w >> lx) | ((w << up) & flip)); }
};

template <typename W> __device__ __forceinline__ TorusRotation<W> torus_rotation(GroupView const &g) {
  return TorusRotation<W>{static_cast<W>(g.fact_m1), static_cast<W>(g.fact_m2), static_cast<W>(g.flip), g.fact_lx,
                          g.nbits - g.fact_lx};
}

// Records the decision on a lane's current candidate in its acceptance bits.
struct LaneBits {
  uint32_t *bits;
  int64_t pos = 0;
  uint32_t word = 0, count = 0;
  __device__ __forceinline__ void decide(bool accept, int64_t &remaining) {
    if (accept) {
      word |= 1u << (pos & 31);
      ++count;
    }
    ++pos;
    --remaining;
    if ((pos & 31) == 0 || remaining == 0) {
      bits[(pos - 1) >> 5] = word;
      word = 0;
    }
  }
};

// Pass 0: a candidate survives when no translation maps it below itself
// (t(x) >= x, and with spin inversion t(x) ^ flip >= x, the identity's
// flipped image included).  Ranked candidates of a fixed Hamming weight,
// stepped by Gosper's rule.
//
// With row tables (RowTables, rows and columns of at most 8 sites) a
// candidate is decided in one step: an image below x needs a top row at
// most x's top row, and the top rows of the translations of x are the
// rotations of its rows, so x survives only if its top row is the least
// rotation of every row, and then only the translations that reach that
// row are formed and compared (the same for x ^ flip).  Otherwise two
// translations per loop iteration (column rotations along a row, then the
// next row), selected without branches; lanes move to their next candidate
// independently.
template <typename W, int SI>
__global__ __launch_bounds__(256) void enum_translation_kernel(EnumParams p, GroupView g, uint32_t *__restrict__ bits,
                                                               uint32_t *__restrict__ counts) {
  __shared__ RowTables rt;
  row_tables_init(g, rt);
  int64_t const lane = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t const first = lane * p.per_lane;
  int64_t remaining = 0;
  if (lane < p.lanes && first < p.total) remaining = min(p.per_lane, p.total - first);
  W x = remaining > 0 ? static_cast<W>(enum_first(p, first)) : W{0};
  LaneBits out{bits + lane * words_per_lane(p.per_lane)};
  TorusRotation<W> const rot = torus_rotation<W>(g);
  W const flip = static_cast<W>(g.flip);
  int const lx = g.fact_lx, ly = g.fact_ly;
  if (row_tables_apply(g)) {
    uint32_t const rm = (1u << lx) - 1;
    int const top = (ly - 1) * lx;
    // t_(0, b) and t_(a, 0) as in state_info_factored
    auto rows_by = [&](W y, int b) -> W {
      return b == 0 ? y : static_cast<W>((y >> (b * lx)) | ((y << (g.nbits - b * lx)) & flip));
    };
    auto columns_by = [&](W y, int a) -> W {
      return static_cast<W>(((y >> a) & static_cast<W>(rt.lo[a])) | ((y << (lx - a)) & static_cast<W>(rt.hi[a])));
    };
    // true when a translation of y is below x (xt: x's top row)
    auto below = [&](W y, uint32_t xt) -> bool {
      uint32_t mr[kRowTableBits];
      uint32_t m = 0xFFu, cmask = 0;
#pragma unroll
      for (int j = 0; j < kRowTableBits; ++j) {
        mr[j] = j < ly ? rt.minrot[static_cast<uint32_t>(y >> (j * lx)) & rm] : 0xFFu;
        m = mr[j] < m ? mr[j] : m;
      }
      if (m < xt) return true;
      if (m > xt) return false;
#pragma unroll
      for (int j = 0; j < kRowTableBits; ++j) cmask |= (j < ly && mr[j] == m ? 1u : 0u) << j;
      bool low = false;
      for (; cmask != 0 && !low; cmask &= cmask - 1) {
        int const j = __ffs(cmask) - 1;
        uint32_t const vj = static_cast<uint32_t>(y >> (j * lx)) & rm;
        W const yb = rows_by(y, j + 1 == ly ? 0 : j + 1);
        for (uint32_t am = rt.amin[vj]; am != 0; am &= am - 1) low = low || columns_by(yb, __ffs(am) - 1) < x;
      }
      return low;
    };
    while (remaining > 0) {
      uint32_t const xt = static_cast<uint32_t>(x >> top) & rm;
      bool const reject = below(x, xt) || (SI != 0 && below(static_cast<W>(x ^ flip), xt));
      out.decide(!reject, remaining);
      if (remaining > 0) x = static_cast<W>(gosper_next(static_cast<uint64_t>(x)));
    }
    if (lane < p.lanes) counts[lane] = out.count;
    return;
  }
  W row = x, w = x; // w = t_(a, b)(x), row = t_(0, b)(x)
  int a = 0, b = 0;
  while (__any(remaining > 0)) {
    if (remaining > 0) {
      bool reject = false;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        bool const live = b < ly;
        bool const low = (SI != 0 && static_cast<W>(w ^ flip) < x) || ((a | b) != 0 && w < x);
        reject = reject || (live && low);
        bool const along = a + 1 < lx;
        W const wc = rot.column(w), rr = rot.row(row);
        w = along ? wc : rr;
        row = along ? row : rr;
        a = along ? a + 1 : 0;
        b = along ? b : b + 1;
      }
      if (reject || b >= ly) {
        out.decide(!reject, remaining);
        if (remaining > 0) x = static_cast<W>(gosper_next(static_cast<uint64_t>(x)));
        row = w = x;
        a = b = 0;
      }
    }
  }
  if (lane < p.lanes) counts[lane] = out.count;
}

// Pass 1 over the survivors of pass 0 (listed candidates): one point element
// q per loop iteration, the identity last (pass 0 has tested its
// translations): q(x) by its Benes network, then every translation of q(x)
// and of its flipped image.  Rejected as soon as an image is below x;
// accepted after the last q when the stabiliser sum exceeds norm_eps.
template <typename W, int D, int SI>
__global__ __launch_bounds__(256) void enum_factored_kernel(EnumParams p, GroupView g, int q_identity,
                                                            uint32_t *__restrict__ bits,
                                                            uint32_t *__restrict__ counts) {
  __shared__ RowTables rt;
  row_tables_init(g, rt);
  bool const tables = row_tables_apply(g);
  int64_t const lane = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t const first = lane * p.per_lane;
  int64_t remaining = 0;
  if (lane < p.lanes && first < p.total) remaining = min(p.per_lane, p.total - first);
  W x = remaining > 0 ? static_cast<W>(enum_first(p, first)) : W{0};
  LaneBits out{bits + lane * words_per_lane(p.per_lane)};
  TorusRotation<W> const rot = torus_rotation<W>(g);
  W const flip = static_cast<W>(g.flip);
  int const lx = g.fact_lx, ly = g.fact_ly, nq = g.fact_nq;
  int i = 0;
  double nn = 0.0;
  while (__any(remaining > 0)) {
    if (remaining > 0) {
      int const q = (q_identity + 1 + i) % nq;
      W const y = group_apply<W, D>(g, g.fact_q[q], x);
      bool lower = false;
      if (tables) {
        // as in state_info_factored: only the translations whose top row can
        // be below x's, or equal to it, are formed
        uint32_t const rm = (1u << lx) - 1;
        int const top = (ly - 1) * lx;
        uint32_t const xt = static_cast<uint32_t>(x >> top) & rm;
        uint32_t const xcls = rt.minrot[xt];
#pragma unroll
        for (int s = 0; s < (SI != 0 ? 2 : 1); ++s) {
          W const ys = s == 1 ? static_cast<W>(y ^ flip) : y;
          uint32_t m = 0xFFu, smask = 0;
#pragma unroll
          for (int j = 0; j < kRowTableBits; ++j) {
            uint32_t const mr = j < ly ? rt.minrot[static_cast<uint32_t>(ys >> (j * lx)) & rm] : 0xFFu;
            m = mr < m ? mr : m;
            smask |= (j < ly && mr == xcls ? 1u : 0u) << j;
          }
          lower = lower || m < xt;
          // images with top row xt: compared with x, and counted when equal
          if (m <= xt) {
            for (; smask != 0; smask &= smask - 1) {
              int const j = __ffs(smask) - 1;
              uint32_t const vj = static_cast<uint32_t>(ys >> (j * lx)) & rm;
              int const b = j + 1 == ly ? 0 : j + 1;
              W const yb = b == 0 ? ys : static_cast<W>((ys >> (b * lx)) | ((ys << (g.nbits - b * lx)) & flip));
              for (int a = 0; a < lx; ++a) {
                if ((((vj >> a) | (vj << (lx - a))) & rm) != xt) continue;
                W const w = static_cast<W>(((yb >> a) & static_cast<W>(rt.lo[a])) |
                                           ((yb << (lx - a)) & static_cast<W>(rt.hi[a])));
                lower = lower || w < x;
                if (w == x) {
                  double const re = g.eig_re[g.fact_elem[(q * ly + b) * lx + a]];
                  nn = nn + ((s == 1 && SI == -1) ? -re : re);
                }
              }
            }
          }
        }
      }
#pragma unroll
      for (int s = 0; s < (tables ? 0 : (SI != 0 ? 2 : 1)); ++s) {
        W row = s == 1 ? static_cast<W>(y ^ flip) : y;
        for (int b = 0; b < ly; ++b) {
          W w = row;
          int const k0 = (q * ly + b) * lx;
          for (int a = 0; a < lx; ++a) {
            lower = lower || w < x;
            if (w == x) {
              double const re = g.eig_re[g.fact_elem[k0 + a]];
              nn = nn + ((s == 1 && SI == -1) ? -re : re);
            }
            w = rot.column(w);
          }
          row = rot.row(row);
        }
      }
      ++i;
      if (lower || i == nq) {
        out.decide(!lower && nn > g.norm_eps, remaining);
        if (remaining > 0) x = static_cast<W>(enum_step(p, first, out.pos, static_cast<uint64_t>(x)));
        i = 0;
        nn = 0.0;
      }
    }
  }
  if (lane < p.lanes) counts[lane] = out.count;
}

// Replays the candidate ranges and writes accepted states (all candidates
// when bits == nullptr) starting at offsets[lane] (lane * per_lane when
// offsets == nullptr).
__global__ __launch_bounds__(256) void enum_write_kernel(EnumParams p,
                                                         uint32_t const *__restrict__ bits,
                                                         uint64_t const *__restrict__ offsets,
                                                         uint64_t *__restrict__ out) {
  int64_t const lane = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (lane >= p.lanes) return;
  int64_t const first = lane * p.per_lane;
  if (first >= p.total) return;
  int64_t const n = min(p.per_lane, p.total - first);
  uint64_t x = enum_first(p, first);
  uint64_t o = offsets != nullptr ? offsets[lane] : static_cast<uint64_t>(first);
  uint32_t const *my_bits = bits != nullptr ? bits + lane * words_per_lane(p.per_lane) : nullptr;
  uint32_t word = 0xffffffffu;
  for (int64_t k = 0; k < n; ++k) {
    if (my_bits != nullptr && (k & 31) == 0) word = my_bits[k >> 5];
    if ((word >> (k & 31)) & 1u) out[o++] = x;
    if (k + 1 < n) x = enum_step(p, first, k + 1, x);
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of uint32 counts into uint64 offsets
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *s, uint64_t &total) {
  int const t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    uint64_t const add = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += add;
    __syncthreads();
  }
  total = s[kScanBlock - 1];
  uint64_t const incl = s[t];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(uint32_t const *__restrict__ in,
                                                                 int64_t n,
                                                                 uint64_t *__restrict__ sums) {
  __shared__ uint64_t s[kScanBlock];
  int64_t const base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  uint64_t acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t const i = base + static_cast<int64_t>(threadIdx.x) * kScanItems + k;
    if (i < n) acc += in[i];
  }
  uint64_t total;
  (void)block_exclusive_scan(acc, s, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of `n` tile sums in place, total at sums[n].
__global__ __launch_bounds__(kScanBlock) void scan_tiles_kernel(uint64_t *__restrict__ sums, int64_t n) {
  __shared__ uint64_t s[kScanBlock];
  uint64_t carry = 0;
  for (int64_t base = 0; base < n; base += kScanBlock) {
    int64_t const i = base + threadIdx.x;
    uint64_t const v = i < n ? sums[i] : 0;
    uint64_t total;
    uint64_t const ex = block_exclusive_scan(v, s, total);
    if (i < n) sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) sums[n] = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(uint32_t const *__restrict__ in,
                                                                int64_t n,
                                                                uint64_t const *__restrict__ sums,
                                                                uint64_t *__restrict__ out) {
  __shared__ uint64_t s[kScanBlock];
  int64_t const base = static_cast<int64_t>(blockIdx.x) * kScanTile +
                       static_cast<int64_t>(threadIdx.x) * kScanItems;
  uint64_t acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t const i = base + k;
    if (i < n) acc += in[i];
  }
  uint64_t total;
  uint64_t run = sums[blockIdx.x] + block_exclusive_scan(acc, s, total);
  for (int k = 0; k < kScanItems; ++k) {
    int64_t const i = base + k;
    if (i < n) {
      out[i] = run;
      run += in[i];
    }
  }
}

// offsets[0 .. n) exclusive sums of counts; returns the total.
bool exclusive_scan(uint32_t const *d_counts, int64_t n, uint64_t *d_offsets, uint64_t &total) {
  int64_t const tiles = (n + kScanTile - 1) / kScanTile;
  DeviceArray<uint64_t> sums;
  if (!sums.allocate(static_cast<size_t>(tiles + 1))) return false;
  scan_reduce_kernel<<<static_cast<unsigned>(tiles), kScanBlock, 0, stream()>>>(d_counts, n, sums.ptr);
  LSAMD_HIP(hipGetLastError());
  scan_tiles_kernel<<<1, kScanBlock, 0, stream()>>>(sums.ptr, tiles);
  LSAMD_HIP(hipGetLastError());
  scan_apply_kernel<<<static_cast<unsigned>(tiles), kScanBlock, 0, stream()>>>(d_counts, n, sums.ptr,
                                                                              d_offsets);
  LSAMD_HIP(hipGetLastError());
  LSAMD_HIP(hipMemcpyAsync(&total, sums.ptr + tiles, sizeof(uint64_t), hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  return true;
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------

namespace {
// Candidates per lane: enough lanes to fill the chip several times over
// (256 CUs x 32 waves x 64 lanes ~ 0.5 M), at least 64 per lane.
// Group elements a candidate must pass in the first filter pass before it
// goes to the full test.
constexpr int kSurvivorElements = 24;

// Candidates per lane for about 2^21 lanes.  Ranked candidates come at least
// 64 to a lane (a lane unranks its first candidate, then steps); listed
// survivors of the first pass each cost a full scan of the group, so they
// may go one to a lane.
int64_t lanes_for(int64_t total, int64_t &per_lane, bool listed = false) {
  int64_t const target_lanes = int64_t{1} << 21;
  per_lane = (total + target_lanes - 1) / target_lanes;
  per_lane = listed ? std::max<int64_t>(1, per_lane) : std::max<int64_t>(64, ((per_lane + 31) / 32) * 32);
  return (total + per_lane - 1) / per_lane;
}
} // namespace

std::atomic<bool> g_factored_enumeration{true};

namespace {
// Scratch of the filter passes, kept between builds (grow-only) so that a
// build makes no hipMalloc / hipFree calls of its own once the buffers are
// large enough: a 6x6 build needs about 0.6 GB of candidate bits and 1 GB
// of first-pass survivors.  Released by release_enumeration_scratch
// (ls_chpl_finalize).  A build that finds them in use by another thread
// allocates its own.
struct EnumScratch {
  std::mutex busy;
  DeviceArray<uint32_t> bits, counts;
  DeviceArray<uint64_t> offsets, survivors;
};
EnumScratch g_enum_scratch;

template <typename T> bool at_least(DeviceArray<T> &a, size_t n) { return a.count >= n || a.allocate(n); }
} // namespace

void release_enumeration_scratch() {
  std::lock_guard<std::mutex> lock(g_enum_scratch.busy);
  g_enum_scratch.bits.reset();
  g_enum_scratch.counts.reset();
  g_enum_scratch.offsets.reset();
  g_enum_scratch.survivors.reset();
}

bool enumerate_states(EnumerationSpec const &spec, uint64_t lower, uint64_t upper,
                      DeviceArray<uint64_t> &out, int64_t &count, int64_t &candidates) {
  count = 0;
  candidates = 0;
  if (!ensure_device()) return false;
  uint64_t const *binom = binomials_device();
  if (binom == nullptr) return false;
  int const n = spec.number_bits;
  uint64_t const mask = n >= 64 ? ~uint64_t{0} : ((uint64_t{1} << n) - 1);

  bool const projected = spec.group != nullptr && spec.group->size > 0;
  EnumParams p{};
  p.binom = binom;
  uint64_t hi = upper;
  if (spec.spinful_per_sector) {
    // (StatesEnumeration.chpl:290-326): x = (xB << n) | xA, xB outer
    int const ns = spec.number_sites;
    uint64_t const smask = (uint64_t{1} << ns) - 1;
    uint64_t const minA = lower & smask, maxA = upper & smask;
    uint64_t const minB = (lower >> ns) & smask, maxB = (upper >> ns) & smask;
    if (__builtin_popcountll(minA) != spec.number_up || __builtin_popcountll(maxA) != spec.number_up ||
        __builtin_popcountll(minB) != spec.number_down || __builtin_popcountll(maxB) != spec.number_down)
      return fail("invalid lower/upper for a spinful fermion basis");
    p.hamming = kSpinful;
    p.sites = ns;
    p.up = spec.number_up;
    p.down = spec.number_down;
    p.minA = minA;
    p.maxA = maxA;
    p.rankA0 = fixed_hamming_state_to_index(minA);
    p.countA = fixed_hamming_state_to_index(maxA) - p.rankA0 + 1;
    p.rankB0 = fixed_hamming_state_to_index(minB);
    uint64_t const countB = fixed_hamming_state_to_index(maxB) - p.rankB0 + 1;
    p.rank0 = 0;
    p.total = static_cast<int64_t>(p.countA * countB);
  } else {
    if (!projected && spec.spin_inversion != 0) hi = std::min(hi, hi ^ mask); // :276-281
    if (lower > hi) {
      out.reset();
      return true;
    }
    p.hamming = spec.hamming_weight;
  }
  if (p.hamming == kSpinful) {
    // set above
  } else if (p.hamming >= 0) {
    if (__builtin_popcountll(lower) != p.hamming || __builtin_popcountll(hi) != p.hamming)
      return fail("lower and upper must have the basis Hamming weight");
    p.rank0 = fixed_hamming_state_to_index(lower);
    p.total = static_cast<int64_t>(fixed_hamming_state_to_index(hi) - p.rank0 + 1);
  } else {
    p.rank0 = lower;
    p.total = static_cast<int64_t>(hi - lower + 1);
  }
  candidates = p.total;
  p.lanes = lanes_for(p.total, p.per_lane);
  int const blocks = grid_for(p.lanes, 256);

  if (!projected) {
    if (!out.allocate(static_cast<size_t>(p.total))) return false;
    timer_begin(kTimerEnumerate);
    enum_write_kernel<<<blocks, 256, 0, stream()>>>(p, nullptr, nullptr, out.ptr);
    LSAMD_HIP(hipGetLastError());
    timer_end(kTimerEnumerate);
    LSAMD_HIP(hipStreamSynchronize(stream()));
    count = p.total;
    return true;
  }

  GroupData const &g = *spec.group;
  GroupView const v = g.view();
  // The factorised filter: fixed Hamming weight, a group of torus
  // translations times point elements with characters +-1, not fermionic.
  int q_identity = -1;
  if (g.factored() && p.hamming >= 0 && g_factored_enumeration.load())
    for (size_t k = 0; k < g.h_fact_q.size(); ++k) {
      bool id = true;
      for (int d = 0; d < g.depth && id; ++d) id = g.h_masks[static_cast<size_t>(g.h_fact_q[k]) * g.depth + d] == 0;
      if (id) q_identity = static_cast<int>(k);
    }
  bool const use_factored = q_identity >= 0;
  // The element order of the element-by-element filter (kind 0), formed on
  // first use: the factorised filter does not read it.
  DeviceArray<int> d_order;
  auto order = [&]() -> int const * {
    if (d_order.ptr == nullptr) {
      std::vector<int> const h_order =
          group_rejection_order(g, p.hamming >= 0 ? p.hamming : (p.hamming == kSpinful ? p.up + p.down : -1),
                                g.number_bits);
      if (!d_order.upload(h_order.data(), h_order.size())) return nullptr;
    }
    return d_order.ptr;
  };
  std::unique_lock<std::mutex> scratch_lock(g_enum_scratch.busy, std::try_to_lock);
  EnumScratch own;
  EnumScratch &sc = scratch_lock.owns_lock() ? g_enum_scratch : own;
  // One filter pass over candidates q (ranked or listed): bits, counts,
  // scan, then the accepted ones written to `dst` in order.  kind 0: the
  // element-by-element filter; 1: translations only (pass 0 of the
  // factorised filter); 2: the factorised test (its pass 1).
  // `dst` receives exactly `kept` states (the scratch survivor list may be
  // longer than its contents).
  auto filter = [&](EnumParams const &q, DeviceArray<uint64_t> &dst, uint64_t &kept, int kind = 0,
                    bool dst_is_scratch = false) -> bool {
    int const nblocks = grid_for(q.lanes, 256);
    DeviceArray<uint32_t> &bits = sc.bits, &counts = sc.counts;
    DeviceArray<uint64_t> &offsets = sc.offsets;
    if (!at_least(bits, static_cast<size_t>(q.lanes * words_per_lane(q.per_lane)))) return false;
    if (!at_least(counts, static_cast<size_t>(q.lanes))) return false;
    if (!at_least(offsets, static_cast<size_t>(q.lanes))) return false;
    int const *d_ord = kind == 0 ? order() : nullptr;
    if (kind == 0 && d_ord == nullptr) return false;
    bool ok = dispatch_group(g, [&]<typename W, int D, int SI>() -> bool {
      if constexpr (SI != kFermionic) {
        if (kind == 1) {
          enum_translation_kernel<W, SI><<<nblocks, 256, 0, stream()>>>(q, v, bits.ptr, counts.ptr);
          LSAMD_HIP(hipGetLastError());
          return true;
        }
        if (kind == 2) {
          enum_factored_kernel<W, D, SI><<<nblocks, 256, 0, stream()>>>(q, v, q_identity, bits.ptr, counts.ptr);
          LSAMD_HIP(hipGetLastError());
          return true;
        }
      }
      int const depth = D > 0 ? D : g.depth;
      size_t const smem = ((static_cast<size_t>(g.size) * depth * sizeof(W) + 7) / 8) * 8 +
                          static_cast<size_t>(g.size) * (sizeof(double) + sizeof(int));
      if (smem <= 64 * 1024) {
        enum_count_kernel<W, D, SI, true><<<nblocks, 256, smem, stream()>>>(q, v, d_ord, bits.ptr, counts.ptr);
      } else {
        enum_count_kernel<W, D, SI, false><<<nblocks, 256, 0, stream()>>>(q, v, d_ord, bits.ptr, counts.ptr);
      }
      LSAMD_HIP(hipGetLastError());
      return true;
    });
    if (!ok) return false;
    kept = 0;
    if (!exclusive_scan(counts.ptr, q.lanes, offsets.ptr, kept)) return false;
    if (!(dst_is_scratch ? at_least(dst, static_cast<size_t>(kept)) : dst.allocate(static_cast<size_t>(kept))))
      return false;
    if (kept > 0) {
      enum_write_kernel<<<nblocks, 256, 0, stream()>>>(q, bits.ptr, offsets.ptr, dst.ptr);
      LSAMD_HIP(hipGetLastError());
    }
    return true;
  };
  timer_begin(kTimerEnumerate);
  uint64_t total = 0;
  if (use_factored) {
    // Pass 0: translations only, over every candidate.  Pass 1: the
    // factorised test of the survivors (about one candidate in lx ly).
    uint64_t nsurv = 0;
    if (!filter(p, sc.survivors, nsurv, 1, true)) return false;
    EnumParams q = p;
    q.list = sc.survivors.ptr;
    q.rank0 = 0;
    q.total = static_cast<int64_t>(nsurv);
    q.lanes = lanes_for(q.total, q.per_lane, true);
    if (nsurv > 0) {
      if (!filter(q, out, total, 2)) return false;
    } else {
      out.reset();
    }
  } else if (g.size <= kSurvivorElements) {
    p.limit = g.size;
    p.survive_only = 0;
    if (!filter(p, out, total)) return false;
  } else {
    // Pass 1: every candidate against the first elements of the scan order
    // (nearly all are rejected by the first one or two).  Pass 2: the
    // survivors, now a short ascending list, against the whole group, spread
    // evenly over the lanes.
    uint64_t nsurv = 0;
    p.limit = kSurvivorElements;
    p.survive_only = 1;
    if (!filter(p, sc.survivors, nsurv, 0, true)) return false;
    EnumParams q = p;
    q.list = sc.survivors.ptr;
    q.rank0 = 0;
    q.total = static_cast<int64_t>(nsurv);
    q.limit = g.size;
    q.survive_only = 0;
    q.lanes = lanes_for(q.total, q.per_lane, true);
    if (nsurv > 0) {
      if (!filter(q, out, total)) return false;
    } else {
      out.reset();
    }
  }
  timer_end(kTimerEnumerate);
  LSAMD_HIP(hipStreamSynchronize(stream()));
  count = static_cast<int64_t>(total);
  return true;
}

} // namespace lsamd
