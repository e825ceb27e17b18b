/*
 * oracle.c -- CPU restatement of the lattice-symmetries hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Every function names the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).  How each function is
 * pinned (golden vectors from the reference's own C sources, brute force,
 * closed forms, known answers) is listed in the header of oracle/oracle.py.
 *
 * Conventions shared with the reference:
 *   - bit i of g(x) is bit p[i] of x                  (haskell/src/LatticeSymmetries/Benes.hs:338-345)
 *   - (pa <> pb)[i] = pb[pa[i]], phases add mod 1      (Benes.hs:100-101, Group.hs:90-96)
 *   - a non-branching term (v,m,l,r,x,s) maps |a> to
 *     v (-1)^popcount(a&s) |a^x> when (a&m)==r          (kernels/reference.c:67-134)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>
#include <omp.h>

#define ORA_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Permutations and group closure                                            */
/* ------------------------------------------------------------------------ */

/* Benes.hs:338-345 (permuteBits'): bit i of the result is bit p[i] of x. */
ORA_EXPORT uint64_t ora_permute_bits(int n, int32_t const *p, uint64_t x) {
  uint64_t y = 0;
  for (int i = 0; i < n; ++i)
    y |= ((x >> p[i]) & 1ULL) << i;
  return y;
}

/* Benes.hs:325-336 (bitPermuteStep / permuteBits): apply a Benes network
   given as `depth` (mask, shift) pairs; masks are read with a stride so the
   [depth][|G|] layout of ls_hs_permutation_group can be used in place. */
ORA_EXPORT uint64_t ora_benes_apply(int depth, uint64_t const *masks,
                                    ptrdiff_t mask_stride,
                                    uint64_t const *shifts, uint64_t x) {
  for (int k = 0; k < depth; ++k) {
    uint64_t const m = masks[k * mask_stride];
    unsigned const d = (unsigned)shifts[k];
    uint64_t const y = ((x >> d) ^ x) & m;
    x = (x ^ y) ^ (y << d);
  }
  return x;
}

static int64_t gcd64(int64_t a, int64_t b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b != 0) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

typedef struct {
  int32_t *perm; /* n entries */
  int64_t num, den;
} ora_sym;

static int sym_cmp(int n, int32_t const *pa, int64_t na, int64_t da,
                   int32_t const *pb, int64_t nb, int64_t db) {
  for (int i = 0; i < n; ++i) {
    if (pa[i] != pb[i]) return pa[i] < pb[i] ? -1 : 1;
  }
  /* Ratio ordering: na/da vs nb/db with positive denominators. */
  __int128 l = (__int128)na * db, r = (__int128)nb * da;
  if (l != r) return l < r ? -1 : 1;
  return 0;
}

/* Group.hs:47-53 (getPeriodicity). */
static int periodicity(int n, int32_t const *p) {
  int32_t *cur = malloc(sizeof(int32_t) * (size_t)n);
  int32_t *tmp = malloc(sizeof(int32_t) * (size_t)n);
  memcpy(cur, p, sizeof(int32_t) * (size_t)n);
  int k = 1;
  for (;;) {
    int is_id = 1;
    for (int i = 0; i < n; ++i)
      if (cur[i] != i) { is_id = 0; break; }
    if (is_id) break;
    /* go (n+1) (p0 <> p): (p0 <> p)[i] = p[p0[i]] */
    for (int i = 0; i < n; ++i) tmp[i] = cur[p[i]];
    memcpy(cur, tmp, sizeof(int32_t) * (size_t)n);
    ++k;
    if (k > 1000000) { k = -1; break; }
  }
  free(cur);
  free(tmp);
  return k;
}

/* Sorted set of symmetries with binary-search insert. */
typedef struct {
  int n, size, cap;
  int32_t *perms;
  int64_t *num, *den;
} ora_symset;

static int symset_find(ora_symset const *s, int32_t const *p, int64_t num,
                       int64_t den, int *pos) {
  int lo = 0, hi = s->size;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    int c = sym_cmp(s->n, s->perms + (size_t)mid * s->n, s->num[mid],
                    s->den[mid], p, num, den);
    if (c < 0) lo = mid + 1;
    else hi = mid;
  }
  *pos = lo;
  return lo < s->size &&
         sym_cmp(s->n, s->perms + (size_t)lo * s->n, s->num[lo], s->den[lo], p,
                 num, den) == 0;
}

static int symset_insert(ora_symset *s, int32_t const *p, int64_t num,
                         int64_t den) {
  int pos;
  if (symset_find(s, p, num, den, &pos)) return 0;
  if (s->size == s->cap) return -1;
  memmove(s->perms + (size_t)(pos + 1) * s->n, s->perms + (size_t)pos * s->n,
          sizeof(int32_t) * (size_t)(s->size - pos) * s->n);
  memmove(s->num + pos + 1, s->num + pos, sizeof(int64_t) * (size_t)(s->size - pos));
  memmove(s->den + pos + 1, s->den + pos, sizeof(int64_t) * (size_t)(s->size - pos));
  memcpy(s->perms + (size_t)pos * s->n, p, sizeof(int32_t) * (size_t)s->n);
  s->num[pos] = num;
  s->den[pos] = den;
  s->size++;
  return 1;
}

/* Group.hs:90-96: phases add modulo one. */
static void phase_add(int64_t na, int64_t da, int64_t nb, int64_t db,
                      int64_t *nc, int64_t *dc) {
  int64_t num = na * db + nb * da, den = da * db;
  int64_t g = gcd64(num, den);
  if (g == 0) g = 1;
  num /= g;
  den /= g;
  if (num >= den) num -= (num / den) * den;
  *nc = num;
  *dc = den;
}

/*
 * Group.hs:106-132 (mkSymmetries) + Group.hs:174-183 (fromGenerators):
 * closure of the generators as an ascending set of (permutation, phase),
 * followed by the consistency check.  Returns |G| >= 1, or
 *   -1 inconsistent symmetries, -2 capacity exceeded, -3 invalid sector,
 *   -4 generators of different lengths / invalid permutation.
 * Phases are returned as reduced fractions num/den.
 */
ORA_EXPORT int ora_group_closure(int n, int ngens, int32_t const *gens,
                                 int32_t const *sectors, int cap,
                                 int32_t *out_perms, int64_t *out_num,
                                 int64_t *out_den) {
  if (ngens <= 0 || n <= 0) return 0;
  /* validate permutations */
  for (int k = 0; k < ngens; ++k) {
    char *seen = calloc((size_t)n, 1);
    for (int i = 0; i < n; ++i) {
      int v = gens[(size_t)k * n + i];
      if (v < 0 || v >= n || seen[v]) { free(seen); return -4; }
      seen[v] = 1;
    }
    free(seen);
  }
  int64_t *gnum = malloc(sizeof(int64_t) * (size_t)ngens);
  int64_t *gden = malloc(sizeof(int64_t) * (size_t)ngens);
  for (int k = 0; k < ngens; ++k) {
    int per = periodicity(n, gens + (size_t)k * n);
    /* Group.hs:78-85 (mkSymmetry) */
    if (sectors[k] < 0 || sectors[k] >= per) { free(gnum); free(gden); return -3; }
    int64_t g = gcd64(sectors[k], per);
    gnum[k] = sectors[k] / (g ? g : 1);
    gden[k] = per / (g ? g : 1);
    if (gnum[k] == 0) gden[k] = 1;
  }
  ora_symset interior = {n, 0, cap, out_perms, out_num, out_den};
  /* boundary sets are kept as plain sorted sets too */
  ora_symset boundary = {n, 0, cap, malloc(sizeof(int32_t) * (size_t)cap * n),
                         malloc(sizeof(int64_t) * (size_t)cap),
                         malloc(sizeof(int64_t) * (size_t)cap)};
  ora_symset next = {n, 0, cap, malloc(sizeof(int32_t) * (size_t)cap * n),
                     malloc(sizeof(int64_t) * (size_t)cap),
                     malloc(sizeof(int64_t) * (size_t)cap)};
  int32_t *tmp = malloc(sizeof(int32_t) * (size_t)n);
  int status = 0;
  for (int i = 0; i < n; ++i) tmp[i] = i;
  symset_insert(&boundary, tmp, 0, 1);
  while (boundary.size > 0 && status == 0) {
    for (int b = 0; b < boundary.size; ++b)
      if (symset_insert(&interior, boundary.perms + (size_t)b * n, boundary.num[b],
                        boundary.den[b]) < 0) { status = -2; break; }
    if (status) break;
    next.size = 0;
    for (int b = 0; b < boundary.size && status == 0; ++b) {
      int32_t const *h = boundary.perms + (size_t)b * n;
      for (int k = 0; k < ngens; ++k) {
        int32_t const *g = gens + (size_t)k * n;
        /* h <> g: (h <> g)[i] = g[h[i]] */
        for (int i = 0; i < n; ++i) tmp[i] = g[h[i]];
        int64_t nn, dd;
        phase_add(boundary.num[b], boundary.den[b], gnum[k], gden[k], &nn, &dd);
        int pos;
        if (symset_find(&interior, tmp, nn, dd, &pos)) continue;
        if (symset_insert(&next, tmp, nn, dd) < 0) { status = -2; break; }
      }
    }
    ora_symset t = boundary;
    boundary = next;
    next = t;
  }
  int size = interior.size;
  if (status == 0) {
    /* Group.hs:124-131: closure must be consistent. */
    for (int a = 0; a < size && status == 0; ++a)
      for (int b = 0; b < size; ++b) {
        int32_t const *pa = out_perms + (size_t)a * n;
        int32_t const *pb = out_perms + (size_t)b * n;
        for (int i = 0; i < n; ++i) tmp[i] = pb[pa[i]];
        int64_t nn, dd;
        phase_add(out_num[a], out_den[a], out_num[b], out_den[b], &nn, &dd);
        int pos;
        if (!symset_find(&interior, tmp, nn, dd, &pos)) { status = -1; break; }
        int per = periodicity(n, tmp);
        if (((int64_t)per * nn) % dd != 0) { status = -1; break; }
      }
  }
  free(boundary.perms); free(boundary.num); free(boundary.den);
  free(next.perms); free(next.num); free(next.den);
  free(tmp); free(gnum); free(gden);
  return status ? status : size;
}

/* Group.hs:115-116: characters cos(-2 pi phi) and sin(-2 pi phi). */
ORA_EXPORT void ora_characters(int size, int64_t const *num,
                               int64_t const *den, double *re, double *im) {
  for (int i = 0; i < size; ++i) {
    double const phi = (double)num[i] / (double)den[i];
    double const arg = -(2.0 * M_PI * phi);
    re[i] = cos(arg);
    im[i] = sin(arg);
  }
}

/* ------------------------------------------------------------------------ */
/* Group handle: images of a state under every group element                 */
/* ------------------------------------------------------------------------ */

typedef struct ora_group {
  int number_bits;   /* n */
  int size;          /* |G| (permutations only) */
  int spin_inversion;
  uint64_t flip_mask;
  double *eig_re, *eig_im;
  /* Direct mode: byte tables tab[g][byte][value]. */
  uint64_t *tab;
  /* Benes mode: masks[depth][size], shifts[depth]. */
  int depth;
  uint64_t *masks, *shifts;
  int use_benes;
  /* Fermionic signs: 0 = spins; 1 = spinless fermions on number_bits sites;
     2 = spinful fermions, sites 0..n/2-1 up and n/2..n-1 down. */
  int fermion_species;
  int32_t *perms; /* size x number_bits, kept for the signs */
  /* Opt-in: stabiliser sums <= norm_eps count as zero (0: reference rule). */
  double norm_eps;
} ora_group;

/* kernels.c:96-98 */
static uint64_t flip_mask_of(int n) {
  return n == 0 ? 0 : (~(uint64_t)0) >> (64 - n);
}

ORA_EXPORT ora_group *ora_group_create(int n, int size, int32_t const *perms,
                                       double const *eig_re,
                                       double const *eig_im, int spin_inversion,
                                       int depth, uint64_t const *masks,
                                       uint64_t const *shifts) {
  ora_group *g = calloc(1, sizeof(ora_group));
  g->number_bits = n;
  g->size = size;
  g->spin_inversion = spin_inversion;
  g->flip_mask = flip_mask_of(n);
  g->eig_re = malloc(sizeof(double) * (size_t)(size > 0 ? size : 1));
  g->eig_im = malloc(sizeof(double) * (size_t)(size > 0 ? size : 1));
  if (size > 0) {
    memcpy(g->eig_re, eig_re, sizeof(double) * (size_t)size);
    memcpy(g->eig_im, eig_im, sizeof(double) * (size_t)size);
  }
  if (perms != NULL && size > 0) {
    g->perms = malloc(sizeof(int32_t) * (size_t)size * n);
    memcpy(g->perms, perms, sizeof(int32_t) * (size_t)size * n);
    g->tab = calloc((size_t)size * 8 * 256, sizeof(uint64_t));
    for (int e = 0; e < size; ++e) {
      int32_t const *p = perms + (size_t)e * n;
      uint64_t *t = g->tab + (size_t)e * 8 * 256;
      /* bit i of g(x) = bit p[i] of x: source bit p[i] lands at i */
      for (int i = 0; i < n; ++i) {
        int const src = p[i];
        int const byte = src / 8, bit = src % 8;
        for (int v = 0; v < 256; ++v)
          if ((v >> bit) & 1) t[byte * 256 + v] |= 1ULL << i;
      }
    }
  }
  if (masks != NULL && depth > 0 && size > 0) {
    g->depth = depth;
    g->masks = malloc(sizeof(uint64_t) * (size_t)depth * size);
    g->shifts = malloc(sizeof(uint64_t) * (size_t)depth);
    memcpy(g->masks, masks, sizeof(uint64_t) * (size_t)depth * size);
    memcpy(g->shifts, shifts, sizeof(uint64_t) * (size_t)depth);
  }
  g->use_benes = (g->tab == NULL);
  return g;
}

/* Select which image evaluator to use: 0 = direct permutation tables (an
   evaluator independent of any Benes network), 1 = Benes masks as the
   reference's Halide kernels do (generator.cpp:5-8, 92-94). */
ORA_EXPORT int ora_group_set_mode(ora_group *g, int use_benes) {
  if (use_benes && g->masks == NULL) return -1;
  if (!use_benes && g->tab == NULL) return -1;
  g->use_benes = use_benes;
  return 0;
}

ORA_EXPORT void ora_group_destroy(ora_group *g) {
  if (g == NULL) return;
  free(g->eig_re);
  free(g->eig_im);
  free(g->tab);
  free(g->masks);
  free(g->shifts);
  free(g->perms);
  free(g);
}

static inline uint64_t image_of(ora_group const *g, int e, uint64_t x) {
  if (g->use_benes) {
    return ora_benes_apply(g->depth, g->masks + e, g->size, g->shifts, x);
  }
  uint64_t const *t = g->tab + (size_t)e * 8 * 256;
  uint64_t y = 0;
  for (int b = 0; b < 8; ++b) y |= t[b * 256 + ((x >> (8 * b)) & 0xFF)];
  return y;
}

ORA_EXPORT void ora_group_images(ora_group const *g, uint64_t x, uint64_t *ys) {
  for (int e = 0; e < g->size; ++e) ys[e] = image_of(g, e, x);
}

/* ------------------------------------------------------------------------ */
/* is_representative: generator.cpp:168-272                                  */
/* ------------------------------------------------------------------------ */

/* The per-element contribution `temp` (generator.cpp:201-222). */
/*
 * Fermionic sign of group element e on state x: with the state written as
 * the ordered product of creation operators (bit order, the Jordan-Wigner
 * order of the sign masks s = bit(i) - 1, Generator.hs:240-247), the element
 * relabels orbital b as the position i with p[i] = b, and the sign is the
 * parity of the inversions among the relabelled occupied orbitals.  The
 * reference has no fermionic symmetries (Basis.hs:941-948); this is the
 * convention of the build, pinned by brute force in the tests.
 */
static int fermion_sign(ora_group const *g, int e, uint64_t x) {
  if (!g->fermion_species) return 1;
  int const n = g->number_bits;
  int32_t const *p = g->perms + (size_t)e * n;
  int inv_of[64];
  for (int i = 0; i < n; ++i) inv_of[p[i]] = i;
  int img[64], k = 0;
  for (int b = 0; b < n; ++b)
    if ((x >> b) & 1) img[k++] = inv_of[b];
  int parity = 0;
  for (int a = 0; a < k; ++a)
    for (int c = a + 1; c < k; ++c) parity ^= img[a] > img[c];
  return parity ? -1 : 1;
}

ORA_EXPORT int ora_group_set_fermionic(ora_group *g, int species) {
  if (g->perms == NULL || g->spin_inversion != 0) return -1;
  g->fermion_species = species;
  return 0;
}

ORA_EXPORT void ora_group_set_norm_eps(ora_group *g, double eps) { g->norm_eps = eps; }

ORA_EXPORT int ora_fermion_sign(ora_group const *g, int e, uint64_t x) {
  return fermion_sign(g, e, x);
}

static inline void is_rep_temp(ora_group const *g, int e, uint64_t x,
                               double *t_n, int *t_f) {
  uint64_t const y = image_of(g, e, x);
  double re = g->eig_re[e];
  if (g->fermion_species && y == x) re = re * (double)fermion_sign(g, e, x);
  if (g->spin_inversion == 0) {
    *t_n = (double)(y == x) * re;
    *t_f = y >= x;
  } else {
    uint64_t const yf = y ^ g->flip_mask;
    double const a = (double)(y == x), b = (double)(yf == x);
    *t_n = (g->spin_inversion == 1 ? (a + b) : (a - b)) * re;
    *t_f = (y >= x) && (yf >= x);
  }
}

/*
 * Reduction order of generator.cpp:224-253 with chunk = `lanes`
 * (natural_vector_size(uint64) of the Halide target; 4 on avx2):
 *   per chunk c:  lane = ((0 + t[cL]) + t[cL+1]) + ... ;  total += lane
 *   chunks stop once the flag is false (m_batch.where, :236), then the
 *   tail elements are added one by one, and is_rep = norm > 0 ? flag : 0.
 */
static void is_rep_one(ora_group const *g, int lanes, uint64_t x,
                       uint8_t *flag_out, double *norm_out) {
  int const G = g->size;
  int const nchunks = G / lanes;
  double total = 0.0;
  int flag = 1;
  for (int c = 0; c < nchunks; ++c) {
    if (!flag) break;
    double lane_n = 0.0;
    int lane_f = 1;
    for (int j = 0; j < lanes; ++j) {
      double tn;
      int tf;
      is_rep_temp(g, c * lanes + j, x, &tn, &tf);
      lane_n = lane_n + tn;
      lane_f = lane_f && tf;
    }
    total = total + lane_n;
    flag = flag && lane_f;
  }
  for (int e = nchunks * lanes; e < G; ++e) {
    double tn;
    int tf;
    is_rep_temp(g, e, x, &tn, &tf);
    total = total + tn;
    flag = flag && tf;
  }
  *norm_out = total;
  *flag_out = (total > g->norm_eps) ? (uint8_t)flag : 0;
}

ORA_EXPORT void ora_is_representative(ora_group const *g, int lanes,
                                      ptrdiff_t batch, uint64_t const *xs,
                                      ptrdiff_t stride, uint8_t *flags,
                                      double *norms) {
  for (ptrdiff_t i = 0; i < batch; ++i)
    is_rep_one(g, lanes, xs[i * stride], flags + i, norms + i);
}

/* ------------------------------------------------------------------------ */
/* state_info: generator.cpp:10-166                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint64_t r;
  double c_re, c_im, n;
} si_acc;

/* generator.cpp:24-40 */
static inline void si_step_impl(si_acc *a, uint64_t x, uint64_t y, double c_re,
                                double c_im) {
  if (y < a->r) {
    a->r = y;
    a->c_re = c_re;
    a->c_im = c_im;
  }
  if (y == x) a->n = a->n + c_re;
}

/* generator.cpp:42-56 */
static inline void si_step(ora_group const *g, si_acc *a, uint64_t x, int e) {
  uint64_t const y = image_of(g, e, x);
  double re = g->eig_re[e], im = g->eig_im[e];
  if (g->fermion_species && (y < a->r || y == x)) {
    double const sg = (double)fermion_sign(g, e, x);
    re = re * sg;
    im = im * sg;
  }
  si_step_impl(a, x, y, re, im);
  if (g->spin_inversion == 1) si_step_impl(a, x, y ^ g->flip_mask, re, im);
  else if (g->spin_inversion == -1)
    si_step_impl(a, x, y ^ g->flip_mask, -re, -im);
}

/*
 * Reduction order of generator.cpp:77-140 with chunk = `lanes`
 * (2 * natural_vector_size(uint64); 8 on avx2): per-lane accumulators seeded
 * with (x, 1, 0, 0), lanes merged left to right keeping the earlier lane on
 * ties (:58-75), then tail elements one by one, norm = sqrt(n / |G|) or
 * sqrt(n / (2|G|)) with spin inversion (:136-140).
 */
static void state_info_one(ora_group const *g, int lanes, uint64_t x,
                           uint64_t *rep, double *chr, double *norm) {
  int const G = g->size;
  int const nchunks = G / lanes;
  si_acc acc[64];
  for (int j = 0; j < lanes; ++j) acc[j] = (si_acc){x, 1.0, 0.0, 0.0};
  for (int c = 0; c < nchunks; ++c)
    for (int j = 0; j < lanes; ++j) si_step(g, &acc[j], x, c * lanes + j);
  si_acc cur = acc[0];
  for (int j = 1; j < lanes; ++j) {
    if (acc[j].r < cur.r) {
      cur.r = acc[j].r;
      cur.c_re = acc[j].c_re;
      cur.c_im = acc[j].c_im;
    }
    cur.n = cur.n + acc[j].n;
  }
  for (int e = nchunks * lanes; e < G; ++e) si_step(g, &cur, x, e);
  *rep = cur.r;
  chr[0] = cur.c_re;
  chr[1] = cur.c_im;
  double const denom = (g->spin_inversion == 0) ? (double)G : (double)(2 * G);
  *norm = (g->norm_eps > 0 && !(cur.n > g->norm_eps)) ? 0.0 : sqrt(cur.n / denom);
}

ORA_EXPORT void ora_state_info(ora_group const *g, int lanes, ptrdiff_t batch,
                               uint64_t const *xs, ptrdiff_t stride,
                               uint64_t *reps, double *chars, double *norms) {
  for (ptrdiff_t i = 0; i < batch; ++i)
    state_info_one(g, lanes, xs[i * stride], reps + i, chars + 2 * i, norms + i);
}

/* ------------------------------------------------------------------------ */
/* Enumeration: chapel/src/StatesEnumeration.chpl                            */
/* ------------------------------------------------------------------------ */

/* StatesEnumeration.chpl:29-32 (Gosper's next state of equal weight). */
ORA_EXPORT uint64_t ora_next_state_fixed_hamming(uint64_t v) {
  uint64_t const t = v | (v - 1);
  return (t + 1) | (((~t & (t + 1)) - 1) >> (__builtin_ctzll(v) + 1));
}

/* Basis.hs:487-511 (binomial table and fixedHammingStateToIndex). */
static uint64_t binom_tab[65][65];
static int binom_ready = 0;
static void init_binom(void) {
  if (binom_ready) return;
  memset(binom_tab, 0, sizeof(binom_tab));
  for (int n = 0; n <= 64; ++n) {
    binom_tab[n][0] = 1;
    for (int k = 1; k <= n; ++k)
      binom_tab[n][k] = binom_tab[n - 1][k - 1] + (k <= n - 1 ? binom_tab[n - 1][k] : 0);
  }
  binom_ready = 1;
}
static uint64_t binom(int n, int k) {
  if (n <= 0 || k > n) return 0; /* Basis.hs:506-509 */
  return binom_tab[n][k];
}

ORA_EXPORT uint64_t ora_fixed_hamming_state_to_index(uint64_t a) {
  init_binom();
  uint64_t i = 0;
  int k = 1;
  while (a != 0) {
    int const c = __builtin_ctzll(a);
    a &= a - 1;
    i += binom(c, k);
    ++k;
  }
  return i;
}

/* Basis.hs:522-545 (fixedHammingIndexToState). */
ORA_EXPORT uint64_t ora_fixed_hamming_index_to_state(uint64_t index, int h) {
  init_binom();
  uint64_t state = 0;
  for (int i = h; i > 0; --i) {
    /* largest c with binom(c, i) <= index */
    int c = i - 1;
    uint64_t contribution = 0;
    while (c + 1 < 64 && binom(c + 1, i) <= index) {
      ++c;
      contribution = binom(c, i);
    }
    state |= 1ULL << c;
    index -= contribution;
  }
  return state;
}

/*
 * StatesEnumeration.chpl:242-268 (_enumerateStatesProjected): walk every state
 * of the Hamming weight of `lower` from lower to upper (inclusive, ascending)
 * and keep those with is_representative flag != 0 and norm > 0.  Writes at
 * most `cap` states; returns the total number of representatives found.
 */
ORA_EXPORT int64_t ora_enumerate_projected(ora_group const *g, int lanes,
                                           uint64_t lower, uint64_t upper,
                                           uint64_t *out, int64_t cap) {
  if (lower > upper) return 0;
  int64_t count = 0;
  uint64_t v = lower;
  for (;;) {
    uint8_t f;
    double nrm;
    is_rep_one(g, lanes, v, &f, &nrm);
    if (f && nrm > 0) {
      if (count < cap) out[count] = v;
      ++count;
    }
    if (v == upper) break;
    v = ora_next_state_fixed_hamming(v);
  }
  return count;
}

/* Parallel variant used for the CPU baseline: the candidate range is split by
   combinadics rank into contiguous chunks (StatesEnumeration.chpl:130-152),
   each chunk is enumerated by one thread, and the per-chunk counts are
   returned (the states themselves are not kept). */
ORA_EXPORT int64_t ora_count_projected_parallel(ora_group const *g, int lanes,
                                                uint64_t lower,
                                                uint64_t num_candidates,
                                                int num_threads) {
  init_binom();
  int const h = __builtin_popcountll(lower);
  uint64_t const r0 = ora_fixed_hamming_state_to_index(lower);
  int64_t total = 0;
  int64_t const nchunks = (int64_t)num_threads * 64;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(num_threads)
  for (int64_t c = 0; c < nchunks; ++c) {
    uint64_t const b = (uint64_t)c * num_candidates / (uint64_t)nchunks;
    uint64_t const e = (uint64_t)(c + 1) * num_candidates / (uint64_t)nchunks;
    if (b == e) continue;
    uint64_t v = ora_fixed_hamming_index_to_state(r0 + b, h);
    for (uint64_t k = b; k < e; ++k) {
      uint8_t f;
      double nrm;
      is_rep_one(g, lanes, v, &f, &nrm);
      if (f && nrm > 0) ++total;
      if (k + 1 < e) v = ora_next_state_fixed_hamming(v);
    }
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* Index: kernels/indexing.c                                                 */
/* ------------------------------------------------------------------------ */

/* indexing.c:273-325 semantics: position of x in the sorted list, -1 when
   absent (a plain lower_bound; the prefix-bucket table of indexing.c:45-92
   only narrows the search window and does not change the answer). */
ORA_EXPORT void ora_state_index(int64_t num_states, uint64_t const *reps,
                                ptrdiff_t batch, uint64_t const *xs,
                                int64_t *out) {
  for (ptrdiff_t i = 0; i < batch; ++i) {
    uint64_t const x = xs[i];
    int64_t lo = 0, hi = num_states;
    while (lo < hi) {
      int64_t const mid = lo + (hi - lo) / 2;
      if (reps[mid] < x) lo = mid + 1;
      else hi = mid;
    }
    out[i] = (lo < num_states && reps[lo] == x) ? lo : -1;
  }
}

/* ------------------------------------------------------------------------ */
/* Term application: kernels/reference.c:67-134                              */
/* ------------------------------------------------------------------------ */

typedef struct {
  int number_terms;
  double const *v; /* interleaved re, im */
  uint64_t const *m, *r, *x, *s;
} ora_terms;

/* reference.c:67-95 (ls_internal_operator_apply_diag_x1). xs may be NULL. */
ORA_EXPORT void ora_apply_diag(int T, double const *v, uint64_t const *m,
                               uint64_t const *r, uint64_t const *s,
                               ptrdiff_t batch, uint64_t const *alphas,
                               double *ys, double const *xs) {
  if (T == 0) {
    memset(ys, 0, sizeof(double) * (size_t)batch);
    return;
  }
  for (ptrdiff_t i = 0; i < batch; ++i) {
    double acc = 0;
    uint64_t const a = alphas[i];
    for (int t = 0; t < T; ++t) {
      int const delta = (a & m[t]) == r[t];
      if (delta) {
        int const sign = 1 - 2 * (__builtin_popcountll(a & s[t]) % 2);
        double const factor =
            xs != NULL ? (double)(delta * sign) * xs[i] : (double)(delta * sign);
        acc += v[2 * t] * factor;
      }
    }
    ys[i] = acc;
  }
}

/* reference.c:97-134 (ls_internal_operator_apply_off_diag_x1). */
ORA_EXPORT void ora_apply_off_diag(int T, double const *v, uint64_t const *m,
                                   uint64_t const *r, uint64_t const *x,
                                   uint64_t const *s, ptrdiff_t batch,
                                   uint64_t const *alphas, uint64_t *betas,
                                   double *coeffs, int64_t *offsets,
                                   double const *xs) {
  if (T == 0) {
    memset(offsets, 0, sizeof(int64_t) * (size_t)(batch + 1));
    return;
  }
  int64_t off = 0;
  offsets[0] = 0;
  for (ptrdiff_t i = 0; i < batch; ++i) {
    uint64_t const a = alphas[i];
    for (int t = 0; t < T; ++t) {
      int const delta = (a & m[t]) == r[t];
      if (delta) {
        int const sign = 1 - 2 * (__builtin_popcountll(a & s[t]) % 2);
        double const factor =
            xs != NULL ? (double)(delta * sign) * xs[i] : (double)(delta * sign);
        coeffs[2 * off] = v[2 * t] * factor;
        coeffs[2 * off + 1] = v[2 * t + 1] * factor;
        betas[off] = a ^ x[t];
        ++off;
      }
    }
    offsets[i + 1] = off;
  }
}

/* ------------------------------------------------------------------------ */
/* Matrix-vector product: chapel/src/DistributedMatrixVector.chpl:34-143,    */
/* 1045-1105 and chapel/src/BatchedOperator.chpl:139-282 (push form).        */
/* ------------------------------------------------------------------------ */

typedef struct {
  int Td, To;
  double const *vd; uint64_t const *md, *rd, *sd;
  double const *vo; uint64_t const *mo, *ro, *xo, *so;
} ora_op;

/*
 * kind: 0 = no projection (BatchedOperator.chpl:139-164),
 *       1 = spin inversion only (:166-205),
 *       2 = permutations (+ inversion) with state_info rescale (:207-253).
 * Real x / y, coefficients cast to real (DistributedMatrixVector.chpl:109,126).
 * The diagonal pass assigns y (:34-44); without diagonal terms y is left as
 * given (:1054-1055).  Returns 0, or -1 when a nonzero coefficient points to
 * a state outside the basis (:127-135; the reference halts there), with the
 * offending row/state in err_info[0..1].
 * The sum is accumulated row by row, term by term (a sequential order of the
 * reference's relaxed atomic adds, ConcurrentAccessor.chpl:31-33).
 */
ORA_EXPORT int ora_matvec_push(ora_group const *g, int lanes, int kind,
                               int number_sites, int spin_inversion,
                               int64_t num_states, uint64_t const *reps,
                               int Td, double const *vd, uint64_t const *md,
                               uint64_t const *rd, uint64_t const *sd,
                               int To, double const *vo, uint64_t const *mo,
                               uint64_t const *ro, uint64_t const *xo,
                               uint64_t const *so, double const *x, double *y,
                               int64_t *err_info) {
  if (Td > 0) ora_apply_diag(Td, vd, md, rd, sd, num_states, reps, y, x);
  if (To == 0) return 0;
  uint64_t const mask = flip_mask_of(number_sites);
  for (int64_t i = 0; i < num_states; ++i) {
    uint64_t const a = reps[i];
    double na = 1.0;
    if (kind == 2) {
      uint64_t rr;
      double ch[2];
      state_info_one(g, lanes, a, &rr, ch, &na);
    }
    for (int t = 0; t < To; ++t) {
      if ((a & mo[t]) != ro[t]) continue;
      int const sign = 1 - 2 * (__builtin_popcountll(a & so[t]) % 2);
      double const factor = (double)sign * x[i];
      double c_re = vo[2 * t] * factor, c_im = vo[2 * t + 1] * factor;
      uint64_t b = a ^ xo[t];
      if (kind == 1) {
        uint64_t const inv = b ^ mask;
        if (inv < b) {
          b = inv;
          c_re *= (double)spin_inversion;
          c_im *= (double)spin_inversion;
        }
      } else if (kind == 2) {
        uint64_t rb;
        double ch[2], nb;
        state_info_one(g, lanes, b, &rb, ch, &nb);
        /* BatchedOperator.chpl:244: cs *= tempCoeffs * norm_b / norm_a */
        double const t_re = (c_re * nb) / na, t_im = (c_im * nb) / na;
        c_re = ch[0] * t_re - ch[1] * t_im;
        c_im = ch[0] * t_im + ch[1] * t_re;
        b = rb;
      }
      double const c = c_re; /* cast to real(64) */
      if (c == 0) continue;
      int64_t idx;
      ora_state_index(num_states, reps, 1, &b, &idx);
      if (idx < 0) {
        if (err_info) { err_info[0] = i; err_info[1] = (int64_t)b; }
        return -1;
      }
      y[idx] += c;
    }
  }
  return 0;
}

/*
 * The full complex matrix M of the projected operator, in the reference's
 * convention (entries as in BatchedOperator.chpl:207-253 without the cast to
 * real): M[idx(rep b), i] += chi_b * v * sign * n_b / n_a.  Written as a
 * coordinate list (rows, cols, re, im) of at most `cap` entries; returns the
 * number of entries or -1 when a state is outside the basis.  Used by the
 * tests to form dense matrices for brute-force checks and to build the
 * complex-vector products that the reference does not offer.
 */
ORA_EXPORT int64_t ora_matrix_coo(ora_group const *g, int lanes, int kind,
                                  int number_sites, int spin_inversion,
                                  int64_t num_states, uint64_t const *reps,
                                  int Td, double const *vd, uint64_t const *md,
                                  uint64_t const *rd, uint64_t const *sd,
                                  int To, double const *vo, uint64_t const *mo,
                                  uint64_t const *ro, uint64_t const *xo,
                                  uint64_t const *so, int64_t cap,
                                  int64_t *rows, int64_t *cols, double *re,
                                  double *im) {
  int64_t k = 0;
  uint64_t const mask = flip_mask_of(number_sites);
  for (int64_t i = 0; i < num_states; ++i) {
    uint64_t const a = reps[i];
    double d = 0;
    for (int t = 0; t < Td; ++t)
      if ((a & md[t]) == rd[t]) {
        int const sign = 1 - 2 * (__builtin_popcountll(a & sd[t]) % 2);
        d += vd[2 * t] * (double)sign;
      }
    if (Td > 0) {
      if (k < cap) { rows[k] = i; cols[k] = i; re[k] = d; im[k] = 0; }
      ++k;
    }
    double na = 1.0;
    if (kind == 2) {
      uint64_t rr;
      double ch[2];
      state_info_one(g, lanes, a, &rr, ch, &na);
    }
    for (int t = 0; t < To; ++t) {
      if ((a & mo[t]) != ro[t]) continue;
      int const sign = 1 - 2 * (__builtin_popcountll(a & so[t]) % 2);
      double c_re = vo[2 * t] * (double)sign, c_im = vo[2 * t + 1] * (double)sign;
      uint64_t b = a ^ xo[t];
      if (kind == 1) {
        uint64_t const inv = b ^ mask;
        if (inv < b) { b = inv; c_re *= spin_inversion; c_im *= spin_inversion; }
      } else if (kind == 2) {
        uint64_t rb;
        double ch[2], nb;
        state_info_one(g, lanes, b, &rb, ch, &nb);
        double const t_re = (c_re * nb) / na, t_im = (c_im * nb) / na;
        c_re = ch[0] * t_re - ch[1] * t_im;
        c_im = ch[0] * t_im + ch[1] * t_re;
        b = rb;
      }
      if (c_re == 0 && c_im == 0) continue;
      int64_t idx;
      ora_state_index(num_states, reps, 1, &b, &idx);
      if (idx < 0) return -1;
      if (k < cap) { rows[k] = idx; cols[k] = i; re[k] = c_re; im[k] = c_im; }
      ++k;
    }
  }
  return k;
}

/*
 * CPU baseline: the push-form product from the source rows row0, row0 +
 * stride, ... (< row1) of the basis, OpenMP-parallel with relaxed atomic adds
 * into y like ConcurrentAccessor.chpl:31-33: the diagonal pass assigns y at
 * those rows (DistributedMatrixVector.chpl:34-71), then each source row adds
 * its off-diagonal pairs wherever they land (y must hold zeros elsewhere).
 * Returns the number of nonzero contributions processed (diagonal entries
 * included).  stride 1 over [0, num_states) is the whole product.
 */
ORA_EXPORT int64_t ora_matvec_push_parallel(
    ora_group const *g, int lanes, int64_t num_states, uint64_t const *reps,
    int64_t row0, int64_t row1, int64_t stride, int Td, double const *vd,
    uint64_t const *md, uint64_t const *rd, uint64_t const *sd, int To,
    double const *vo, uint64_t const *mo, uint64_t const *ro,
    uint64_t const *xo, uint64_t const *so, double const *x, double *y,
    int num_threads) {
  int64_t nnz = 0;
  if (row0 < 0) row0 = 0;
  if (row1 > num_states) row1 = num_states;
  if (stride < 1) stride = 1;
  int64_t const count = row1 > row0 ? (row1 - row0 + stride - 1) / stride : 0;
#pragma omp parallel for schedule(static) reduction(+ : nnz) num_threads(num_threads)
  for (int64_t q = 0; q < count; ++q) {
    int64_t const i = row0 + q * stride;
    uint64_t const a = reps[i];
    double d = 0;
    for (int t = 0; t < Td; ++t)
      if ((a & md[t]) == rd[t]) {
        int const sign = 1 - 2 * (__builtin_popcountll(a & sd[t]) % 2);
        d += vd[2 * t] * ((double)sign * x[i]);
      }
    y[i] = d;
    ++nnz;
  }
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nnz) num_threads(num_threads)
  for (int64_t q = 0; q < count; ++q) {
    int64_t const i = row0 + q * stride;
    uint64_t const a = reps[i];
    uint64_t rr;
    double ch[2], na;
    state_info_one(g, lanes, a, &rr, ch, &na);
    for (int t = 0; t < To; ++t) {
      if ((a & mo[t]) != ro[t]) continue;
      int const sign = 1 - 2 * (__builtin_popcountll(a & so[t]) % 2);
      double const factor = (double)sign * x[i];
      double const c_re0 = vo[2 * t] * factor, c_im0 = vo[2 * t + 1] * factor;
      uint64_t rb;
      double nb;
      state_info_one(g, lanes, a ^ xo[t], &rb, ch, &nb);
      double const t_re = (c_re0 * nb) / na, t_im = (c_im0 * nb) / na;
      double const c = ch[0] * t_re - ch[1] * t_im;
      if (c == 0) continue;
      int64_t idx;
      ora_state_index(num_states, reps, 1, &rb, &idx);
      if (idx >= 0) {
#pragma omp atomic
        y[idx] += c;
      }
      ++nnz;
    }
  }
  return nnz;
}

/*
 * Selected rows of y = Re(M) x, each computed from the pairs that its own
 * representative emits: y_i = d_i x_i + sum Re(C) x_idx(rep beta), with C as
 * in BatchedOperator.chpl:232-247.  This is row i of Re(M)^T x: equal to the
 * push form (row i of Re(M) x) only when Re(M) is symmetric, which holds for
 * Hermitian operators that respect the group (the tests pin the equality on
 * such sectors and show the difference on an operator that is not).  Used to
 * spot-check rows of full-size products.  Returns -1 on a state outside the
 * basis.
 */
ORA_EXPORT int ora_matvec_rows(ora_group const *g, int lanes, int64_t num_states,
                               uint64_t const *reps, int64_t nrows,
                               int64_t const *rows, int Td, double const *vd,
                               uint64_t const *md, uint64_t const *rd,
                               uint64_t const *sd, int To, double const *vo,
                               uint64_t const *mo, uint64_t const *ro,
                               uint64_t const *xo, uint64_t const *so,
                               double const *x, double *out) {
  int status = 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < nrows; ++q) {
    int64_t const i = rows[q];
    uint64_t const a = reps[i];
    double d = 0;
    for (int t = 0; t < Td; ++t)
      if ((a & md[t]) == rd[t]) {
        int const sign = 1 - 2 * (__builtin_popcountll(a & sd[t]) % 2);
        d += vd[2 * t] * ((double)sign * x[i]);
      }
    uint64_t rr;
    double ch[2], na;
    state_info_one(g, lanes, a, &rr, ch, &na);
    double acc = 0;
    for (int t = 0; t < To; ++t) {
      if ((a & mo[t]) != ro[t]) continue;
      int const sign = 1 - 2 * (__builtin_popcountll(a & so[t]) % 2);
      uint64_t rb;
      double nb;
      state_info_one(g, lanes, a ^ xo[t], &rb, ch, &nb);
      double const t_re = (vo[2 * t] * (double)sign * nb) / na;
      double const t_im = (vo[2 * t + 1] * (double)sign * nb) / na;
      double const c = ch[0] * t_re - ch[1] * t_im;
      if (c == 0) continue;
      int64_t idx;
      ora_state_index(num_states, reps, 1, &rb, &idx);
      if (idx < 0) {
        status = -1;
        continue;
      }
      acc += c * x[idx];
    }
    out[q] = d + acc;
  }
  return status;
}

ORA_EXPORT int ora_num_threads(void) { return omp_get_max_threads(); }
