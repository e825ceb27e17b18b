"""Python access to the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.

How each oracle function is pinned (see tests/test_oracle.py):
  ora_apply_diag / ora_apply_off_diag
      bit-exact against golden vectors written by the reference's own
      kernels/reference.c (tests/golden/, produced by gen_golden.py from a
      gcc build of that file: oracle/_ref).
  ora_state_index
      bit-exact against golden vectors of the reference's
      kernels/indexing.c (same route), plus absent / edge states.
  ora_group_closure / ora_characters
      group orders of the reference's model files (chain 32: 64, square
      4x4 and 6x6: 128 and 288) and closure/consistency properties;
      characters against cos/sin of 2 pi k / N closed forms.
  ora_permute_bits / ora_benes_apply
      BenesSpec.hs:34-42 fixed vectors; direct-vs-network agreement.
  ora_state_info / ora_is_representative / enumeration
      Burnside's lemma for the number of representatives, brute-force orbit
      minima over explicit permutations (tiny systems), the norm identity
      norm^2 |G| = sum over the stabiliser of chi, and E0 known answers
      (python/example/getting_started.py:51: -18.06178542; HPhi energies in
      test/*/HPhi/output/zvo_energy.dat) through dense ED.
  ora_matvec_push / ora_matrix_coo
      dense ED of the projected matrix against the eigenvalues of the full
      (unsymmetrised) Hamiltonian restricted to the sector, brute force.
The Halide kernels (kernels/generator.cpp) cannot be built here, so the
exact floating-point summation order of their norms is restated, not
compared: for sectors whose characters are not +-1 the norm > 0 decision
on numerically-zero norms is "parity unpinned" beyond the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_lib = None

_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so" if False else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_permute_bits.restype = C.c_uint64
        L.ora_permute_bits.argtypes = [C.c_int, _i32p, C.c_uint64]
        L.ora_benes_apply.restype = C.c_uint64
        L.ora_benes_apply.argtypes = [C.c_int, _u64p, C.c_ssize_t, _u64p, C.c_uint64]
        L.ora_group_closure.restype = C.c_int
        L.ora_group_closure.argtypes = [C.c_int, C.c_int, _i32p, _i32p, C.c_int, _i32p, _i64p, _i64p]
        L.ora_characters.restype = None
        L.ora_characters.argtypes = [C.c_int, _i64p, _i64p, _f64p, _f64p]
        L.ora_group_create.restype = C.c_void_p
        L.ora_group_create.argtypes = [C.c_int, C.c_int, _i32p, _f64p, _f64p, C.c_int, C.c_int, _u64p, _u64p]
        L.ora_group_set_mode.restype = C.c_int
        L.ora_group_set_mode.argtypes = [C.c_void_p, C.c_int]
        L.ora_group_destroy.restype = None
        L.ora_group_destroy.argtypes = [C.c_void_p]
        L.ora_group_images.restype = None
        L.ora_group_images.argtypes = [C.c_void_p, C.c_uint64, _u64p]
        L.ora_is_representative.restype = None
        L.ora_is_representative.argtypes = [C.c_void_p, C.c_int, C.c_ssize_t, _u64p, C.c_ssize_t, _u8p, _f64p]
        L.ora_state_info.restype = None
        L.ora_state_info.argtypes = [C.c_void_p, C.c_int, C.c_ssize_t, _u64p, C.c_ssize_t, _u64p, _f64p, _f64p]
        L.ora_next_state_fixed_hamming.restype = C.c_uint64
        L.ora_next_state_fixed_hamming.argtypes = [C.c_uint64]
        L.ora_fixed_hamming_state_to_index.restype = C.c_uint64
        L.ora_fixed_hamming_state_to_index.argtypes = [C.c_uint64]
        L.ora_fixed_hamming_index_to_state.restype = C.c_uint64
        L.ora_fixed_hamming_index_to_state.argtypes = [C.c_uint64, C.c_int]
        L.ora_enumerate_projected.restype = C.c_int64
        L.ora_enumerate_projected.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, _u64p, C.c_int64]
        L.ora_count_projected_parallel.restype = C.c_int64
        L.ora_count_projected_parallel.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_int]
        L.ora_state_index.restype = None
        L.ora_state_index.argtypes = [C.c_int64, _u64p, C.c_ssize_t, _u64p, _i64p]
        L.ora_apply_diag.restype = None
        L.ora_apply_diag.argtypes = [C.c_int, _f64p, _u64p, _u64p, _u64p, C.c_ssize_t, _u64p, _f64p, _f64p]
        L.ora_apply_off_diag.restype = None
        L.ora_apply_off_diag.argtypes = [C.c_int, _f64p, _u64p, _u64p, _u64p, _u64p, C.c_ssize_t, _u64p, _u64p,
                                         _f64p, _i64p, _f64p]
        L.ora_matvec_push.restype = C.c_int
        L.ora_matvec_push.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, _u64p,
                                      C.c_int, _f64p, _u64p, _u64p, _u64p,
                                      C.c_int, _f64p, _u64p, _u64p, _u64p, _u64p, _f64p, _f64p, _i64p]
        L.ora_matrix_coo.restype = C.c_int64
        L.ora_matrix_coo.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, _u64p,
                                     C.c_int, _f64p, _u64p, _u64p, _u64p,
                                     C.c_int, _f64p, _u64p, _u64p, _u64p, _u64p, C.c_int64, _i64p, _i64p,
                                     _f64p, _f64p]
        L.ora_matvec_push_parallel.restype = C.c_int64
        L.ora_matvec_push_parallel.argtypes = [C.c_void_p, C.c_int, C.c_int64, _u64p, C.c_int64, C.c_int64, C.c_int64,
                                               C.c_int, _f64p, _u64p, _u64p, _u64p,
                                               C.c_int, _f64p, _u64p, _u64p, _u64p, _u64p, _f64p, _f64p, C.c_int]
        L.ora_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def permute_bits(perm, x: int) -> int:
    p = np.ascontiguousarray(perm, dtype=np.int32)
    return int(lib().ora_permute_bits(p.size, _p(p, _i32p), x))


def group_closure(n: int, generators, sectors, cap: int = 1 << 16):
    gens = np.ascontiguousarray(np.asarray(generators, dtype=np.int32).reshape(-1, n))
    secs = np.ascontiguousarray(sectors, dtype=np.int32)
    perms = np.zeros((cap, n), dtype=np.int32)
    num = np.zeros(cap, dtype=np.int64)
    den = np.zeros(cap, dtype=np.int64)
    k = lib().ora_group_closure(n, gens.shape[0], _p(gens, _i32p), _p(secs, _i32p), cap, _p(perms, _i32p),
                                _p(num, _i64p), _p(den, _i64p))
    if k < 0:
        raise ValueError(f"group closure failed with status {k}")
    re = np.zeros(k)
    im = np.zeros(k)
    lib().ora_characters(k, _p(num, _i64p), _p(den, _i64p), _p(re, _f64p), _p(im, _f64p))
    return perms[:k].copy(), num[:k].copy(), den[:k].copy(), re, im


class Group:
    """Oracle group handle (direct permutation tables; optional Benes masks)."""

    def __init__(self, n, perms, eig_re, eig_im, spin_inversion, masks=None, shifts=None):
        self.n = n
        self.perms = np.ascontiguousarray(perms, dtype=np.int32)
        self.size = self.perms.shape[0]
        re = np.ascontiguousarray(eig_re, dtype=np.float64)
        im = np.ascontiguousarray(eig_im, dtype=np.float64)
        self.spin_inversion = spin_inversion
        mp = sp = None
        depth = 0
        if masks is not None:
            self._masks = np.ascontiguousarray(masks, dtype=np.uint64)
            self._shifts = np.ascontiguousarray(shifts, dtype=np.uint64)
            depth = self._masks.shape[0]
            mp, sp = _p(self._masks, _u64p), _p(self._shifts, _u64p)
        self.h = lib().ora_group_create(n, self.size, _p(self.perms, _i32p), _p(re, _f64p), _p(im, _f64p),
                                        spin_inversion, depth, mp, sp)

    @staticmethod
    def from_generators(n, generators, sectors, spin_inversion=0):
        perms, num, den, re, im = group_closure(n, generators, sectors)
        g = Group(n, perms, re, im, spin_inversion)
        g.num, g.den = num, den
        return g

    def use_benes(self, flag: bool):
        if lib().ora_group_set_mode(self.h, 1 if flag else 0) != 0:
            raise ValueError("mode not available")

    def images(self, x: int) -> np.ndarray:
        out = np.zeros(self.size, dtype=np.uint64)
        lib().ora_group_images(self.h, x, _p(out, _u64p))
        return out

    def is_representative(self, xs, lanes: int = 4):
        xs = np.ascontiguousarray(xs, dtype=np.uint64)
        f = np.zeros(xs.size, dtype=np.uint8)
        n = np.zeros(xs.size)
        lib().ora_is_representative(self.h, lanes, xs.size, _p(xs, _u64p), 1, _p(f, _u8p), _p(n, _f64p))
        return f, n

    def state_info(self, xs, lanes: int = 8):
        xs = np.ascontiguousarray(xs, dtype=np.uint64)
        r = np.zeros(xs.size, dtype=np.uint64)
        c = np.zeros((xs.size, 2))
        n = np.zeros(xs.size)
        lib().ora_state_info(self.h, lanes, xs.size, _p(xs, _u64p), 1, _p(r, _u64p), _p(c, _f64p), _p(n, _f64p))
        return r, c[:, 0] + 1j * c[:, 1], n

    def enumerate(self, lower: int, upper: int, cap: int = 1 << 26, lanes: int = 4) -> np.ndarray:
        out = np.zeros(cap, dtype=np.uint64)
        k = lib().ora_enumerate_projected(self.h, lanes, lower, upper, _p(out, _u64p), cap)
        if k > cap:
            raise ValueError("capacity exceeded")
        return out[:k].copy()

    def count_parallel(self, lower: int, num_candidates: int, threads: int, lanes: int = 4) -> int:
        return int(lib().ora_count_projected_parallel(self.h, lanes, lower, num_candidates, threads))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_group_destroy(self.h)
            self.h = None


def next_state(v: int) -> int:
    return int(lib().ora_next_state_fixed_hamming(v))


def state_to_index(v: int) -> int:
    return int(lib().ora_fixed_hamming_state_to_index(v))


def index_to_state(i: int, h: int) -> int:
    return int(lib().ora_fixed_hamming_index_to_state(i, h))


def enumerate_unprojected(n: int, h, spin_inversion: int = 0):
    """All states of weight h (or all states) as the unprojected enumeration
    (StatesEnumeration.chpl:269-289)."""
    if h is None:
        hi = (1 << n) - 1
        if spin_inversion:
            hi = min(hi, hi ^ ((1 << n) - 1))
        return np.arange(0, hi + 1, dtype=np.uint64)
    lo = (1 << h) - 1
    hi = lo << (n - h) if not spin_inversion else lo << (n - h - 1)
    if spin_inversion:
        hi = min(hi, hi ^ ((1 << n) - 1))
    out = []
    v = lo
    while True:
        out.append(v)
        if v == hi:
            break
        v = next_state(v)
    return np.array(out, dtype=np.uint64)


def state_index(reps, xs):
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    xs = np.ascontiguousarray(xs, dtype=np.uint64)
    out = np.zeros(xs.size, dtype=np.int64)
    lib().ora_state_index(reps.size, _p(reps, _u64p), xs.size, _p(xs, _u64p), _p(out, _i64p))
    return out


class Terms:
    """Term arrays (v interleaved re/im)."""

    def __init__(self, terms):
        n = len(terms)
        self.count = n
        self.v = np.zeros(2 * max(n, 1))
        self.m = np.zeros(max(n, 1), dtype=np.uint64)
        self.r = np.zeros(max(n, 1), dtype=np.uint64)
        self.x = np.zeros(max(n, 1), dtype=np.uint64)
        self.s = np.zeros(max(n, 1), dtype=np.uint64)
        for i, t in enumerate(terms):
            self.v[2 * i], self.v[2 * i + 1] = complex(t.v).real, complex(t.v).imag
            self.m[i], self.r[i], self.x[i], self.s[i] = t.m, t.r, t.x, t.s

    def args(self, with_x=True):
        a = [self.count, _p(self.v, _f64p), _p(self.m, _u64p), _p(self.r, _u64p)]
        if with_x:
            a.append(_p(self.x, _u64p))
        a.append(_p(self.s, _u64p))
        return a


def apply_diag(terms: Terms, alphas, xs=None):
    a = np.ascontiguousarray(alphas, dtype=np.uint64)
    y = np.zeros(a.size)
    xp = None if xs is None else _p(np.ascontiguousarray(xs, dtype=np.float64), _f64p)
    lib().ora_apply_diag(*terms.args(with_x=False), a.size, _p(a, _u64p), _p(y, _f64p), xp)
    return y


def apply_off_diag(terms: Terms, alphas, xs=None):
    a = np.ascontiguousarray(alphas, dtype=np.uint64)
    cap = max(a.size * terms.count, 1)
    b = np.zeros(cap, dtype=np.uint64)
    c = np.zeros(2 * cap)
    o = np.zeros(a.size + 1, dtype=np.int64)
    xs_arr = None if xs is None else np.ascontiguousarray(xs, dtype=np.float64)
    xp = None if xs_arr is None else _p(xs_arr, _f64p)
    lib().ora_apply_off_diag(*terms.args(), a.size, _p(a, _u64p), _p(b, _u64p), _p(c, _f64p), _p(o, _i64p), xp)
    k = int(o[-1])
    return b[:k], c[:2 * k:2] + 1j * c[1:2 * k:2], o


def projection_kind(group, spin_inversion):
    if group is not None and group.size > 0:
        return 2
    return 1 if spin_inversion else 0


def matvec_push(group, number_sites, spin_inversion, reps, diag: Terms, off: Terms, x, y=None):
    """The reference's real product (push form, sequential order)."""
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(reps.size) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
    err = np.zeros(2, dtype=np.int64)
    kind = projection_kind(group, spin_inversion)
    rc = lib().ora_matvec_push(group.h if group is not None else None, 8, kind, number_sites, spin_inversion,
                               reps.size, _p(reps, _u64p), *diag.args(with_x=False), *off.args(),
                               _p(x, _f64p), _p(y, _f64p), _p(err, _i64p))
    if rc != 0:
        raise ValueError(f"state {err[1]} (from row {err[0]}) is not in the basis")
    return y


def matrix_dense(group, number_sites, spin_inversion, reps, diag: Terms, off: Terms, complex_values=True):
    """Dense matrix M (M[idx(rep beta), i] += C) of the projected operator."""
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    n = reps.size
    cap = n * (off.count + 1) + 1
    rows = np.zeros(cap, dtype=np.int64)
    cols = np.zeros(cap, dtype=np.int64)
    re = np.zeros(cap)
    im = np.zeros(cap)
    kind = projection_kind(group, spin_inversion)
    k = lib().ora_matrix_coo(group.h if group is not None else None, 8, kind, number_sites, spin_inversion,
                             n, _p(reps, _u64p), *diag.args(with_x=False), *off.args(), cap,
                             _p(rows, _i64p), _p(cols, _i64p), _p(re, _f64p), _p(im, _f64p))
    if k < 0:
        raise ValueError("a state is not in the basis")
    M = np.zeros((n, n), dtype=np.complex128)
    np.add.at(M, (rows[:k], cols[:k]), re[:k] + 1j * im[:k])
    return M if complex_values else M.real


def matvec_push_parallel(group, reps, diag: Terms, off: Terms, x, y, row0, row1, threads, stride=1):
    """Push-form product from source rows row0, row0 + stride, ... < row1
    (OpenMP, atomic adds into y, which must be zero elsewhere); returns the
    number of contributions."""
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    return int(lib().ora_matvec_push_parallel(group.h, 8, reps.size, _p(reps, _u64p), row0, row1, stride,
                                              *diag.args(with_x=False), *off.args(), _p(x, _f64p),
                                              _p(y, _f64p), threads))


def lanczos_ground_state(matvec, n, steps=100, seed=42):
    """Plain FP64 Lanczos with full reorthogonalisation (numpy); returns the
    lowest Ritz value after `steps` steps (or fewer on breakdown)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    v /= np.linalg.norm(v)
    V = [v]
    alphas, betas = [], []
    for _ in range(min(steps, n)):
        w = matvec(V[-1])
        a = float(np.dot(V[-1], w))
        alphas.append(a)
        w = w - a * V[-1] - (betas[-1] * V[-2] if betas else 0)
        for u in V:
            w -= np.dot(u, w) * u
        b = float(np.linalg.norm(w))
        if b < 1e-12 or len(V) >= n:
            break
        betas.append(b)
        V.append(w / b)
    T = np.diag(alphas) + np.diag(betas[:len(alphas) - 1], 1) + np.diag(betas[:len(alphas) - 1], -1)
    return float(np.linalg.eigvalsh(T)[0])


def matvec_rows(group, reps, diag: Terms, off: Terms, x, rows):
    """Selected rows of Re(M) x (see ora_matvec_rows)."""
    L = lib()
    L.ora_matvec_rows.restype = C.c_int
    L.ora_matvec_rows.argtypes = [C.c_void_p, C.c_int, C.c_int64, _u64p, C.c_int64, _i64p,
                                  C.c_int, _f64p, _u64p, _u64p, _u64p,
                                  C.c_int, _f64p, _u64p, _u64p, _u64p, _u64p, _f64p, _f64p]
    reps = np.ascontiguousarray(reps, dtype=np.uint64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(rows.size)
    rc = L.ora_matvec_rows(group.h, 8, reps.size, _p(reps, _u64p), rows.size, _p(rows, _i64p),
                           *diag.args(with_x=False), *off.args(), _p(x, _f64p), _p(out, _f64p))
    if rc != 0:
        raise ValueError("a state is not in the basis")
    return out
