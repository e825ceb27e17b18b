"""GPU parity: the HIP engine against the oracle and the reference's golden
vectors, through the C ABI.  Integer / index outputs must be bit-exact;
floating-point products must agree to 1e-12 relative (FP64)."""
import ctypes as C
import sys

import numpy as np
import pytest

import helpers as H
import lsamd
from lsamd import _abi, models
from lsamd.expr import Term
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def symmetries(gens, sectors):
    return lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])


GROUPS = [
    # name, n, generators, sectors, spin inversion, Hamming weight
    ("chain5_k2", 5, [models.chain_translation(5)], [2], 0, 2),          # runtime-depth path
    ("chain12_k3", 12, [models.chain_translation(12)], [3], 0, 6),       # 32-bit, depth 7
    ("chain12_pi_p_m", 12, [models.chain_translation(12), models.chain_parity(12)], [6, 1], -1, 6),
    ("chain16_k0", 16, [models.chain_translation(16), models.chain_parity(16)], [0, 0], 1, 8),
    ("chain20_k5", 20, [models.chain_translation(20)], [5], 1, 10),      # complex characters
    ("chain32", 32, [models.chain_translation(32), models.chain_parity(32)], [0, 0], 1, 16),  # depth 9
    ("square4", 16, models.square_generators(4), [0, 0, 0, 0, 0], 1, 8),
    ("square6", 36, models.square_generators(6), [0, 0, 0, 0, 0], 1, 18),  # 64-bit, depth 11
    ("chain40_k7", 40, [models.chain_translation(40)], [7], -1, 20),
]


def random_weight_states(n, h, count, rng):
    out = np.zeros(count, dtype=np.uint64)
    for i in range(count):
        bits = rng.choice(n, size=h, replace=False)
        out[i] = np.uint64(sum(1 << int(b) for b in bits))
    return out


@pytest.mark.parametrize("name,n,gens,sectors,si,h", GROUPS, ids=[g[0] for g in GROUPS])
def test_state_info_and_is_representative_bit_exact(engine, name, n, gens, sectors, si, h):
    say("state_info/is_representative", name)
    rng = np.random.default_rng(abs(hash(name)) % (2 ** 32))
    S = symmetries(gens, sectors)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    xs = np.concatenate([random_weight_states(n, h, 3000, rng),
                         rng.integers(0, 1 << n, size=500, dtype=np.uint64)])
    b = lsamd.DeviceBasis(n, h, si if si else None, S)
    r, c, nrm = b.state_info(xs)
    orr, oc, onrm = og.state_info(xs)
    assert np.array_equal(r, orr)
    assert np.array_equal(c, oc)
    assert np.array_equal(nrm, onrm, equal_nan=True)
    f, fn = b.is_representative(xs)
    of, ofn = og.is_representative(xs)
    assert np.array_equal(f, of)
    assert np.array_equal(fn, ofn, equal_nan=True)
    # the same through the reference-named kernel entry points (host arrays,
    # strided input)
    sb = lsamd.SpinBasis(n, h, si if si else None, S)
    strided = np.repeat(xs, 2)
    r2, c2, n2 = np.zeros(xs.size, np.uint64), np.zeros(xs.size, np.complex128), np.zeros(xs.size)
    data = sb._kernels.state_info_data
    engine.ls_hs_state_info_halide_kernel(
        xs.size, strided.ctypes.data_as(C.POINTER(C.c_uint64)), 2, r2.ctypes.data_as(C.POINTER(C.c_uint64)), 1,
        c2.ctypes.data_as(C.POINTER(_abi.ls_hs_scalar)), n2.ctypes.data_as(C.POINTER(C.c_double)), data)
    _abi.check()
    assert np.array_equal(r2, orr) and np.array_equal(c2, oc) and np.array_equal(n2, onrm, equal_nan=True)


ENUM_SECTORS = [
    ("chain16_k0_pp", 16, 8, [models.chain_translation(16), models.chain_parity(16)], [0, 0], 1),
    ("chain20_k3", 20, 10, [models.chain_translation(20)], [3], 0),
    ("chain20_k10_pm", 20, 10, [models.chain_translation(20), models.chain_parity(20)], [10, 1], -1),
    ("chain22_k4", 22, 11, [models.chain_translation(22)], [4], 1),
    ("square4_a2", 16, 8, models.square_generators(4), [0, 0, 1, 1, 0], 0),
    ("chain24_k0", 24, 12, [models.chain_translation(24), models.chain_parity(24)], [0, 0], 1),
    ("square4_h7_pipi", 16, 7, models.square_generators(4), [2, 2, 0, 0, 0], 0),
]


@pytest.mark.parametrize("name,n,h,gens,sectors,si", ENUM_SECTORS, ids=[e[0] for e in ENUM_SECTORS])
def test_enumeration_bit_exact(engine, name, n, h, gens, sectors, si):
    say("enumeration", name)
    S = symmetries(gens, sectors)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    b = lsamd.DeviceBasis(n, h, si if si else None, S)
    count = b.build()
    want = og.enumerate(b.min_state_estimate, b.max_state_estimate)
    got = b.states
    assert count == want.size
    assert np.array_equal(got, want)
    # reference ABI path: ls_hs_build_representatives -> engine table
    sb = lsamd.SpinBasis(n, h, si if si else None, S)
    sb.build()
    assert np.array_equal(sb.states, want)


def test_enumeration_without_permutations(engine):
    for n, h, si in [(14, 7, 0), (14, 7, 1), (12, 5, 0)]:
        say("unprojected", n, h, si)
        b = lsamd.DeviceBasis(n, h, si if si else None, None)
        b.build()
        assert np.array_equal(b.states, O.enumerate_unprojected(n, h, si))
    b = lsamd.DeviceBasis(10, None, None, None)
    b.build()
    assert np.array_equal(b.states, np.arange(1 << 10, dtype=np.uint64))


@pytest.mark.parametrize("case", ["idx_a", "idx_b", "idx_c"])
def test_state_index_matches_reference_golden(engine, golden, case):
    say("index", case)
    nbits, prefix = (int(v) for v in golden[f"{case}_nbits"])
    reps = golden[f"{case}_reps"]
    q = golden[f"{case}_queries"]
    arr = _abi.chpl_external_array(reps.ctypes.data, reps.size, None)
    data = engine.ls_hs_create_state_index_binary_search_kernel_data(C.byref(arr), nbits, prefix)
    _abi.check()
    out = np.zeros(q.size, dtype=np.int64)
    engine.ls_hs_state_index_binary_search_kernel(q.size, q.ctypes.data_as(C.POINTER(C.c_uint64)), 1,
                                                  out.ctypes.data_as(C.POINTER(C.c_ssize_t)), 1, data)
    _abi.check()
    engine.ls_hs_destroy_state_index_binary_search_kernel_data(data)
    assert np.array_equal(out, golden[f"{case}_indices"])


def _golden_terms(golden, name, key):
    v = golden[f"{name}_{key}_v"]
    return [Term(complex(v[i]), *(int(golden[f"{name}_{key}_{f}"][i]) for f in "mlrxs")) for i in range(v.size)]


@pytest.mark.parametrize("name", ["spin", "fermion"])
def test_term_application_matches_reference_golden(engine, golden, name):
    say("terms", name)
    nbits = int(golden[f"{name}_nbits"][0])
    diag, off = _golden_terms(golden, name, "diag"), _golden_terms(golden, name, "off")
    b = lsamd.SpinBasis(nbits)  # only carries the bit count for the terms
    op = lsamd.Operator.__new__(lsamd.Operator)
    op.basis = b
    op._diag = lsamd._TermArrays(diag, nbits)
    op._off = lsamd._TermArrays(off, nbits)
    op._op = _abi.ls_hs_operator()
    op._op.basis = b.pointer
    op._op.diag_terms = C.pointer(op._diag.struct)
    op._op.off_diag_terms = C.pointer(op._off.struct)
    a, xs = golden[f"{name}_alphas"], golden[f"{name}_xs"]
    assert np.array_equal(op.apply_diag(a, xs), golden[f"{name}_diag_ys"])
    assert np.array_equal(op.apply_diag(a), golden[f"{name}_diag_ys_nox"])
    betas, coeffs, offsets = op.apply_off_diag(a, xs)
    assert np.array_equal(offsets, golden[f"{name}_off_offsets"])
    assert np.array_equal(betas, golden[f"{name}_off_betas"])
    assert np.array_equal(coeffs, golden[f"{name}_off_coeffs"])


MATVEC_SECTORS = [
    ("chain16_k0_pp", 16, 8, [models.chain_translation(16), models.chain_parity(16)], [0, 0], 1,
     models.heisenberg_chain(16, 1.0, 0.25)),
    ("chain20_j1j2_k0_pm", 20, 10, [models.chain_translation(20), models.chain_parity(20)], [0, 1], -1,
     models.j1j2_chain(20, 1.0, 0.5)),
    ("chain18_k9", 18, 9, [models.chain_translation(18)], [9], 1, models.j1j2_chain(18, 1.0, 0.5)),
    ("square4_j1j2", 16, 8, models.square_generators(4), [0, 0, 0, 0, 0], 1, models.j1j2_square(4, 1.0, 0.5)),
    ("chain16_k4_cplx", 16, 8, [models.chain_translation(16)], [4], 0, models.heisenberg_chain(16)),
]


@pytest.mark.parametrize("name,n,h,gens,sectors,si,expr", MATVEC_SECTORS, ids=[m[0] for m in MATVEC_SECTORS])
def test_matvec_matches_oracle(engine, name, n, h, gens, sectors, si, expr):
    say("matvec", name)
    S = symmetries(gens, sectors)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, si)
    sb = lsamd.SpinBasis(n, h, si if si else None, S)
    sb.build()
    reps = sb.states.copy()
    diag, off = expr.split()
    rng = np.random.default_rng(42)
    x = rng.standard_normal(reps.size)
    want = O.matvec_push(og, n, si, reps, O.Terms(diag), O.Terms(off), x)
    scale = np.abs(want).max()
    # reference entry point (CSR cached on first use)
    op = lsamd.Operator(sb, expr)
    y = op.matvec(x)
    assert np.abs(y - want).max() <= RTOL64 * scale
    # device API: on-the-fly pull form and the CSR cache
    db = lsamd.DeviceBasis(n, h, si if si else None, S)
    db.build()
    dop = lsamd.DeviceOperator(db, expr)
    y_otf = dop.matvec(x, "otf")
    assert np.abs(y_otf - want).max() <= RTOL64 * scale
    dop.build_csr(96)
    y_csr = dop.matvec(x, "csr")
    assert np.abs(y_csr - want).max() <= RTOL64 * scale
    # complex vectors: y = M x with the full complex matrix
    if reps.size <= 4000:
        M = O.matrix_dense(og, n, si, reps, O.Terms(diag), O.Terms(off))
        xc = x + 1j * rng.standard_normal(reps.size)
        yc = dop.matvec_c(xc)
        wc = M @ xc
        assert np.abs(yc - wc).max() <= RTOL64 * np.abs(wc).max()
    # FP32 block of 8 against the FP64 block of 8
    X = rng.standard_normal((reps.size, 8))
    Y64 = dop.matvec_block(X, 64)
    ref8 = np.stack([O.matvec_push(og, n, si, reps, O.Terms(diag), O.Terms(off), X[:, q]) for q in range(8)], 1)
    assert np.abs(Y64 - ref8).max() <= RTOL64 * np.abs(ref8).max()
    Y32 = dop.matvec_block(X, 32)
    assert np.abs(Y32 - ref8).max() <= 2e-5 * np.abs(ref8).max()


def test_every_csr_schedule_gives_the_same_product(engine):
    """Lanes per row x non-temporal stream loads x die-aware persistent grid:
    all schedules of the FP64 and FP32 block products against the oracle."""
    n, h = 20, 10
    gens = [models.chain_translation(n), models.chain_parity(n)]
    S = symmetries(gens, [0, 1])
    expr = models.j1j2_chain(n, 1.0, 0.5)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, -1)
    db = lsamd.DeviceBasis(n, h, -1, S)
    db.build()
    reps = db.states
    diag, off = expr.split()
    rng = np.random.default_rng(11)
    x = rng.standard_normal(reps.size)
    want = O.matvec_push(og, n, -1, reps, O.Terms(diag), O.Terms(off), x)
    X = rng.standard_normal((reps.size, 8))
    ref8 = np.stack([O.matvec_push(og, n, -1, reps, O.Terms(diag), O.Terms(off), X[:, q]) for q in range(8)], 1)
    # the value-table CSR (default) and the plain CSR store the same matrix
    plain = lsamd.DeviceOperator(db, expr)
    plain.set_csr_packing(False)
    plain.build_csr(96)
    assert not plain.csr_info["packed"]
    dop = lsamd.DeviceOperator(db, expr)
    dop.build_csr(96)
    info = dop.csr_info
    say("csr info", info)
    assert info["packed"] and 0 < info["table_size"] <= 256
    for a, b in zip(dop.csr_arrays(), plain.csr_arrays()):
        assert np.array_equal(a, b)
    assert np.abs(plain.matvec(x, "csr") - want).max() <= RTOL64 * np.abs(want).max()
    assert np.abs(plain.matvec_block(X, 32) - ref8).max() <= 2e-5 * np.abs(ref8).max()
    for flags in (0, 1, 2, 3, 6, 7):
        for lanes in (1, 8, 64):
            say("schedule", lanes, flags)
            dop.set_schedule(lanes, flags, lanes, flags)
            y = dop.matvec(x, "csr")
            assert np.abs(y - want).max() <= RTOL64 * np.abs(want).max()
            Y32 = dop.matvec_block(X, 32)
            assert np.abs(Y32 - ref8).max() <= 2e-5 * np.abs(ref8).max()
            Y64 = dop.matvec_block(X, 64)
            assert np.abs(Y64 - ref8).max() <= RTOL64 * np.abs(ref8).max()
    dop.autotune()
    sched = dop.schedule
    assert sched["persistent_blocks"] % 8 == 0
    assert np.abs(dop.matvec(x, "csr") - want).max() <= RTOL64 * np.abs(want).max()
    with pytest.raises(lsamd.EngineError):
        dop.set_schedule(3, 0, 8, 0)


def test_block_products_take_the_schedule_of_their_gather_width(engine):
    """Block products that gather at most 16 B of x per stored entry (FP32
    blocks of 1, 2, 4; FP64 blocks of 1, 2) run on the FP64 schedule, wider
    ones on the FP32 schedule.  The lanes per row fix the order of a row's
    sum, so with the two schedules at 1 and 64 lanes each block size must
    give the bits of the schedule it belongs to, and match the oracle."""
    n, h = 20, 10
    gens = [models.chain_translation(n), models.chain_parity(n)]
    S = symmetries(gens, [0, 1])
    expr = models.j1j2_chain(n, 1.0, 0.5)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, -1)
    db = lsamd.DeviceBasis(n, h, -1, S)
    db.build()
    reps = db.states
    diag, off = expr.split()
    rng = np.random.default_rng(17)
    X = rng.standard_normal((reps.size, 16))
    ref = np.stack([O.matvec_push(og, n, -1, reps, O.Terms(diag), O.Terms(off), X[:, q]) for q in range(16)], 1)
    dop = lsamd.DeviceOperator(db, expr)
    dop.build_csr(96)
    for prec, narrow_ks, wide_ks, tol in ((32, (1, 2, 4), (8, 16), 2e-5), (64, (1, 2), (4, 8, 16), RTOL64)):
        for k in narrow_ks + wide_ks:
            Xk = X[:, :k]
            dop.set_schedule(1, 7, 64, 0)
            mixed = dop.matvec_block(Xk, prec)
            assert np.abs(mixed - ref[:, :k]).max() <= tol * np.abs(ref[:, :k]).max(), (prec, k)
            lanes, flags = (1, 7) if k in narrow_ks else (64, 0)
            dop.set_schedule(lanes, flags, lanes, flags)
            same = dop.matvec_block(Xk, prec)
            say("block schedule", prec, k, lanes, flags)
            assert np.array_equal(mixed.view(np.uint32 if prec == 32 else np.uint64),
                                  same.view(np.uint32 if prec == 32 else np.uint64)), (prec, k)


def test_matvec_rejects_operator_breaking_symmetry(engine):
    say("error path")
    n, h = 12, 6
    S = symmetries([models.chain_translation(n)], [0])
    sb = lsamd.SpinBasis(n, h, None, S)
    sb.build()
    # a transverse field changes the Hamming weight: its betas have nonzero
    # coefficients but are not in the basis -> the reference halts with
    # "invalid index" (DistributedMatrixVector.chpl:127-135)
    from lsamd.expr import sigma_x
    expr = models.heisenberg_chain(n) + sigma_x(0) * 0.3
    op = lsamd.Operator(sb, expr)
    with pytest.raises(lsamd.EngineError, match="invalid index"):
        op.matvec(np.ones(sb.number_states))


def test_lanczos_precision_forms(engine):
    """FP64 Lanczos, FP32-product Lanczos + FP64 correction, and FP32 block
    Lanczos + FP64 correction against dense diagonalisation of the oracle's
    matrix (chain of 20, J1-J2, k = 0, P+, Z+)."""
    say("lanczos")
    n, h = 20, 10
    gens = [models.chain_translation(n), models.chain_parity(n)]
    S = symmetries(gens, [0, 0])
    expr = models.j1j2_chain(n, 1.0, 0.5)
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 1)
    db = lsamd.DeviceBasis(n, h, 1, S)
    db.build()
    reps = db.states
    diag, off = expr.split()
    M = O.matrix_dense(og, n, 1, reps, O.Terms(diag), O.Terms(off)).real
    exact = np.linalg.eigvalsh(M)
    dop = lsamd.DeviceOperator(db, expr)
    dop.build_csr(96)
    r64 = dop.lanczos("fp64", max_steps=300, k=1, tol=1e-13)
    assert abs(r64["eigenvalues"][0] - exact[0]) <= 1e-10 * abs(exact[0])
    assert r64["residual_norm"] < 1e-6
    r32 = dop.lanczos("fp32", max_steps=300, k=1, tol=1e-8, refine_steps=30)
    assert abs(r32["eigenvalues"][0] - exact[0]) <= 1e-10 * abs(exact[0])
    rb = dop.block_lanczos("fp32", block_size=8, max_blocks=12, k=4, tol=1e-7, refine_steps=30)
    assert abs(rb["eigenvalues"][0] - exact[0]) <= 1e-10 * abs(exact[0])
    # thick-restart block Lanczos: FP32 block product, then FP64 cycles
    for mode in ("fp32", "fp64"):
        be = dop.block_eigensolve(mode, block_size=8, nev=4, max_blocks=10, keep_blocks=2, max_cycles=40, tol=1e-8)
        assert be["converged"], be
        assert np.abs(be["eigenvalues"] - exact[:4]).max() <= 1e-10 * abs(exact[0]), (mode, be["eigenvalues"])
        assert be["max_residual"] < 1e-7
        if mode == "fp32":
            assert be["fp64_cycles"] >= 1 and be["cycles"] > be["fp64_cycles"]
    # the next three Ritz values of the FP32 block run, at FP32 accuracy
    assert np.allclose(rb["eigenvalues"][1:4], exact[1:4], rtol=1e-5)
    rotf = dop.lanczos("otf", max_steps=300, k=1, tol=1e-13)
    assert abs(rotf["eigenvalues"][0] - exact[0]) <= 1e-10 * abs(exact[0])


def test_reference_kernels_are_reentrant(engine):
    """The reference's Chapel layer calls state_info, is_representative and
    state_index concurrently from its tasks, each with its own buffers
    (chapel/src/DistributedMatrixVector.chpl:929-965): eight host threads
    calling the reference-named entry points at once give the serial results
    bit for bit."""
    from concurrent.futures import ThreadPoolExecutor
    n, h = 24, 12
    S = symmetries([models.chain_translation(n)], [5])  # complex characters
    sb = lsamd.SpinBasis(n, h, 1, S)
    sb.build()
    reps = sb.states.copy()
    rng = np.random.default_rng(24)
    batches = []
    for _ in range(32):
        bits = np.argsort(rng.random((3000, n)), axis=1)[:, :h].astype(np.uint64)
        states = np.bitwise_or.reduce(np.uint64(1) << bits, axis=1)
        batches.append(np.concatenate([states, reps[rng.choice(reps.size, 1000)]]))

    def run(xs):
        r, c, nrm = sb.state_info(xs)
        f, fn = sb.is_representative(xs)
        return r, c, nrm, f, fn, sb.index(xs)

    serial = [run(b) for b in batches]
    assert any((s[5] >= 0).any() for s in serial) and any((s[5] < 0).any() for s in serial)
    for rnd in range(2):
        say("concurrent round", rnd)
        with ThreadPoolExecutor(max_workers=8) as pool:
            got = list(pool.map(run, batches))
        for g, s in zip(got, serial):
            for a, b in zip(g, s):
                assert np.array_equal(a, b, equal_nan=True)
