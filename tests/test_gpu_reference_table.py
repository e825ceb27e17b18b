"""The reference's engine table (ls_chpl_kernels, kernels/lattice_symmetries_types.h:170-180,
filled by ls_chpl_init as chapel/src/LatticeSymmetries.chpl:17-32 does) called through its
function pointers, and ls_hs_unchecked_set_representatives (kernels/reference.c:196-211)."""
import ctypes as C
import sys

import numpy as np
import pytest

import lsamd
from lsamd import _abi, models
from lsamd.expr import Term
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_ARR = C.POINTER(_abi.chpl_external_array)
_OP = C.POINTER(_abi.ls_hs_operator)
_U64P = C.POINTER(C.c_uint64)


class ChplKernels(C.Structure):
    _fields_ = [
        ("enumerate_states", C.CFUNCTYPE(None, C.POINTER(_abi.ls_hs_basis), C.c_uint64, C.c_uint64, _ARR)),
        ("operator_apply_off_diag", C.CFUNCTYPE(None, _OP, C.c_int64, _U64P, _ARR, _ARR, _ARR, C.c_int64)),
        ("operator_apply_diag", C.CFUNCTYPE(None, _OP, C.c_int64, _U64P, _ARR, C.c_int64)),
        ("matrix_vector_product", C.CFUNCTYPE(None, _OP, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))),
    ]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def table():
    lib = lsamd.load()
    lib.ls_chpl_init()
    p = lib.ls_hs_internal_get_chpl_kernels()
    return C.cast(p, C.POINTER(ChplKernels)).contents


def external(arr, dtype, count):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    ptr = C.cast(arr.elts, C.POINTER(np.ctypeslib.as_ctypes_type(dtype)))
    return np.ctypeslib.as_array(ptr, shape=(count,)).copy()


def golden_operator(golden, name):
    def terms(key):
        v = golden[f"{name}_{key}_v"]
        return [Term(complex(v[i]), *(int(golden[f"{name}_{key}_{f}"][i]) for f in "mlrxs")) for i in range(v.size)]

    nbits = int(golden[f"{name}_nbits"][0])
    b = lsamd.SpinBasis(nbits)
    op = lsamd.Operator.__new__(lsamd.Operator)
    op.basis = b
    op._diag = lsamd._TermArrays(terms("diag"), nbits)
    op._off = lsamd._TermArrays(terms("off"), nbits)
    op._op = _abi.ls_hs_operator()
    op._op.basis = b.pointer
    op._op.diag_terms = C.pointer(op._diag.struct)
    op._op.off_diag_terms = C.pointer(op._off.struct)
    return b, op


@pytest.mark.parametrize("name", ["spin", "fermion"])
def test_table_term_application_matches_reference_golden(engine, golden, name):
    """operator_apply_diag / operator_apply_off_diag of the table against
    vectors written by the reference's own kernels/reference.c; the table
    entries apply the terms without x (BatchedOperator.chpl:298-357), so
    their coefficients times x_alpha are the golden coefficients with x."""
    say("table terms", name)
    k = table()
    b, op = golden_operator(golden, name)
    alphas = np.ascontiguousarray(golden[f"{name}_alphas"], dtype=np.uint64)
    xs = golden[f"{name}_xs"]
    n = alphas.size
    diag = _abi.chpl_external_array()
    k.operator_apply_diag(op.pointer, n, alphas.ctypes.data_as(_U64P), C.byref(diag), 1)
    ys = external(diag, np.float64, diag.num_elts)
    assert np.array_equal(ys, golden[f"{name}_diag_ys_nox"])
    lsamd.load().ls_hs_internal_destroy_external_array(C.byref(diag))
    betas, coeffs, offsets = _abi.chpl_external_array(), _abi.chpl_external_array(), _abi.chpl_external_array()
    k.operator_apply_off_diag(op.pointer, n, alphas.ctypes.data_as(_U64P), C.byref(betas), C.byref(coeffs),
                              C.byref(offsets), 1)
    off = external(offsets, np.int64, offsets.num_elts)
    assert np.array_equal(off, golden[f"{name}_off_offsets"])
    total = int(off[-1])
    bs = external(betas, np.uint64, total)
    cs = external(coeffs, np.float64, 2 * total).view(np.complex128)
    assert np.array_equal(bs, golden[f"{name}_off_betas"])
    rows = np.repeat(np.arange(n), np.diff(off))
    assert np.array_equal(cs * xs[rows], golden[f"{name}_off_coeffs"])
    for a in (betas, coeffs, offsets):
        lsamd.load().ls_hs_internal_destroy_external_array(C.byref(a))
    del op, b


def test_table_enumeration_and_product(engine):
    """enumerate_states and matrix_vector_product through the table on a
    symmetric sector, against the oracle."""
    n, h = 14, 7
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), 0),
                          lsamd.Symmetry(models.chain_parity(n), 1)])
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, -1)
    sb = lsamd.SpinBasis(n, h, -1, S)
    k = table()
    out = _abi.chpl_external_array()
    lo, hi = sb.min_state_estimate, sb.max_state_estimate
    k.enumerate_states(sb.pointer, lo, hi, C.byref(out))
    got = external(out, np.uint64, out.num_elts)
    lsamd.load().ls_hs_internal_destroy_external_array(C.byref(out))
    want = og.enumerate(lo, hi)
    assert np.array_equal(got, want)
    sb.build()
    expr = models.j1j2_chain(n, 1.0, 0.5)
    op = lsamd.Operator(sb, expr)
    x = np.random.default_rng(5).standard_normal(want.size)
    y = np.zeros_like(x)
    k.matrix_vector_product(op.pointer, 1, x.ctypes.data_as(C.POINTER(C.c_double)),
                            y.ctypes.data_as(C.POINTER(C.c_double)))
    d, o = expr.split()
    w = O.matvec_push(og, n, -1, want, O.Terms(d), O.Terms(o), x)
    assert np.abs(y - w).max() <= 1e-12 * np.abs(w).max()


@pytest.mark.parametrize("cache_bits", [1, 8, 22])
def test_unchecked_set_representatives(engine, cache_bits):
    """A basis whose representatives are set from outside (reference.c:196-211):
    the index and the product work on the adopted list."""
    n, h = 12, 6
    S = lsamd.Symmetries([lsamd.Symmetry(models.chain_translation(n), 3)])
    og = O.Group(n, S.elements(), S.characters.real, S.characters.imag, 0)
    reps = np.ascontiguousarray(og.enumerate((1 << h) - 1, ((1 << h) - 1) << (n - h)))
    sb = lsamd.SpinBasis(n, h, None, S)
    arr = _abi.chpl_external_array()
    arr.elts = reps.ctypes.data_as(C.c_void_p)
    arr.num_elts = reps.size
    arr.freer = None
    lsamd.load().ls_hs_unchecked_set_representatives(sb.pointer, C.byref(arr), cache_bits)
    assert sb.number_states == reps.size
    assert np.array_equal(sb.index(reps), np.arange(reps.size))
    missing = np.array([(1 << h) - 1 + (1 << (n - 1))], dtype=np.uint64)  # weight 7: never a representative
    assert sb.index(missing)[0] == -1
    expr = models.heisenberg_chain(n, 1.0, 0.25)
    op = lsamd.Operator(sb, expr)
    x = np.random.default_rng(cache_bits).standard_normal(reps.size)
    d, o = expr.split()
    w = O.matvec_push(og, n, 0, reps, O.Terms(d), O.Terms(o), x)
    assert np.abs(op.matvec(x) - w).max() <= 1e-12 * np.abs(w).max()
    del op
    sb.__del__()
    del reps
