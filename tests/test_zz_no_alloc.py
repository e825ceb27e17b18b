"""ls_amd_state_info_noalloc against ls_hs_state_info_halide_kernel, the
entry point it mirrors: the same representatives, characters and norms bit
for bit (NaN positions included) on two calls in a row, the second smaller,
with strided arrays, and a clear error past its maximum batch."""
import ctypes as C
import sys

import numpy as np
import pytest

import lsamd
from lsamd import _abi, models

pytestmark = pytest.mark.gpu

MAX_BATCH = 1 << 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def group_data(n, gens, sectors, si):
    S = lsamd.Symmetries([lsamd.Symmetry(g, s) for g, s in zip(gens, sectors)])
    data = lsamd.load().ls_internal_create_halide_kernel_data(S.pointer, si)
    assert data
    return S, data


def random_states(n, h, count, seed):
    rng = np.random.default_rng(seed)
    out = np.zeros(count, dtype=np.uint64)
    for i in range(count):
        bits = rng.choice(n, size=h, replace=False)
        out[i] = np.uint64(sum(1 << int(b) for b in bits))
    return out


def call(fn, xs, x_stride, out_stride, data):
    n = xs.size // x_stride
    betas = np.zeros(n * out_stride, dtype=np.uint64)
    chars = np.zeros(2 * n)
    norms = np.zeros(n)
    rc = fn(n, xs.ctypes.data_as(C.POINTER(C.c_uint64)), x_stride,
            betas.ctypes.data_as(C.POINTER(C.c_uint64)), out_stride,
            chars.ctypes.data_as(C.POINTER(_abi.ls_hs_scalar)), norms.ctypes.data_as(C.POINTER(C.c_double)), data)
    return rc, betas[::out_stride].copy(), chars, norms


def same_bits(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("case", ["square6", "chain20_k5"])
def test_noalloc_state_info_equals_the_reference_entry_point(engine, case):
    lib = lsamd.load()
    if case == "square6":
        n, h, S_data = 36, 18, group_data(36, models.square_generators(6), [0] * 5, 1)
    else:
        n, h, S_data = 20, 10, group_data(20, [models.chain_translation(20)], [5], 1)
    S, data = S_data
    try:
        for seed, count, stride in ((1, MAX_BATCH, 1), (2, 1000, 1), (3, 777, 3)):
            say("noalloc", case, count, stride)
            xs = random_states(n, h, count * stride, seed)
            ref_betas = np.zeros(count * stride, dtype=np.uint64)
            ref_chars = np.zeros(2 * count)
            ref_norms = np.zeros(count)
            lib.ls_hs_state_info_halide_kernel(count, xs.ctypes.data_as(C.POINTER(C.c_uint64)), stride,
                                               ref_betas.ctypes.data_as(C.POINTER(C.c_uint64)), stride,
                                               ref_chars.ctypes.data_as(C.POINTER(_abi.ls_hs_scalar)),
                                               ref_norms.ctypes.data_as(C.POINTER(C.c_double)), data)
            _abi.check("state_info")
            rc, betas, chars, norms = call(lib.ls_amd_state_info_noalloc, xs, stride, stride, data)
            assert rc == 0, _abi.last_error()
            assert np.array_equal(betas, ref_betas[::stride])
            assert same_bits(chars, ref_chars)
            assert same_bits(norms, ref_norms)
        # past the maximum: an error, no work
        xs = random_states(n, h, MAX_BATCH + 1, 4)
        rc, _, _, _ = call(lib.ls_amd_state_info_noalloc, xs, 1, 1, data)
        assert rc == -1 and "maximum" in _abi.last_error()
        lib.ls_amd_clear_error()
    finally:
        lib.ls_internal_destroy_halide_kernel_data(data)
