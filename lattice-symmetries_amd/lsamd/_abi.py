"""ctypes view of liblsamd.so (include/lsamd.h, include/lattice_symmetries_types.h).

The structs mirror the reference's kernels/lattice_symmetries_types.h
(lines 27-180) member for member, so the same objects the reference's
Haskell layer builds can be handed to the engine.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_ROOT = os.path.dirname(_HERE)
REPO_ROOT = os.path.dirname(PACKAGE_ROOT)
LIBRARY_PATH = os.path.join(PACKAGE_ROOT, "liblsamd.so")
HEADER_PATHS = [
    os.path.join(REPO_ROOT, "include", "lsamd.h"),
    os.path.join(REPO_ROOT, "include", "lattice_symmetries_types.h"),
]


class chpl_external_array(C.Structure):
    _fields_ = [("elts", C.c_void_p), ("num_elts", C.c_uint64), ("freer", C.c_void_p)]


class ls_hs_scalar(C.Structure):
    _fields_ = [("real", C.c_double), ("imag", C.c_double)]


state_index_kernel_type = C.CFUNCTYPE(
    None, C.c_ssize_t, C.POINTER(C.c_uint64), C.c_ssize_t, C.POINTER(C.c_ssize_t), C.c_ssize_t, C.c_void_p
)
is_representative_kernel_type = C.CFUNCTYPE(
    None, C.c_ssize_t, C.POINTER(C.c_uint64), C.c_ssize_t, C.POINTER(C.c_uint8), C.POINTER(C.c_double), C.c_void_p
)
state_info_kernel_type = C.CFUNCTYPE(
    None,
    C.c_ssize_t,
    C.POINTER(C.c_uint64),
    C.c_ssize_t,
    C.POINTER(C.c_uint64),
    C.c_ssize_t,
    C.POINTER(ls_hs_scalar),
    C.POINTER(C.c_double),
    C.c_void_p,
)


class ls_hs_basis_kernels(C.Structure):
    _fields_ = [
        ("state_info_kernel", C.c_void_p),
        ("state_info_data", C.c_void_p),
        ("is_representative_kernel", C.c_void_p),
        ("is_representative_data", C.c_void_p),
        ("state_index_kernel", C.c_void_p),
        ("state_index_data", C.c_void_p),
    ]


class ls_hs_permutation_group(C.Structure):
    _fields_ = [
        ("refcount", C.c_int),
        ("number_bits", C.c_int),
        ("number_shifts", C.c_int),
        ("number_masks", C.c_int),
        ("masks", C.POINTER(C.c_uint64)),
        ("shifts", C.POINTER(C.c_uint64)),
        ("eigvals_re", C.POINTER(C.c_double)),
        ("eigvals_im", C.POINTER(C.c_double)),
        ("haskell_payload", C.c_void_p),
    ]


LS_HS_SPIN, LS_HS_SPINFUL_FERMION, LS_HS_SPINLESS_FERMION = 0, 1, 2


class ls_hs_basis(C.Structure):
    _fields_ = [
        ("refcount", C.c_int),
        ("number_sites", C.c_int),
        ("number_particles", C.c_int),
        ("number_up", C.c_int),
        ("particle_type", C.c_int),
        ("spin_inversion", C.c_int),
        ("state_index_is_identity", C.c_bool),
        ("requires_projection", C.c_bool),
        ("kernels", C.POINTER(ls_hs_basis_kernels)),
        ("representatives", chpl_external_array),
        ("haskell_payload", C.c_void_p),
    ]


class ls_hs_nonbranching_terms(C.Structure):
    _fields_ = [
        ("number_terms", C.c_int),
        ("number_bits", C.c_int),
        ("v", C.POINTER(ls_hs_scalar)),
        ("m", C.POINTER(C.c_uint64)),
        ("l", C.POINTER(C.c_uint64)),
        ("r", C.POINTER(C.c_uint64)),
        ("x", C.POINTER(C.c_uint64)),
        ("s", C.POINTER(C.c_uint64)),
    ]


class ls_hs_operator(C.Structure):
    _fields_ = [
        ("refcount", C.c_int),
        ("basis", C.POINTER(ls_hs_basis)),
        ("off_diag_terms", C.POINTER(ls_hs_nonbranching_terms)),
        ("diag_terms", C.POINTER(ls_hs_nonbranching_terms)),
        ("haskell_payload", C.c_void_p),
    ]


class ls_amd_lanczos_result(C.Structure):
    _fields_ = [
        ("steps", C.c_int),
        ("refine_steps", C.c_int),
        ("number_eigenvalues", C.c_int),
        ("eigenvalues", C.c_double * 8),
        ("rayleigh", C.c_double),
        ("residual_norm", C.c_double),
        ("matvec_seconds", C.c_double),
        ("matvec_calls", C.c_int64),
        ("total_seconds", C.c_double),
    ]


class ls_amd_eig_info(C.Structure):
    _fields_ = [
        ("cycles", C.c_int),
        ("fp64_cycles", C.c_int),
        ("blocks", C.c_int),
        ("converged", C.c_int),
        ("products", C.c_int64),
        ("product_seconds", C.c_double),
        ("total_seconds", C.c_double),
        ("max_residual", C.c_double),
    ]


_P = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "ls_hs_fatal_error": (None, [C.c_char_p, C.c_int, C.c_char_p]),
    "ls_hs_set_exception_handler": (None, [_P]),
    "ls_hs_error": (None, [C.c_char_p]),
    "ls_hs_internal_destroy_external_array": (None, [C.POINTER(chpl_external_array)]),
    "ls_hs_internal_read_refcount": (C.c_int, [C.POINTER(C.c_int)]),
    "ls_hs_internal_write_refcount": (None, [C.POINTER(C.c_int), C.c_int]),
    "ls_hs_internal_inc_refcount": (C.c_int, [C.POINTER(C.c_int)]),
    "ls_hs_internal_dec_refcount": (C.c_int, [C.POINTER(C.c_int)]),
    "ls_internal_operator_apply_diag_x1": (None, [C.POINTER(ls_hs_operator), C.c_ssize_t, _u64p, _f64p, _f64p]),
    "ls_internal_operator_apply_off_diag_x1": (
        None,
        [C.POINTER(ls_hs_operator), C.c_ssize_t, _u64p, _u64p, C.POINTER(ls_hs_scalar), C.POINTER(C.c_ssize_t), _f64p],
    ),
    "ls_hs_state_index": (None, [C.POINTER(ls_hs_basis), C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(C.c_ssize_t), C.c_ssize_t]),
    "ls_hs_is_representative": (None, [C.POINTER(ls_hs_basis), C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(C.c_uint8), _f64p]),
    "ls_hs_state_info": (
        None,
        [C.POINTER(ls_hs_basis), C.c_ssize_t, _u64p, C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(ls_hs_scalar), _f64p],
    ),
    "ls_hs_build_representatives": (None, [C.POINTER(ls_hs_basis), C.c_uint64, C.c_uint64]),
    "ls_hs_unchecked_set_representatives": (None, [C.POINTER(ls_hs_basis), C.POINTER(chpl_external_array), C.c_int]),
    "ls_hs_internal_get_chpl_kernels": (_P, []),
    "ls_hs_internal_set_chpl_kernels": (None, [_P]),
    "ls_hs_create_state_index_binary_search_kernel_data": (_P, [C.POINTER(chpl_external_array), C.c_int, C.c_int]),
    "ls_hs_destroy_state_index_binary_search_kernel_data": (None, [_P]),
    "ls_hs_state_index_binary_search_kernel": (None, [C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(C.c_ssize_t), C.c_ssize_t, _P]),
    "ls_internal_create_halide_kernel_data": (_P, [C.POINTER(ls_hs_permutation_group), C.c_int]),
    "ls_internal_destroy_halide_kernel_data": (None, [_P]),
    "ls_hs_is_representative_halide_kernel": (None, [C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(C.c_uint8), _f64p, _P]),
    "ls_hs_state_info_halide_kernel": (
        None,
        [C.c_ssize_t, _u64p, C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(ls_hs_scalar), _f64p, _P],
    ),
    "ls_chpl_init": (None, []),
    "ls_chpl_finalize": (None, []),
    "ls_amd_forget_operator": (None, [C.POINTER(ls_hs_operator)]),
    "ls_amd_operator_cache_entries": (C.c_int64, []),
    "ls_chpl_enumerate_representatives": (None, [C.POINTER(ls_hs_basis), C.c_uint64, C.c_uint64, C.POINTER(chpl_external_array)]),
    "ls_chpl_operator_apply_diag": (None, [C.POINTER(ls_hs_operator), C.c_int64, _u64p, C.POINTER(chpl_external_array), C.c_int64]),
    "ls_chpl_operator_apply_off_diag": (
        None,
        [C.POINTER(ls_hs_operator), C.c_int64, _u64p, C.POINTER(chpl_external_array), C.POINTER(chpl_external_array),
         C.POINTER(chpl_external_array), C.c_int64],
    ),
    "ls_chpl_matrix_vector_product": (None, [C.POINTER(ls_hs_operator), C.c_int, _f64p, _f64p]),
    # additive API
    "ls_amd_device_available": (C.c_int, []),
    "ls_amd_last_error": (C.c_char_p, []),
    "ls_amd_clear_error": (None, []),
    "ls_amd_device_malloc": (_P, [C.c_size_t]),
    "ls_amd_device_free": (None, [_P]),
    "ls_amd_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t]),
    "ls_amd_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t]),
    "ls_amd_synchronize": (C.c_int, []),
    "ls_amd_permutation_group_create": (C.POINTER(ls_hs_permutation_group), [C.c_int, C.c_int, _i32p, _i32p]),
    "ls_amd_permutation_group_destroy": (None, [C.POINTER(ls_hs_permutation_group)]),
    "ls_amd_permutation_group_elements": (C.c_int, [C.POINTER(ls_hs_permutation_group), _i32p]),
    "ls_amd_basis_create": (_P, [C.c_int, C.c_int, C.c_int, C.POINTER(ls_hs_permutation_group)]),
    "ls_amd_fermion_basis_create": (_P, [C.c_int, C.c_int, C.c_int, C.POINTER(ls_hs_permutation_group)]),
    "ls_amd_basis_destroy": (None, [_P]),
    "ls_amd_basis_set_norm_tolerance": (C.c_int, [_P, C.c_double]),
    "ls_amd_basis_min_state": (C.c_uint64, [_P]),
    "ls_amd_basis_max_state": (C.c_uint64, [_P]),
    "ls_amd_basis_build": (C.c_int64, [_P, C.c_uint64, C.c_uint64]),
    "ls_amd_basis_number_states": (C.c_int64, [_P]),
    "ls_amd_basis_number_candidates": (C.c_int64, [_P]),
    "ls_amd_basis_states_device": (_P, [_P]),
    "ls_amd_basis_copy_states": (C.c_int, [_P, _u64p]),
    "ls_amd_basis_set_states": (C.c_int, [_P, _u64p, C.c_int64]),
    "ls_amd_basis_state_info_device": (C.c_int, [_P, C.c_int64, _P, _P, _P, _P]),
    "ls_amd_basis_is_representative_device": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "ls_amd_basis_index_device": (C.c_int, [_P, C.c_int64, _P, _P]),
    "ls_amd_state_info_noalloc": (
        C.c_int,
        [C.c_ssize_t, _u64p, C.c_ssize_t, _u64p, C.c_ssize_t, C.POINTER(ls_hs_scalar), _f64p, _P],
    ),
    "ls_amd_operator_create": (_P, [_P, C.POINTER(ls_hs_nonbranching_terms), C.POINTER(ls_hs_nonbranching_terms)]),
    "ls_amd_operator_destroy": (None, [_P]),
    "ls_amd_matvec_f64": (C.c_int, [_P, _P, _P, C.c_int]),
    "ls_amd_matvec_c128": (C.c_int, [_P, _P, _P]),
    "ls_amd_operator_build_csr": (C.c_int64, [_P, C.c_int]),
    "ls_amd_operator_csr_nnz": (C.c_int64, [_P]),
    "ls_amd_operator_copy_csr": (C.c_int, [_P, _i64p, C.POINTER(C.c_uint32), _f64p, _f64p]),
    "ls_amd_csr_matvec_f64": (C.c_int, [_P, _P, _P]),
    "ls_amd_csr_matvec_f32": (C.c_int, [_P, C.c_int, _P, _P]),
    "ls_amd_csr_matvec_f64_block": (C.c_int, [_P, C.c_int, _P, _P]),
    "ls_amd_operator_autotune": (C.c_int, [_P]),
    "ls_amd_operator_is_hermitian": (C.c_int, [_P, C.c_int]),
    "ls_amd_operator_set_csr_packing": (C.c_int, [_P, C.c_int]),
    "ls_amd_operator_set_factored": (C.c_int, [_P, C.c_int]),
    "ls_amd_operator_uses_factored": (C.c_int, [_P]),
    "ls_amd_operator_csr_info": (C.c_int, [_P, _i64p]),
    "ls_amd_operator_schedule": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "ls_amd_operator_set_schedule": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ls_amd_set_factored_enumeration": (C.c_int, [C.c_int]),
    "ls_amd_last_kernel_ms": (C.c_double, [C.c_int]),
    "ls_amd_lanczos": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, _P,
                                 C.POINTER(ls_amd_lanczos_result)]),
    "ls_amd_block_lanczos": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int,
                                       _P, C.POINTER(ls_amd_lanczos_result)]),
    "ls_amd_dense_eigh": (C.c_int, [C.c_int, _f64p, _f64p, _f64p]),
    "ls_amd_dense_eigh_hermitian": (C.c_int, [C.c_int, _f64p, _f64p, _f64p]),
    "ls_amd_model_from_yaml": (_P, [C.c_char_p]),
    "ls_amd_model_destroy": (None, [_P]),
    "ls_amd_model_basis_info": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "ls_amd_model_symmetry": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ls_amd_model_number_terms": (C.c_int64, [_P, C.c_int, C.c_int]),
    "ls_amd_model_terms": (C.c_int, [_P, C.c_int, C.c_int, _f64p, _u64p, _u64p, _u64p, _u64p, _u64p]),
    "ls_amd_model_create_basis": (_P, [_P]),
    "ls_amd_model_create_operator": (_P, [_P, C.c_int, _P]),
    "ls_amd_block_eigensolve": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_uint64, _f64p, _f64p, _P,
                                          C.POINTER(ls_amd_eig_info)]),
    "ls_amd_block_eigenvectors": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                            C.c_double, C.c_uint64, _f64p, _f64p, _P,
                                            C.POINTER(ls_amd_eig_info)]),
    "ls_amd_sector_batch_create": (_P, [C.POINTER(_P), C.c_int]),
    "ls_amd_sector_batch_destroy": (None, [_P]),
    "ls_amd_sector_batch_count": (C.c_int, [_P]),
    "ls_amd_sector_batch_number_states": (C.c_int64, [_P]),
    "ls_amd_sector_batch_nnz": (C.c_int64, [_P]),
    "ls_amd_sector_batch_offsets": (C.c_int, [_P, _i64p]),
    "ls_amd_sector_batch_set_schedule": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "ls_amd_sector_batch_get_schedule": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                   C.POINTER(C.c_int)]),
    "ls_amd_sector_batch_autotune": (C.c_int, [_P, C.c_int]),
    "ls_amd_sector_batch_matvec_f32": (C.c_int, [_P, C.c_int, _P, _P]),
    "ls_amd_sector_batch_matvec_f64": (C.c_int, [_P, C.c_int, _P, _P]),
    "ls_amd_sector_batch_eigensolve": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                                 C.c_double, C.c_uint64, _f64p, _f64p,
                                                 C.POINTER(ls_amd_eig_info)]),
}


def declared_functions() -> list[str]:
    """Function names declared in the public headers."""
    names = []
    for path in HEADER_PATHS:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith(("ls_hs_", "ls_internal_", "ls_chpl_", "ls_amd_")) and not name.endswith("_type"):
                # skip function-pointer members "(*name)" and typedef names
                start = m.start()
                if text[max(0, start - 2):start].endswith("(*"):
                    continue
                names.append(name)
    return sorted(set(names))


_lib = None
_lock = threading.Lock()
_errors: list[str] = []
_handler_ref = None
HANDLER_TYPE = C.CFUNCTYPE(None, C.c_char_p)


class EngineError(RuntimeError):
    pass


def _record_error(message):
    _errors.append(message.decode("utf-8", "replace") if message else "")


def load(path: str | None = None):
    """Loads liblsamd.so (fails loudly when it has not been built) and
    installs an error handler that records messages instead of aborting."""
    global _lib, _handler_ref
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIBRARY_PATH
        if not os.path.exists(p):
            raise EngineError(f"{p} does not exist: build it with `make -C lattice-symmetries_amd`")
        lib = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _handler_ref = HANDLER_TYPE(_record_error)
        lib.ls_hs_set_exception_handler(C.cast(_handler_ref, C.c_void_p))
        _lib = lib
        return lib


def take_errors() -> list[str]:
    out = list(_errors)
    _errors.clear()
    return out


def check(what: str = ""):
    """Raises EngineError when the engine reported an error since the last check."""
    errs = take_errors()
    if errs:
        raise EngineError((what + ": " if what else "") + "; ".join(errs))


def last_error() -> str:
    return (load().ls_amd_last_error() or b"").decode()
