"""Model files in the reference's YAML format (chapel/data/<model>.yaml),
read by the engine's C++ front-end (ls_amd_model_*)."""
import ctypes as C
from typing import List, Optional

import numpy as np

from . import _abi
from ._abi import load
from .expr import Expr, Term

PARTICLES = ("spin", "spinless-fermion", "spinful-fermion")


class Model:
    """Basis header, symmetries and operators of a model file."""

    def __init__(self, text: str):
        lib = load()
        self._h = lib.ls_amd_model_from_yaml(text.encode("utf-8"))
        if not self._h:
            raise ValueError(_abi.last_error())
        info = (C.c_int * 10)()
        if lib.ls_amd_model_basis_info(self._h, info) != 0:
            raise ValueError(_abi.last_error())
        self.particle = PARTICLES[info[0]]
        self.number_sites = info[1]
        self.hamming_weight = None if info[2] < 0 else info[2]
        self.spin_inversion = None if info[3] == 0 else info[3]
        self.number_up = None if info[4] < 0 else info[4]
        self.number_down = None if info[5] < 0 else info[5]
        self.number_particles = None if info[6] < 0 else info[6]
        self.has_hamiltonian = bool(info[8])
        self.number_observables = info[9]
        self.symmetries = []
        for k in range(info[7]):
            perm = np.zeros(self.number_sites, dtype=np.int32)
            sector = C.c_int32()
            if lib.ls_amd_model_symmetry(self._h, k, perm.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(sector)) != 0:
                raise ValueError(_abi.last_error())
            self.symmetries.append((perm, int(sector.value)))

    @staticmethod
    def from_file(path: str) -> "Model":
        with open(path, "r", encoding="utf-8") as f:
            return Model(f.read())

    def __del__(self):
        if getattr(self, "_h", None) and _abi._lib is not None:
            _abi._lib.ls_amd_model_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def terms(self, which: int = -1, diagonal: bool = False) -> List[Term]:
        """Non-branching terms of the hamiltonian (which = -1) or observable k."""
        lib = load()
        n = lib.ls_amd_model_number_terms(self._h, which, int(diagonal))
        if n < 0:
            raise ValueError(_abi.last_error())
        v = np.zeros(2 * max(n, 1))
        arrs = [np.zeros(max(n, 1), dtype=np.uint64) for _ in range(5)]
        u64 = C.POINTER(C.c_uint64)
        rc = lib.ls_amd_model_terms(self._h, which, int(diagonal), v.ctypes.data_as(C.POINTER(C.c_double)),
                                    *[a.ctypes.data_as(u64) for a in arrs])
        if rc != 0:
            raise ValueError(_abi.last_error())
        m, l, r, x, s = arrs
        return [Term(complex(v[2 * k], v[2 * k + 1]), int(m[k]), int(l[k]), int(r[k]), int(x[k]), int(s[k]))
                for k in range(n)]

    def expr(self, which: int = -1) -> Expr:
        out = Expr()
        for t in self.terms(which, True) + self.terms(which, False):
            out = out + Expr.of(t)
        return out

    def device_basis(self):
        """Device basis of the model (ls_amd_model_create_basis), not yet built."""
        from . import DeviceBasis, EngineError
        h = load().ls_amd_model_create_basis(self._h)
        if not h:
            raise EngineError(_abi.last_error())
        return DeviceBasis._adopt(h, self)

    def device_operator(self, basis, which: int = -1):
        from . import DeviceOperator, EngineError
        h = load().ls_amd_model_create_operator(self._h, which, basis.handle)
        if not h:
            raise EngineError(_abi.last_error())
        return DeviceOperator._adopt(h, basis)
