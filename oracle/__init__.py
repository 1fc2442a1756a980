"""The oracle: a CPU (C++/OpenMP) restatement of vidyut3d's per-step path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never uses it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import lib_path, build_case

_LIBS: dict[str, ctypes.CDLL] = {}
_P = ctypes.c_void_p


def load(case: str, build_if_missing: bool = True, variant: str = "") -> ctypes.CDLL:
    """The oracle library of a case; variant "fast" is the -O3 x86-64-v3
    build the CPU baseline times (the parity tests use the default)."""
    key = case + ("_" + variant if variant else "")
    if key in _LIBS:
        return _LIBS[key]
    p = lib_path(case, variant)
    if not os.path.exists(p):
        if not build_if_missing:
            raise FileNotFoundError(p)
        build_case(case, variant)
    lib = ctypes.CDLL(p)
    lib.or_create.restype = _P
    lib.or_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.or_destroy.argtypes = [_P]
    lib.or_last_error.restype = ctypes.c_char_p
    lib.or_last_error.argtypes = [_P]
    for name, args in {
        "or_dims": [_P, _P], "or_get_state": [_P, _P], "or_set_state": [_P, _P],
        "or_set_time": [_P, ctypes.c_double], "or_time_scales": [_P, _P],
        "or_solve_potential": [_P, ctypes.c_double, _P],
        "or_rxn_src": [_P, ctypes.c_double, _P],
        "or_photoionization": [_P, ctypes.c_double, ctypes.c_int, _P, _P, _P],
        "or_expl_src": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P],
        "or_implicit_solve": [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P, _P],
        "or_step": [_P, ctypes.c_int, _P],
        "or_currents": [_P, _P, ctypes.c_int],
        "or_num_params": [_P],
        "or_set_params": [_P, _P, ctypes.c_int],
        "or_num_levels": [_P],
        "or_level_box": [_P, ctypes.c_int, _P],
        "or_get_level_state": [_P, ctypes.c_int, _P],
        "or_set_level_state": [_P, ctypes.c_int, _P],
    }.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    lib.or_num_threads.restype = ctypes.c_int
    _LIBS[key] = lib
    return lib


class Oracle:
    """CPU restatement instance with the same operator names as the GPU one."""

    def __init__(self, case: str, inputs: str, variant: str = ""):
        self.lib = load(case, variant=variant)
        err = ctypes.create_string_buffer(1024)
        self.h = self.lib.or_create(inputs.encode(), err, 1024)
        if not self.h:
            raise RuntimeError(err.value.decode())
        d = np.zeros(8, np.int32)
        self.lib.or_dims(self.h, d.ctypes.data)
        self.nx, self.ny, self.nz, self.nvar, self.nspec, self.dim, self.eidx = [int(x) for x in d[:7]]

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.or_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.nvar, self.nz, self.ny, self.nx)

    def get_state(self):
        out = np.empty(self.shape)
        self._chk(self.lib.or_get_state(self.h, out.ctypes.data))
        return out

    def set_state(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        self._chk(self.lib.or_set_state(self.h, s.ctypes.data))

    def set_time(self, t):
        self._chk(self.lib.or_set_time(self.h, t))

    def find_time_scales(self):
        o = np.zeros(3)
        self._chk(self.lib.or_time_scales(self.h, o.ctypes.data))
        return tuple(float(x) for x in o)

    def solve_potential(self, t):
        it = ctypes.c_int(0)
        self._chk(self.lib.or_solve_potential(self.h, t, ctypes.byref(it)))
        return it.value

    def photoionization(self, t, sph):
        """One photoionization term on the current state: (S, rhs, a)."""
        n = self.shape[1:]
        S, r, a = (np.zeros(n) for _ in range(3))
        self._chk(self.lib.or_photoionization(self.h, t, sph, S.ctypes.data, r.ctypes.data, a.ctypes.data))
        return S, r, a

    def update_rxnsrc(self, t):
        out = np.empty((self.nspec + 1, self.nz, self.ny, self.nx))
        self._chk(self.lib.or_rxn_src(self.h, t, out.ctypes.data))
        return out

    def update_explsrc(self, startspec, numspec, t, dt):
        out = np.empty((self.nspec + 1, self.nz, self.ny, self.nx))
        self._chk(self.lib.or_expl_src(self.h, startspec, numspec, t, dt, out.ctypes.data))
        return out

    def implicit_solve_scalar(self, sp, t, dt, expl):
        it = ctypes.c_int(0)
        expl = np.ascontiguousarray(expl, dtype=np.float64)
        self._chk(self.lib.or_implicit_solve(self.h, sp, t, dt, expl.ctypes.data, ctypes.byref(it)))
        return it.value

    def evolve(self, nsteps):
        info = np.zeros(8)
        self._chk(self.lib.or_step(self.h, nsteps, info.ctypes.data))
        return info

    def num_levels(self) -> int:
        return self.lib.or_num_levels(self.h)

    def level_box(self, lev: int):
        b = np.zeros(6, np.int32)
        self._chk(self.lib.or_level_box(self.h, lev, b.ctypes.data))
        return tuple(int(x) for x in b[:3]), tuple(int(x) for x in b[3:])

    def get_level_state(self, lev: int) -> np.ndarray:
        _, n = self.level_box(lev)
        out = np.empty((self.nvar, n[2], n[1], n[0]))
        self._chk(self.lib.or_get_level_state(self.h, lev, out.ctypes.data))
        return out

    def set_level_state(self, lev: int, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.float64)
        self._chk(self.lib.or_set_level_state(self.h, lev, s.ctypes.data))

    def num_threads(self):
        return self.lib.or_num_threads()

    def currents(self, reset: bool = False) -> np.ndarray:
        """[nlocs, 4]: conduction, displacement, area of the last step and
        the time integral of conduction + displacement current."""
        out = np.zeros(4 * 10)
        n = self.lib.or_currents(self.h, out.ctypes.data, int(reset))
        return out[:4 * n].reshape(n, 4)

    def num_params(self) -> int:
        return self.lib.or_num_params(self.h)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        self._chk(self.lib.or_set_params(self.h, p.ctypes.data, len(p)))
