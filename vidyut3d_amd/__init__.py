"""vidyut3d_amd: MI355X-native vidyut3d per-step plasma-fluid path.

Python host mirror of the reference's Vidyut operator interface for this path
(Source/Vidyut.H:78-191): ``solve_potential``, ``update_rxnsrc``,
``update_explsrc``, ``implicit_solve_scalar``, ``find_time_scales`` and
``evolve`` call the per-case HIP library (``lib/libvd_<case>.so``) through its
C ABI (``include/vidyut3d_amd.h``).  There is no CPU fallback: loading fails
loudly if the library for a case has not been built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CASES_DIR = os.path.join(ROOT, "cases")

VD_OK = 0
_ERRS = {1: "VD_E_BADARG", 2: "VD_E_NOCONV", 3: "VD_E_HIP", 4: "VD_E_NAN", 5: "VD_E_INTERNAL", 6: "VD_E_IO"}


class VidyutError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRS.get(code, code)}: {msg}")
        self.code = code


class SolveInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("bnorm", ctypes.c_double),
                ("resid0", ctypes.c_double), ("resid", ctypes.c_double)]


class BC(ctypes.Structure):
    """vd_bc: per direction and side the inputs' bc code, a device pointer to
    the side's boundary values (or None), and Robin a, b."""
    _fields_ = [("kind", (ctypes.c_int * 2) * 3), ("value", (ctypes.c_void_p * 2) * 3),
                ("robin_a", (ctypes.c_double * 2) * 3), ("robin_b", (ctypes.c_double * 2) * 3)]


class StepInfo(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("poisson_iters", ctypes.c_int), ("species_iters", ctypes.c_int),
                ("dt_edrift", ctypes.c_double), ("dt_ediff", ctypes.c_double),
                ("dt_diel_relax", ctypes.c_double)]


_LIBS: dict[str, ctypes.CDLL] = {}

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_SIGS = {
    "vd_create": (_I, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "vd_destroy": (None, [_P]),
    "vd_last_error": (ctypes.c_char_p, [_P]),
    "vd_global_error": (ctypes.c_char_p, []),
    "vd_get_dims": (_I, [_P, _P]),
    "vd_init_state": (_I, [_P]),
    "vd_set_state": (_I, [_P, _P]),
    "vd_get_state": (_I, [_P, _P]),
    "vd_state_device_ptr": (_I, [_P, ctypes.POINTER(_P)]),
    "vd_set_time": (_I, [_P, _D]),
    "vd_time_scales": (_I, [_P, _P]),
    "vd_solve_potential": (_I, [_P, _D, _P]),
    "vd_reaction_source": (_I, [_P, _D, _P]),
    "vd_explicit_source": (_I, [_P, _D, _D, _P]),
    "vd_implicit_solve": (_I, [_P, _I, _D, _D, _P, _P]),
    "vd_attach_state": (_I, [_P, _P]),
    "vd_poisson_solve": (_I, [_P, _P, _P, _P, _D, _D, _P, _P]),
    "vd_helmholtz_solve": (_I, [_P, _I, _P, _P, _D, _D, _P, _P, _D, _D, _P, _P]),
    "vd_reaction_src": (_I, [_P, _P, _D, _P]),
    "vd_explicit_rhs": (_I, [_P, _P, _P, _I, _I, _P, _P, _D, _D, _P]),
    "vd_explicit_rhs_noalloc": (_I, [_P, _P, _P, _I, _I, _P, _P, _D, _D, _P]),
    "vd_time_scales_state": (_I, [_P, _P, _P]),
    "vd_fill_ghosts": (_I, [_P, _P, _I, _I, _I]),
    "vd_write_plotfile": (_I, [_P, ctypes.c_char_p]),
    "vd_write_checkpoint": (_I, [_P, ctypes.c_char_p]),
    "vd_read_checkpoint": (_I, [_P, ctypes.c_char_p]),
    "vd_get_step": (_I, [_P, _P, _P]),
    "vd_step": (_I, [_P, _I, _P]),
    "vd_synchronize": (_I, [_P]),
    "vd_stream": (_P, [_P]),
    "vd_kernel_timing": (_I, [_P, ctypes.c_longlong]),
    "vd_kernel_times": (_I, [_P, _P, _P, _I]),
    "vd_kernel_class_count": (_I, []),
    "vd_kernel_class_name": (ctypes.c_char_p, [_I]),
    "vd_currents": (_I, [_P, _P, _I]),
    "vd_num_params": (_I, [_P]),
    "vd_get_params": (_I, [_P, _P, _I]),
    "vd_set_params": (_I, [_P, _P, _I]),
    "vd_objective_forward": (_I, [_P, _I, _I, _P]),
    "vd_adjoint_gradient": (_I, [_P, _P, _I]),
    "vd_num_levels": (_I, [_P]),
    "vd_get_level_box": (_I, [_P, _I, _P]),
    "vd_get_level_state": (_I, [_P, _I, _P]),
    "vd_set_level_state": (_I, [_P, _I, _P]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)


def lib_path(case: str) -> str:
    return os.path.join(PKG, "lib", f"libvd_{case}.so")


def load_library(case: str) -> ctypes.CDLL:
    """Load the HIP library of a case (raises if it has not been built)."""
    if case in _LIBS:
        return _LIBS[case]
    path = lib_path(case)
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m vidyut3d_amd.build {case}`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[case] = lib
    return lib


def read_inputs(case: str, overrides: dict | None = None) -> str:
    """Inputs text of a case with key overrides (ParmParse format)."""
    with open(os.path.join(CASES_DIR, case, "inputs")) as f:
        lines = f.read().splitlines()
    overrides = dict(overrides or {})
    out = []
    for ln in lines:
        key = ln.split("#", 1)[0].split("=", 1)[0].strip()
        if key and key in overrides:
            v = overrides.pop(key)
            out.append(f"{key} = {_fmt(v)}")
        else:
            out.append(ln)
    for k, v in overrides.items():
        out.append(f"{k} = {_fmt(v)}")
    return "\n".join(out) + "\n"


def _fmt(v) -> str:
    if isinstance(v, (list, tuple)):
        return " ".join(str(x) for x in v)
    return str(v)


@dataclass
class Dims:
    nx: int
    ny: int
    nz: int
    nvar: int
    nspec: int
    dim: int
    eidx: int
    ntransported: int

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * self.nz


class Vidyut:
    """One vidyut3d solver instance on the GPU for a case directory."""

    def __init__(self, case: str, inputs: str | None = None, overrides: dict | None = None):
        self.case = case
        self.lib = load_library(case)
        text = inputs if inputs is not None else read_inputs(case, overrides)
        h = _P()
        rc = self.lib.vd_create(text.encode(), ctypes.byref(h))
        if rc != VD_OK:
            raise VidyutError(rc, self.lib.vd_global_error().decode())
        self.h = h
        d = np.zeros(8, np.int32)
        self._chk(self.lib.vd_get_dims(self.h, d.ctypes.data))
        self.dims = Dims(*[int(x) for x in d])

    def _chk(self, rc: int):
        if rc != VD_OK:
            raise VidyutError(rc, self.lib.vd_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.vd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------
    def init_state(self):
        self._chk(self.lib.vd_init_state(self.h))

    def get_state(self) -> np.ndarray:
        d = self.dims
        out = np.empty((d.nvar, d.nz, d.ny, d.nx), np.float64)
        self._chk(self.lib.vd_get_state(self.h, out.ctypes.data))
        return out

    def set_state(self, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.float64)
        assert s.size == self.dims.nvar * self.dims.ncells
        self._chk(self.lib.vd_set_state(self.h, s.ctypes.data))

    def set_time(self, t: float):
        self._chk(self.lib.vd_set_time(self.h, t))

    # -- refinement levels (phi_new[lev] of the reference) ----------------------
    def num_levels(self) -> int:
        return int(self.lib.vd_num_levels(self.h))

    def level_box(self, lev: int) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
        """(origin in global level indices, extents) of level lev's box."""
        b = np.zeros(6, np.int32)
        self._chk(self.lib.vd_get_level_box(self.h, lev, b.ctypes.data))
        return tuple(int(x) for x in b[:3]), tuple(int(x) for x in b[3:])

    def get_level_state(self, lev: int) -> np.ndarray:
        _, n = self.level_box(lev)
        out = np.empty((self.dims.nvar, n[2], n[1], n[0]), np.float64)
        self._chk(self.lib.vd_get_level_state(self.h, lev, out.ctypes.data))
        return out

    def set_level_state(self, lev: int, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.float64)
        _, n = self.level_box(lev)
        assert s.size == self.dims.nvar * n[0] * n[1] * n[2]
        self._chk(self.lib.vd_set_level_state(self.h, lev, s.ctypes.data))

    def state_device_ptr(self) -> int:
        p = _P()
        self._chk(self.lib.vd_state_device_ptr(self.h, ctypes.byref(p)))
        return p.value

    def stream(self) -> int:
        return self.lib.vd_stream(self.h)

    # -- operators (names of the reference's Vidyut members) ------------------
    def find_time_scales(self) -> tuple[float, float, float]:
        o = np.zeros(3)
        self._chk(self.lib.vd_time_scales(self.h, o.ctypes.data))
        return tuple(float(x) for x in o)

    def solve_potential(self, time: float) -> SolveInfo:
        info = SolveInfo()
        self._chk(self.lib.vd_solve_potential(self.h, time, ctypes.byref(info)))
        return info

    def update_rxnsrc(self, time: float) -> np.ndarray:
        d = self.dims
        out = np.empty((d.nspec + 1, d.nz, d.ny, d.nx))
        self._chk(self.lib.vd_reaction_source(self.h, time, out.ctypes.data))
        return out

    def update_explsrc(self, time: float, dt: float) -> np.ndarray:
        d = self.dims
        out = np.empty((d.nspec + 1, d.nz, d.ny, d.nx))
        self._chk(self.lib.vd_explicit_source(self.h, time, dt, out.ctypes.data))
        return out

    def implicit_solve_scalar(self, sp: int, time: float, dt: float, expl: np.ndarray) -> SolveInfo:
        expl = np.ascontiguousarray(expl, dtype=np.float64)
        info = SolveInfo()
        self._chk(self.lib.vd_implicit_solve(self.h, sp, time, dt, expl.ctypes.data, ctypes.byref(info)))
        return info

    def evolve(self, nsteps: int) -> StepInfo:
        info = StepInfo()
        self._chk(self.lib.vd_step(self.h, nsteps, ctypes.byref(info)))
        return info

    def synchronize(self):
        self._chk(self.lib.vd_synchronize(self.h))

    # -- per-kernel-class event timing ------------------------------------------
    # -- plotfiles and checkpoints (Source/Plot.cpp) ----------------------------
    def write_plotfile(self, path: str):
        """WritePlotFile of every level into directory `path`."""
        self._chk(self.lib.vd_write_plotfile(self.h, os.fsencode(path)))

    def write_checkpoint(self, path: str):
        """WriteCheckpointFile into directory `path`."""
        self._chk(self.lib.vd_write_checkpoint(self.h, os.fsencode(path)))

    def read_checkpoint(self, path: str):
        """ReadCheckpointFile: restart every level, istep and time."""
        self._chk(self.lib.vd_read_checkpoint(self.h, os.fsencode(path)))

    def step_count(self) -> tuple[int, float]:
        """(istep, time) after the steps taken."""
        n, t = ctypes.c_int(), ctypes.c_double()
        self._chk(self.lib.vd_get_step(self.h, ctypes.byref(n), ctypes.byref(t)))
        return n.value, t.value

    def kernel_timing(self, enable: bool = True, classes: list[str] | None = None):
        """Time kernel classes with HIP events (all classes, or the named ones)."""
        if not enable:
            mask = 0
        elif classes is None:
            mask = -1
        else:
            n = self.lib.vd_kernel_class_count()
            names = [self.lib.vd_kernel_class_name(i).decode() for i in range(n)]
            mask = 0
            for c in classes:
                mask |= 1 << names.index(c)
        self._chk(self.lib.vd_kernel_timing(self.h, mask))

    def kernel_times(self) -> dict[str, tuple[float, int]]:
        n = self.lib.vd_kernel_class_count()
        ms = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        self._chk(self.lib.vd_kernel_times(self.h, ms.ctypes.data, cnt.ctypes.data, n))
        return {self.lib.vd_kernel_class_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    # -- electrode currents and run-time parameters ---------------------------
    def currents(self, reset: bool = False) -> np.ndarray:
        """[nlocs, 4]: conduction current, displacement current and collector
        area after the last step, and the time integral of conduction +
        displacement current since the last reset."""
        out = np.zeros(40)
        n = self.lib.vd_currents(self.h, out.ctypes.data, int(reset))
        if n < 0:
            raise VidyutError(-n, "vd_currents")
        return out[:4 * n].reshape(n, 4)

    def num_params(self) -> int:
        return int(self.lib.vd_num_params(self.h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params())
        self._chk(self.lib.vd_get_params(self.h, p.ctypes.data, len(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float64)
        self._chk(self.lib.vd_set_params(self.h, p.ctypes.data, len(p)))

    def objective_forward(self, nsteps: int, checkpoint_every: int = 20) -> float:
        """Time-integrated current through the electrode of current location
        0 over nsteps steps from the current state (checkpointed for the
        adjoint)."""
        J = ctypes.c_double(0.0)
        self._chk(self.lib.vd_objective_forward(self.h, nsteps, checkpoint_every, ctypes.byref(J)))
        return J.value

    def adjoint_gradient(self) -> np.ndarray:
        """dJ/dp of the last objective_forward for the run-time parameters."""
        g = np.zeros(self.num_params())
        self._chk(self.lib.vd_adjoint_gradient(self.h, g.ctypes.data, len(g)))
        return g
