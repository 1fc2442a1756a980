"""Shared helpers of the test suite: case inputs at test sizes and the
per-field relative error used as the parity measure."""
from __future__ import annotations

import numpy as np

from vidyut3d_amd import read_inputs


def small_inputs(case: str, n=None, extra: dict | None = None) -> str:
    ov = dict(extra or {})
    if n is not None:
        ov["amr.n_cell"] = n
    return read_inputs(case, ov)


def field_rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max|a-b| / max|b| over one field (0 when both vanish)."""
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den


def worst_field(a: np.ndarray, b: np.ndarray, comps) -> tuple[float, int]:
    worst, wc = 0.0, -1
    for c in comps:
        e = field_rel_err(a[c], b[c])
        if e > worst or wc < 0:
            worst, wc = e, c
    return worst, wc


def efield_rel_err(a: np.ndarray, b: np.ndarray, efx: int, dim: int) -> float:
    """max over components of max|E_a - E_b| / max|E_b| with |E_b| the vector
    magnitude: the E field is one field, so a component that is small next to
    the others is measured against the field, not against itself."""
    mag = np.sqrt(sum(b[efx + d] ** 2 for d in range(dim)))
    den = float(np.max(mag))
    num = max(float(np.max(np.abs(a[efx + d] - b[efx + d]))) for d in range(dim))
    return num / den if den > 0 else num


class HipBuffer:
    """Caller-owned device buffer (float64), allocated with hipMalloc through
    the HIP runtime the case library already loaded in this process (so the
    device pointers are the library's own, whatever else the process has
    loaded).  Create a case context before the first buffer."""

    _rt = None

    @classmethod
    def runtime(cls):
        import ctypes
        import os
        if cls._rt is None:
            rt = ctypes.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | ctypes.RTLD_GLOBAL)
            rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            rt.hipFree.argtypes = [ctypes.c_void_p]
            rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy, rt.hipDeviceSynchronize):
                f.restype = ctypes.c_int
            cls._rt = rt
        return cls._rt

    def __init__(self, shape, fill=None):
        import ctypes
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = ctypes.c_void_p()
        assert self.runtime().hipMalloc(ctypes.byref(p), self.nbytes) == 0
        self.p = p
        if fill is not None:
            self.upload(np.broadcast_to(np.asarray(fill, dtype=np.float64), self.shape))

    @property
    def ptr(self):
        return self.p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        assert self.runtime().hipMemcpy(self.p, a.ctypes.data, self.nbytes, 1) == 0   # host to device

    def download(self):
        out = np.empty(self.shape, np.float64)
        assert self.runtime().hipDeviceSynchronize() == 0
        assert self.runtime().hipMemcpy(out.ctypes.data, self.p, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            self.runtime().hipFree(self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
