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
