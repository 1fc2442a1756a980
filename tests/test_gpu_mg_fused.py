"""The fused plane-marching smoothing passes of the 3D constant-coefficient
multigrid levels against the separate colour-sweep, residual/restriction and
prolongation kernels they replace: vidyut.mg_fused_smoothing = 2 (fused on
every eligible level, whatever its size) against 0.  Both
orders of operations are the same, so whole steps must agree to rounding
(the fused passes are written to be bitwise identical; the bar allows for
a different fused-multiply-add choice of the compiler).

Sizes leave partial tiles in x and y (64 x 16 tiles), several z chunks and
a fused second level; the cases cover Dirichlet/Neumann sides (air3d,
lmea3d), periodic x and y (air3d), uniform-coefficient species solves
(air3d), species solves where only the uniform components run fused and the
others the per-colour kernels (lmea3d, whose electron diffusivity varies),
the FP32 Poisson cycles of the mixed form, and refined levels (static
refinement), whose passes hand out the residual instead of restricting it
and smooth the already prolongated correction."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _run(case, n, extra, steps):
    from vidyut3d_amd import Vidyut

    out = []
    for fused in (2, 0):
        ov = dict(extra)
        ov["vidyut.mg_fused_smoothing"] = fused
        g = Vidyut(case, inputs=small_inputs(case, n=n, extra=ov))
        g.init_state()
        info = g.evolve(steps)
        out.append((g.get_state(), info.poisson_iters))
        g.close()
    return out


@pytest.mark.parametrize(
    "case,n,extra",
    [
        ("air3d", (72, 40, 36), {}),
        ("air3d", (80, 48, 40), {"geometry.is_periodic": "1 1 0"}),
        ("air3d", (72, 40, 36), {"vidyut.mixed_precision": 1}),
        ("lmea3d", (72, 40, 36), {}),
    ],
)
def test_fused_passes_match_separate_kernels(case, n, extra):
    (a, ia), (b, ib) = _run(case, n, extra, 3)
    assert ia == ib
    worst = 0.0
    same = 0
    for c in range(a.shape[0]):
        den = float(np.max(np.abs(b[c])))
        d = float(np.max(np.abs(a[c] - b[c])))
        same += int(d == 0.0)
        worst = max(worst, d / den if den > 0 else d)
    print(case, n, extra, "fields identical", same, "of", a.shape[0], "worst rel", worst)
    assert worst <= TOL


@pytest.mark.parametrize(
    "case,n,levels",
    [
        ("air3d", (72, 40, 36), 1),
        ("lmea3d", (48, 40, 36), 2),
    ],
)
def test_fused_passes_on_refined_levels(case, n, levels):
    from vidyut3d_amd import Vidyut

    out = []
    for fused in (2, 0):
        ov = {"vidyut.mg_fused_smoothing": fused, "amr.max_level": levels}
        g = Vidyut(case, inputs=small_inputs(case, n=n, extra=ov))
        g.init_state()
        assert g.num_levels() == levels + 1
        info = g.evolve(3)
        out.append(([g.get_level_state(lev) for lev in range(levels + 1)], info.poisson_iters))
        g.close()
    (a, ia), (b, ib) = out
    assert ia == ib
    for lev in range(levels + 1):
        worst = 0.0
        for c in range(a[lev].shape[0]):
            den = float(np.max(np.abs(b[lev][c])))
            d = float(np.max(np.abs(a[lev][c] - b[lev][c])))
            worst = max(worst, d / den if den > 0 else d)
        print(case, n, "level", lev, "worst rel", worst)
        assert worst <= TOL, (lev, worst)
