"""The marching WENO5 flux kernel for y and z faces (vidyut.flux_march = 1)
against the per-face kernel (0).  Both reconstruct with the same formulas;
the marching kernel groups the smoothness indicators by cell triple, so the
two agree to rounding, not bit for bit.  Compared: the explicit source
(flux divergence, reactions off) of air3d (WENO5-Z and JS, eight
transported species) after a potential solve, with walls and with periodic
x and y sides, several chunks of 32 faces per column, and in the mixed
form, whose FP32 faces the marching kernel builds from triples carrying
their own power-of-two exponents (the per-face kernel scales each five-cell
stencil).  The bars allow for the cancellation of neighbouring face fluxes
in the divergence: 1e-9 for FP64 faces, 1e-5 for FP32 faces (unit roundoff
6e-8, the two kernels rounding differently); an indexing or scaling mistake
shows up at order one."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

pytestmark = pytest.mark.gpu

TOL = 1e-9
TOL_F32 = 1e-5


@pytest.mark.parametrize(
    "n,extra",
    [
        ((40, 72, 68), {}),
        ((48, 64, 60), {"geometry.is_periodic": "1 1 0"}),
        ((40, 72, 68), {"vidyut.weno_scheme": 1}),
        ((40, 72, 68), {"vidyut.mixed_precision": 1}),
        ((48, 64, 60), {"vidyut.mixed_precision": 1, "vidyut.weno_scheme": 2, "geometry.is_periodic": "1 1 0"}),
    ],
)
def test_march_matches_per_face_kernel(n, extra):
    from vidyut3d_amd import Vidyut

    out = []
    for march in (1, 0):
        ov = dict(extra)
        ov["vidyut.flux_march"] = march
        ov["vidyut.do_reactions"] = 0
        g = Vidyut("air3d", inputs=small_inputs("air3d", n=n, extra=ov))
        g.init_state()
        g.solve_potential(0.0)
        out.append(g.update_explsrc(0.0, 5e-13))
        g.close()
    a, b = out
    worst, moved = 0.0, 0
    for c in range(a.shape[0] - 1):      # species sources (the last row is the energy);
        # species with zero initial density have no flux
        den = float(np.max(np.abs(b[c])))
        if den == 0.0:                   # a species that is not transported
            assert float(np.max(np.abs(a[c]))) == 0.0
            continue
        moved += 1
        worst = max(worst, float(np.max(np.abs(a[c] - b[c]))) / den)
    print(n, extra, "species with transport", moved, "worst rel", worst)
    assert moved >= 3 and worst <= (TOL_F32 if extra.get("vidyut.mixed_precision") else TOL)
