"""Mixed-precision form on the GPU: FP32 V-cycles as the preconditioner of
FP64 iterative refinement (residual and update in FP64) and FP32 WENO5 face
fluxes, against the FP64 form of the same library.  Bar: <= 1e-8 relative
field error (L2 and max) after several steps, with the Poisson residual
still converged to the FP64 tolerance."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs, field_rel_err, efield_rel_err

pytestmark = pytest.mark.gpu

CASE = "air3d"
NS = 10
POT, EFX, REF = NS + 2, NS + 3, NS + 9
TRANSPORTED = list(range(8))


def _pair(n):
    from vidyut3d_amd import Vidyut

    base = small_inputs(CASE, n=n)
    g64 = Vidyut(CASE, inputs=base)
    g32 = Vidyut(CASE, inputs=small_inputs(CASE, n=n, extra={"vidyut.mixed_precision": 1}))
    g64.init_state()
    g32.set_state(g64.get_state())
    return g64, g32


def test_mixed_potential_refinement_converges():
    g64, g32 = _pair((32, 24, 40))
    i64 = g64.solve_potential(0.0)
    i32 = g32.solve_potential(0.0)
    # FP64 residual of the mixed solve meets the same stopping test
    assert i32.resid <= max(1e-10, 1e-10 * max(i32.bnorm, i32.resid0)) * 1.0000001
    assert abs(i32.iters - i64.iters) <= 2
    a, b = g32.get_state(), g64.get_state()
    assert field_rel_err(a[POT], b[POT]) <= 1e-8
    assert efield_rel_err(a, b, EFX, 3) <= 1e-8


def test_mixed_steps_field_error():
    g64, g32 = _pair((32, 24, 40))
    g64.evolve(6)
    g32.evolve(6)
    a, b = g32.get_state(), g64.get_state()
    worst = 0.0
    for c in TRANSPORTED + [POT, REF]:
        mx = field_rel_err(a[c], b[c])
        l2 = float(np.linalg.norm(a[c] - b[c]) / max(np.linalg.norm(b[c]), 1e-300))
        worst = max(worst, mx, l2)
        print("field", c, "max", mx, "l2", l2)
    ee = efield_rel_err(a, b, EFX, 3)
    print("E", ee)
    assert worst <= 1e-8 and ee <= 1e-8


def test_mixed_weno_explicit_source():
    """FP32 WENO faces (power-of-two scaled) against FP64: relative error of
    the explicit source at FP32 level."""
    g64, g32 = _pair((32, 24, 40))
    g64.solve_potential(0.0)
    g32.set_state(g64.get_state())
    e64 = g64.update_explsrc(0.0, 5e-13)
    e32 = g32.update_explsrc(0.0, 5e-13)
    for sp in TRANSPORTED:
        assert np.all(np.isfinite(e32[sp])), sp
        err = field_rel_err(e32[sp], e64[sp])
        print("expl", sp, err)
        assert err <= 1e-4
