"""vd_explicit_rhs_noalloc, the entry point that allocates nothing (its face
fluxes in a file-scope device scratch array), against vd_explicit_rhs (the
context's face buffers) through the C ABI on caller-owned device buffers (hipMalloc'd by the test):
bitwise equal results (same kernels, same order) on two calls in a row, the
second on a smaller context and species range, so stale scratch contents of
the first call cannot leak into the second; and the size limit is enforced.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests._common import HipBuffer, small_inputs

pytestmark = pytest.mark.gpu


def _i3(v):
    return (ctypes.c_int * 3)(*([int(x) for x in v] + [0] * (3 - len(v))))


def _ctx(case, n):
    from vidyut3d_amd import Vidyut
    g = Vidyut(case, inputs=small_inputs(case, n=n))
    g.init_state()
    g.evolve(2)
    return g


def _both(g, start, num, lo, hi, fused):
    d = g.dims
    st = g.get_state()
    s = HipBuffer(st.shape, st)
    shape = (d.nspec + 1, d.nz, d.ny, d.nx)
    rxn = HipBuffer(shape, 0.0)
    a = HipBuffer(shape, np.nan)
    b = HipBuffer(shape, np.nan)
    g._chk(g.lib.vd_reaction_src(g.h, s.ptr, 0.0, rxn.ptr))
    rx = None if fused else rxn.ptr
    g._chk(g.lib.vd_explicit_rhs(g.h, s.ptr, rx, start, num, _i3(lo), _i3(hi), 0.0, 5e-13, a.ptr))
    g._chk(g.lib.vd_explicit_rhs_noalloc(g.h, s.ptr, rx, start, num, _i3(lo), _i3(hi), 0.0, 5e-13, b.ptr))
    g.synchronize()
    return a.download(), b.download()


def test_noalloc_matches_allocating_entry_point_twice():
    # first call: air3d 32x24x40, all eight transported species (fused
    # reactions); second: a smaller grid and one species with rxn as input
    g1 = _ctx("air3d", (32, 24, 40))
    a, b = _both(g1, 0, 8, (2, 2, 3), (2, 2, 3), True)
    print("call 1: max |a|", float(np.max(np.abs(a))))
    assert np.all(np.isfinite(a)) and np.any(a[:8])
    assert np.array_equal(a, b)
    g2 = _ctx("air3d", (16, 12, 20))
    a, b = _both(g2, 7, 1, (2, 2, 3), (2, 2, 3), False)
    print("call 2: max |a|", float(np.max(np.abs(a))))
    assert np.all(np.isfinite(a)) and np.any(a[7])
    assert np.array_equal(a, b)


def test_noalloc_rejects_inputs_beyond_the_scratch():
    # 8 species x 3 directions x ~2.2M faces > 8M doubles of scratch
    g = _ctx("air3d", (128, 128, 128))
    d = g.dims
    s = HipBuffer((d.nvar, d.nz, d.ny, d.nx), 0.0)
    out = HipBuffer((d.nspec + 1, d.nz, d.ny, d.nx), 0.0)
    rc = g.lib.vd_explicit_rhs_noalloc(g.h, s.ptr, None, 0, 8, _i3((2, 2, 3)), _i3((2, 2, 3)), 0.0, 5e-13,
                                       out.ptr)
    assert rc == 1                     # VD_E_BADARG
    assert "scratch" in g.lib.vd_last_error(g.h).decode()
