"""The device-resident C-ABI operators on caller-owned device buffers
(hipMalloc'd by the caller through the process's HIP runtime), through ctypes: vd_reaction_src and vd_explicit_rhs
(update_rxnsrc_at_all_levels / update_explsrc_at_all_levels with the
reference's species range and side codes, Source/ScalarSolve.cpp:81-216)
against the oracle, and vd_attach_state (the context advancing a caller's
buffer in place) against the context's own buffer.  Bars: 1e-10 relative per
field (explicit source 1e-11 as the member-level test), bitwise for the
attached state."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests._common import HipBuffer, small_inputs, field_rel_err

pytestmark = pytest.mark.gpu


def _i3(v):
    a = (ctypes.c_int * 3)(*([int(x) for x in v] + [0] * (3 - len(v))))
    return a


def _pair(case, n):
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    text = small_inputs(case, n=n)
    o = Oracle(case, text)
    g = Vidyut(case, inputs=text)
    o.evolve(2)   # fields and charge separation
    return o, g


def _kv(text):
    out = {}
    for ln in text.splitlines():
        ln = ln.split("#", 1)[0]
        if "=" in ln:
            k, v = ln.split("=", 1)
            out[k.strip()] = [int(x) for x in v.split()] if "_bc_" in k else v.strip()
    return out


# (case, grid, species groups: (start, num, bc key prefix))
GROUPS = {
    "air3d": ((24, 20, 32), [(0, 1, "eden"), (1, 6, "ion"), (7, 1, "neutral")]),
    "argon2d": ((64, 48), [(0, 1, "eden"), (1, 1, "ion"), (2, 1, "neutral")]),
}


@pytest.mark.parametrize("case", sorted(GROUPS))
@pytest.mark.parametrize("fused", [False, True])
def test_explicit_rhs_matches_oracle(case, fused):
    n, groups = GROUPS[case]
    o, g = _pair(case, n)
    s = o.get_state()
    d = g.dims
    dev_s = HipBuffer(s.shape, s)
    rxn = HipBuffer((d.nspec + 1,) + s.shape[1:], np.nan)
    out = HipBuffer(rxn.shape, np.nan)
    g._chk(g.lib.vd_reaction_src(g.h, dev_s.ptr, 0.0, rxn.ptr))
    g.synchronize()
    r_o = o.update_rxnsrc(0.0)
    r_g = rxn.download()
    for sp in range(d.nspec + 1):
        assert field_rel_err(r_g[sp], r_o[sp]) <= 1e-12, sp
    inputs = _kv(small_inputs(case, n=n))
    dt = float(inputs["vidyut.dt"])
    for start, num, key in groups:
        lo, hi = inputs[f"vidyut.{key}_bc_lo"], inputs[f"vidyut.{key}_bc_hi"]
        rx = None if fused else rxn.ptr
        g._chk(g.lib.vd_explicit_rhs(g.h, dev_s.ptr, rx, start, num, _i3(lo), _i3(hi), 0.0, dt, out.ptr))
        g.synchronize()
        e_g = out.download()
        e_o = o.update_explsrc(start, num, 0.0, dt)
        for sp in range(d.nspec + 1):
            if start <= sp < start + num:
                err = field_rel_err(e_g[sp], e_o[sp])
                print(case, "fused" if fused else "rxn input", "species", sp, err)
                assert err <= 1e-11, (sp, err)
            else:
                assert not np.any(e_g[sp]), sp


def test_explicit_rhs_rejects_bad_ranges():
    from vidyut3d_amd import VidyutError
    o, g = _pair("argon2d", (32, 32))
    s = o.get_state()
    dev_s = HipBuffer(s.shape, s)
    out = HipBuffer((5, 1, 32, 32), 0.0)
    p = dev_s.ptr
    q = out.ptr
    for start, num, lo, hi in [(3, 3, (2, 0), (2, 0)), (0, 1, (2, 1), (2, 0)), (0, 1, (0, 0), (2, 0)),
                               (-1, 1, (2, 0), (2, 0))]:
        rc = g.lib.vd_explicit_rhs(g.h, p, None, start, num, _i3(lo), _i3(hi), 0.0, 1e-12, q)
        assert rc == 1, (start, num, lo, hi, rc)          # VD_E_BADARG
    with pytest.raises(VidyutError):
        g._chk(g.lib.vd_explicit_rhs(g.h, p, None, 0, 1, None, None, 0.0, 1e-12, q))


@pytest.mark.parametrize("case,n", [("argon2d", (64, 48)), ("air3d", (24, 20, 32))])
def test_attached_state_is_advanced_in_place(case, n):
    from vidyut3d_amd import Vidyut

    text = small_inputs(case, n=n)
    own = Vidyut(case, inputs=text)
    own.init_state()
    s0 = own.get_state()
    ext = Vidyut(case, inputs=text)
    buf = HipBuffer(s0.shape, s0)
    ext._chk(ext.lib.vd_attach_state(ext.h, buf.ptr))
    assert ext.state_device_ptr() == buf.ptr.value
    own.evolve(3)
    ext.evolve(3)
    a = buf.download()                # the caller's buffer holds the result
    assert np.array_equal(a, own.get_state())
    assert np.array_equal(ext.get_state(), a)
    # detaching returns to the context's own (untouched) buffer
    ext._chk(ext.lib.vd_attach_state(ext.h, None))
    assert ext.state_device_ptr() != buf.ptr.value


@pytest.mark.parametrize("case,n", [("argon2d", (48, 40)), ("air3d", (24, 20, 32))])
def test_time_scales_of_a_caller_state(case, n):
    """vd_time_scales_state on a caller buffer equals find_time_scales of the
    context holding the same state (bitwise: the same kernel and reductions),
    for the initial state and for a perturbed one."""
    from vidyut3d_amd import Vidyut
    g = Vidyut(case, inputs=small_inputs(case, n=n))
    g.init_state()
    g.solve_potential(0.0)
    s0 = g.get_state()
    rng = np.random.default_rng(3)
    s1 = s0.copy()
    s1[0] *= 1.0 + rng.uniform(0.0, 0.5, s1[0].shape)        # electron density (dielectric scale)
    got = []
    for s in (s0, s1):
        buf = HipBuffer(s.shape, s)
        out = (ctypes.c_double * 3)()
        g._chk(g.lib.vd_time_scales_state(g.h, buf.ptr, out))
        g.set_state(s)
        ref = g.find_time_scales()
        print(case, list(out), list(ref))
        assert list(out) == list(ref)
        got.append(list(out))
    assert got[0][2] != got[1][2]          # the caller's buffer is what was read


@pytest.mark.parametrize("case,n,extra,ngrow", [
    ("argon2d", (20, 16), None, 3),                                   # x walls, y periodic
    ("air3d", (12, 10, 14), None, 2),                                 # walls everywhere
    ("air3d", (12, 10, 14), {"geometry.is_periodic": "1 0 1"}, 3),    # mixed
])
def test_fill_ghosts_matches_wrap_and_edge_padding(case, n, extra, ngrow):
    """vd_fill_ghosts against numpy padding, axis by axis: 'wrap' on periodic
    directions, 'edge' (foextrap) on the others; components outside
    [comp0, comp0 + ncomp) and the valid cells untouched."""
    from vidyut3d_amd import Vidyut
    g = Vidyut(case, inputs=small_inputs(case, n=n, extra=extra))
    d = g.dims
    per = [int(x) for x in ((extra or {}).get("geometry.is_periodic") or
                            ("0 1" if case == "argon2d" else "0 0 0")).split()]
    dim = d.dim
    rng = np.random.default_rng(11)
    valid = rng.standard_normal((d.nvar, d.nz, d.ny, d.nx))
    gz = ngrow if dim > 2 else 0
    fab = np.full((d.nvar, d.nz + 2 * gz, d.ny + 2 * ngrow, d.nx + 2 * ngrow), np.nan)
    fab[:, gz:gz + d.nz, ngrow:ngrow + d.ny, ngrow:ngrow + d.nx] = valid
    buf = HipBuffer(fab.shape, fab)
    comp0, ncomp = 2, 3
    g._chk(g.lib.vd_fill_ghosts(g.h, buf.ptr, comp0, ncomp, ngrow))
    g.synchronize()
    out = buf.download()
    exp = fab.copy()
    for c in range(comp0, comp0 + ncomp):
        v = valid[c]
        for ax, dd in ((2, 0), (1, 1), (0, 2)):        # array axis of direction dd
            if dd >= dim:
                continue
            pad = [(0, 0)] * 3
            pad[ax] = (ngrow, ngrow)
            v = np.pad(v, pad, mode="wrap" if per[dd] else "edge")
        exp[c] = v
    assert np.array_equal(out[comp0:comp0 + ncomp], exp[comp0:comp0 + ncomp])
    # other components: valid cells kept, ghosts not written
    for c in list(range(comp0)) + list(range(comp0 + ncomp, d.nvar)):
        assert np.array_equal(out[c], fab[c], equal_nan=True)
    # argument errors
    assert g.lib.vd_fill_ghosts(g.h, buf.ptr, d.nvar - 1, 2, ngrow) == 1
    assert g.lib.vd_fill_ghosts(g.h, buf.ptr, 0, 1, 0) == 1
