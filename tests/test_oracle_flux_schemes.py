"""Pins of the oracle's second-order WAF flux and of the three WENO5
variants (Source/compute_explicit_flux.H:16-203), from properties the
schemes' definitions fix rather than from the oracle's own formulas.

WAF with the MINBEE limiter (Toro, TVD WAF): at |c| = 1 the limiter is 1 and
the scheme is upwind, so forward Euler translates a profile exactly one
cell; for 0 < c < 1 it is TVD (no new extrema, total variation does not
grow) and second-order accurate on smooth data, where upwind is first order.

WENO5 (finite-difference form, point values treated as cell averages of the
flux's primitive): with ideal weights the reconstruction is exact for
quadratics, whose smoothness indicators are all equal, so every variant
gives the exact derivative there.  The variants differ in how weights leave
the ideal ones:
  * WENO-Z (Borges et al. 2008) uses tau5 = |beta_0 - beta_2|; on data
    mirror-symmetric about the stencil centre beta_0 = beta_2, so its weights
    are exactly ideal and the left-biased value is the linear 5th-order one,
    (-a + 14 b + 47 c)/60 for the stencil (a, b, c, b, a);
  * WENO-JS (Jiang & Shu 1996) weights deviate there (beta_1 differs);
  * WENO-M (Henrick et al. 2005) maps the JS weights through g_k, which
    fixes the ideal weights to third order, so it deviates much less than JS.
Face fluxes are recovered from the oracle's divergence by a running sum from
a face in a constant region (where every scheme gives v * n exactly).
"""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

CASE = "argon2d"
NS = 4
E_ID, ARP = 0, 1
EFX, EFY, REF = NS + 3, NS + 4, NS + 9
L = 0.01
MU_E = -0.03

PERIODIC = {"geometry.is_periodic": "1 1", "vidyut.pot_bc_lo": "0 0", "vidyut.pot_bc_hi": "0 0",
            "vidyut.eden_bc_lo": "0 0", "vidyut.eden_bc_hi": "0 0",
            "vidyut.ion_bc_lo": "0 0", "vidyut.ion_bc_hi": "0 0",
            "vidyut.neutral_bc_lo": "0 0", "vidyut.neutral_bc_hi": "0 0", "prob.noise": 0.0}


def _oracle(n, extra):
    from oracle import Oracle

    ov = dict(PERIODIC)
    ov.update(extra)
    return Oracle(CASE, small_inputs(CASE, n=(n, 4), extra=ov))


def _set_profile(o, prof, ex=-1e5):
    """Electron density prof along x, uniform drift field, no chemistry that
    matters (E/N tiny, no ions)."""
    s = o.get_state()
    s[E_ID] = prof[None, None, :]
    s[EFX] = ex
    s[EFY] = 0.0
    s[REF] = 1e-3
    s[ARP] = 0.0
    o.set_state(s)


def _div(o, prof, dt, ex=-1e5):
    _set_profile(o, prof, ex)
    return o.update_explsrc(E_ID, 1, 0.0, dt)[E_ID, 0, 0]


# ---------------------------------------------------------------------------
# WAF + MINBEE (hyp_order = 2)

def _waf_advect(n, c, nsteps, prof0):
    o = _oracle(n, {"vidyut.hyp_order": 2})
    dx = L / n
    v = MU_E * -1e5
    dt = c * dx / v
    p = prof0.copy()
    for _ in range(nsteps):
        p = p + dt * _div(o, p, dt)
    return p


def test_waf_exact_translation_at_courant_one():
    n = 40
    x = (np.arange(n) + 0.5) * L / n
    prof = 1e16 * (1.2 + np.where(np.abs(x - 0.3 * L) < 0.1 * L, 1.0, 0.0)
                   + 0.5 * np.sin(2 * np.pi * x / L))
    p = _waf_advect(n, 1.0, 3, prof)
    assert np.allclose(p, np.roll(prof, 3), rtol=1e-13, atol=0)


def test_waf_is_tvd_on_a_square_wave():
    n = 64
    x = (np.arange(n) + 0.5) * L / n
    prof = 1e16 * (1.0 + np.where(np.abs(x - 0.35 * L) < 0.15 * L, 2.0, 0.0))
    tv0 = np.sum(np.abs(np.diff(np.r_[prof, prof[:1]])))
    p = prof.copy()
    o = _oracle(n, {"vidyut.hyp_order": 2})
    dx = L / n
    dt = 0.45 * dx / (MU_E * -1e5)
    for _ in range(40):
        p = p + dt * _div(o, p, dt)
        tv = np.sum(np.abs(np.diff(np.r_[p, p[:1]])))
        assert tv <= tv0 * (1 + 1e-12)
        assert p.max() <= prof.max() * (1 + 1e-12) and p.min() >= prof.min() * (1 - 1e-12)
        tv0 = tv
    # the limiter keeps some sharpening: more cells in the plateau's upper
    # half than first-order upwind leaves after the same steps
    o1 = _oracle(n, {"vidyut.hyp_order": 1})
    q = prof.copy()
    for _ in range(40):
        q = q + dt * _div(o1, q, dt)
    assert np.sum(p > 2.5e16) > np.sum(q > 2.5e16)


def test_waf_second_order_on_smooth_data():
    errs = []
    for n in (32, 64, 128):
        x = (np.arange(n) + 0.5) * L / n
        prof = 1e16 * (2 + np.sin(2 * np.pi * x / L))
        p = _waf_advect(n, 0.5, 2 * n, prof)    # one period
        errs.append(np.mean(np.abs(p - prof)) / 1e16)
    order = np.log2(errs[1] / errs[2])
    assert order > 1.7, (errs, order)


# ---------------------------------------------------------------------------
# WENO5 variants (hyp_order = 5, weno_scheme 1 JS, 2 M, 3 Z)

@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_weno_exact_for_quadratics(scheme):
    """Non-periodic x (homogeneous Neumann); cells at least three from a wall
    see only the quadratic, so -d(n v)/dx is exact there."""
    n = 24
    ov = dict(PERIODIC)
    ov.update({"geometry.is_periodic": "0 1", "vidyut.pot_bc_lo": "1 0", "vidyut.pot_bc_hi": "1 0",
               "vidyut.eden_bc_lo": "2 0", "vidyut.eden_bc_hi": "2 0",
               "vidyut.hyp_order": 5, "vidyut.weno_scheme": scheme})
    from oracle import Oracle

    o = Oracle(CASE, small_inputs(CASE, n=(n, 4), extra=ov))
    x = (np.arange(n) + 0.5) * L / n
    prof = 1e16 * (1.0 + 0.7 * x / L + 2.3 * (x / L) ** 2)
    e = _div(o, prof, 1e-12)
    v = MU_E * -1e5
    exact = -v * 1e16 * (0.7 / L + 2 * 2.3 * x / L**2)
    inner = slice(3, n - 3)
    assert np.max(np.abs(e[inner] - exact[inner])) <= 1e-11 * np.max(np.abs(exact))


def _bump_face_value(scheme):
    """Left-biased face value at i0+1/2 over the stencil i0-2..i0+2 of a
    bump mirror-symmetric about cell i0, and the linear 5th-order value."""
    n = 40
    i0 = 17
    c0 = 1e16
    m = np.abs(np.arange(n) - i0)
    K = 12
    prof = c0 * (1.0 + np.where(m < K, 0.8 * np.cos(np.pi * m / (2 * K)) ** 2, 0.0))
    o = _oracle(n, {"vidyut.hyp_order": 5, "vidyut.weno_scheme": scheme})
    dx = L / n
    v = MU_E * -1e5
    e = _div(o, prof, 1e-12)
    # F_{i+1/2} = F_{i-1/2} - dx e_i, starting from face 0 (constant region)
    F = v * c0 - dx * np.cumsum(e)
    face = F[i0] / v        # face i0 + 1/2
    a, b, c = prof[i0 - 2], prof[i0 - 1], prof[i0]
    lin = (-a + 14 * b + 47 * c) / 60
    return face, lin, c0


def test_weno_variants_on_a_symmetric_stencil():
    fz, lin, c0 = _bump_face_value(3)
    fj, _, _ = _bump_face_value(1)
    fm, _, _ = _bump_face_value(2)
    dz, dj, dm = abs(fz - lin), abs(fj - lin), abs(fm - lin)
    # WENO-Z: ideal weights exactly (rounding of the running sum only)
    assert dz <= 1e-13 * c0, dz / c0
    # WENO-JS leaves the ideal weights visibly; WENO-M, whose mapping is
    # flat to second order at the ideal weights, far less
    assert dj >= 1e-7 * c0, dj / c0
    assert dm <= 0.05 * dj, (dm / c0, dj / c0)
    assert dm >= 1e-10 * c0, dm / c0


@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_weno_non_oscillatory_at_a_jump(scheme):
    """A step advected with forward Euler at CFL 0.3 stays within its bounds
    to a small fraction of the jump (ENO property)."""
    n = 48
    x = (np.arange(n) + 0.5) * L / n
    prof = 1e16 * (1.0 + np.where(np.abs(x - 0.4 * L) < 0.2 * L, 1.0, 0.0))
    o = _oracle(n, {"vidyut.hyp_order": 5, "vidyut.weno_scheme": scheme})
    dt = 0.3 * (L / n) / (MU_E * -1e5)
    p = prof.copy()
    for _ in range(10):
        p = p + dt * _div(o, p, dt)
    assert p.max() <= 2e16 * (1 + 0.02) and p.min() >= 1e16 * (1 - 0.02)
