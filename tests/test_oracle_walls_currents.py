"""Pins of the oracle's inhomogeneous-Neumann species walls (boundary code 3)
and electrode currents against their physical meaning, independently of the
oracle's own code.

Walls (Source/PlasmaChem.H:75-200): the code-3 boundary sets the normal
gradient so that the diffusive flux leaving the wall is
  electrons  n_e cbar_e / 4,  cbar_e = sqrt(8 k_B T_e / (pi m_e));
  ions       mu n E_d (outward part only, zero when the drift points inward);
  neutrals   n cbar / 4,        cbar = sqrt(8 k_B T_g / (pi m)).
One implicit step with a fixed wall flux Gamma is the linear system
  n/dt - div(D grad n) = n_old/dt,  with the boundary row's face flux = Gamma,
written here from that meaning (boundary rows get -Gamma/h on the right) and
solved directly; the oracle must reproduce it.  A flipped sign (gain instead
of loss), a dropped 1/4, T_g in place of T_e, or an ion flux on the wrong
wall all fail.

Currents (Source/CurrentDens.cpp:17-456): between Dirichlet plates with a
charge-free uniform plasma the potential is linear, so the conduction
current through an electrode is the ohmic sum over charged species of
q mu n E A (outward normal), and the displacement current is
eps0 (E_new - E_old)/dt A.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests._common import small_inputs

K_B = 1.380649e-23
ME = 9.10938188e-31
M_AMU = 1.66054e-27
ECHARGE = 1.60217662e-19
EPS0 = 8.854187817e-12

# air3d component map (cases/air3d/Chemistry.H): 10 species, then EEN, ETEMP,
# POT, EFX, EFY, EFZ
A_NS = 10
A_E, A_N2P, A_O2M, A_N2S = 0, 1, 5, 7
A_ETEMP, A_EFX, A_EFZ = A_NS + 1, A_NS + 3, A_NS + 5
A_MW_N2S = 28.0134
A_D = {A_E: 0.15, A_N2P: 5.0e-6, A_O2M: 5.5e-6, A_N2S: 5.0e-6}
A_MU = {A_N2P: 2.0e-4, A_O2M: -2.2e-4}


def _air3d(nz=16):
    from oracle import Oracle

    text = small_inputs("air3d", n=(4, 4, nz), extra={
        "vidyut.linsolve_reltol": 1e-13, "vidyut.linsolve_abstol": 1e-30})
    return Oracle("air3d", text)


def _one_step_direct(n_old, D, h, dt, g_lo, g_hi):
    """1D implicit diffusion step with the outward wall fluxes g_lo, g_hi:
    n/dt - d/dz(D dn/dz) = n_old/dt, flux out of the lo face g_lo and out of
    the hi face g_hi (homogeneous Neumann in the other directions)."""
    nz = len(n_old)
    main = np.full(nz, 1.0 / dt + 2 * D / h**2)
    main[0] -= D / h**2
    main[-1] -= D / h**2
    off = np.full(nz - 1, -D / h**2)
    A = sp.diags([off, main, off], [-1, 0, 1], format="csc")
    rhs = n_old / dt
    rhs[0] -= g_lo / h
    rhs[-1] -= g_hi / h
    return spla.spsolve(A, rhs)


def _wall_case(sp_id, ez, te, n0, expect_flux):
    """Uniform state with E = (0, 0, ez) and T_e = te; one implicit solve of
    species sp_id with zero explicit source; compare with the direct solve
    of the physical wall fluxes expect_flux(n0) -> (g_lo, g_hi)."""
    nz = 16
    o = _air3d(nz)
    s = o.get_state()
    for c in range(A_NS):
        s[c] = 1e14
    s[sp_id] = n0
    s[A_ETEMP] = te
    s[A_EFX:A_EFZ + 1] = 0.0
    s[A_EFZ] = ez
    o.set_state(s)
    h = 0.02 / nz
    g_lo, g_hi = expect_flux(n0)
    # dt so that the boundary cell loses about 1% in the step
    rate = max(g_lo, g_hi) / n0 / h
    dt = 0.01 / rate
    expl = np.zeros((A_NS + 1,) + s.shape[1:])
    o.implicit_solve_scalar(sp_id, 0.0, dt, expl)
    out = o.get_state()[sp_id]
    prof = _one_step_direct(np.full(nz, n0), A_D[sp_id], h, dt, g_lo, g_hi)
    # uniform in x and y: every column is the 1D profile
    assert np.max(np.abs(out - prof[:, None, None])) <= 1e-11 * n0
    # particle balance: total change = -dt (g_lo + g_hi) * wall area
    area = 0.02 * 0.02
    dN = np.sum(out - n0) * (0.02 / out.shape[2]) * (0.02 / out.shape[1]) * h
    loss = dt * (g_lo + g_hi) * area
    assert dN == pytest.approx(-loss, rel=1e-9)
    return dt


@pytest.mark.parametrize("te", [1.0e4, 4.0e4])
def test_electron_wall_thermal_flux(te):
    cbar = np.sqrt(8 * K_B * te / np.pi / ME)
    _wall_case(A_E, 0.0, te, 1e16, lambda n: (n * cbar / 4, n * cbar / 4))


@pytest.mark.parametrize("ez", [2e4, -2e4])
def test_positive_ion_wall_drift_flux(ez):
    """N2+ drifts along E: only the wall it drifts into takes mu n |E|."""
    v = A_MU[A_N2P] * ez
    _wall_case(A_N2P, ez, 2e4, 1e16,
               lambda n: (max(-v, 0.0) * n, max(v, 0.0) * n))


def test_negative_ion_wall_drift_flux():
    """O2- drifts against E (charge-signed mobility): E_z > 0 sends it to
    the z-lo wall only."""
    ez = 3e4
    v = A_MU[A_O2M] * ez
    assert v < 0
    _wall_case(A_O2M, ez, 2e4, 1e16,
               lambda n: (max(-v, 0.0) * n, max(v, 0.0) * n))


def test_neutral_wall_thermal_flux():
    """N2* leaves both walls at n cbar/4 with the gas temperature (300 K in
    cases/air3d/inputs) and its own mass, whatever T_e is."""
    cbar = np.sqrt(8 * K_B * 300.0 / np.pi / (A_MW_N2S * M_AMU))
    _wall_case(A_N2S, 0.0, 5e4, 1e16, lambda n: (n * cbar / 4, n * cbar / 4))


# ---------------------------------------------------------------------------
# Electrode currents on argonrf2d: 1 cm gap, x-hi at V(t) = 500 sin(2 pi f t),
# x-lo grounded, y periodic; current location 0 = x-hi, 1 = x-lo.
R_NS = 4
R_E, R_ARP, R_ARM, R_AR = 0, 1, 2, 3
R_MU_E, R_MU_I = -0.03, 1.0e-4


def test_conduction_and_displacement_currents():
    from oracle import Oracle

    nx, ny = 24, 6
    L = 0.01
    dt = 1e-11
    f = 13.56e6
    amp = 500.0
    text = small_inputs("argonrf2d", n=(nx, ny), extra={
        "vidyut.do_transport": 0, "vidyut.do_reactions": 0,
        "vidyut.linsolve_reltol": 1e-13, "vidyut.linsolve_abstol": 1e-30})
    o = Oracle("argonrf2d", text)
    s = o.get_state()
    n = 3e15
    s[R_E] = n
    s[R_ARP] = n
    s[R_ARM] = 0.0
    o.set_state(s)
    o.evolve(1)          # potential at t = 0: V = 0, no displacement current yet
    c1 = o.currents()
    assert abs(c1[0, 0]) <= 1e-12 and c1[0, 1] == 0.0
    area = L   # 2D: collector area per unit depth = the electrode's length
    integral = 0.0
    E_prev = 0.0
    for step in (2, 3):  # potentials at t = dt, 2 dt
        o.evolve(1)
        c = o.currents()
        out = o.get_state()
        # the plasma stays uniform: no transport, no reactions, Neumann walls
        assert np.max(np.abs(out[R_E] - n)) <= 1e-12 * n
        V = amp * np.sin(2 * np.pi * f * (step - 1) * dt)
        E = -V / L
        j_cond = ECHARGE * n * E * (R_MU_I - R_MU_E)   # ions + electrons (charge -1)
        j_disp = EPS0 * (E - E_prev) / dt
        # location 0: x-hi, outward normal +x; location 1: x-lo, outward -x
        assert c[0, 2] == pytest.approx(area, rel=1e-14)
        assert c[0, 0] == pytest.approx(j_cond * area, rel=1e-9)
        assert c[0, 1] == pytest.approx(j_disp * area, rel=1e-9)
        assert c[1, 0] == pytest.approx(-j_cond * area, rel=1e-9)
        assert c[1, 1] == pytest.approx(-j_disp * area, rel=1e-9)
        # the time integral: step 1 adds nothing, later steps dt (cond + disp)
        integral += dt * (j_cond + j_disp) * area
        assert c[0, 3] == pytest.approx(integral, rel=1e-9)
        E_prev = E
