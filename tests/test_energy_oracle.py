"""Closed-form pins of the oracle's electron Joule heating
(Source/ElecEnergySource.cpp:72-294) on the 3D argon LMEA case.

With no space charge and Dirichlet x-sides held at 0 V and V, the potential
is linear and E = -V/L at every face.  With n_e linear in x (slope g) and
uniform Te, the face electron current is q (mu n_f E - D g) and the cell
heating 0.5 * sum over its two x-faces of current * E is
    J = q E (mu n_c E - D g)
(n_c the cell value; y and z faces carry no field).  One step changes the
gradient of a linear profile only next to the walls (drift shifts it
uniformly, diffusion of a linear profile vanishes), so interior cells hold
J to the solver tolerance (the potential and species solves stop at 1e-10
relative residual, which leaves about 1e-8 relative in E^2).  A dropped
factor 0.5, a sign error or a missing diffusion term (0.5% of J at the
gradient used) each fails this."""
from __future__ import annotations

import numpy as np

ECHARGE = 1.60217662e-19
NS = 6
EEN, ETEMP, EJH = NS, NS + 1, NS + 6


def _run(V: float, slope: float):
    from oracle import Oracle
    from vidyut3d_amd import read_inputs

    n = (24, 8, 8)
    text = read_inputs("lmea3d", {"amr.n_cell": "24 8 8", "vidyut.voltage_profile": 0,
                                  "vidyut.voltage_amp_hi": f"{V} 0.0 0.0", "vidyut.do_spacechrg": 0,
                                  "vidyut.do_reactions": 0})
    o = Oracle("lmea3d", text)
    s = o.get_state()
    L = 0.02
    x = (np.arange(n[0]) + 0.5) * L / n[0]
    n0 = 1e15
    ne = n0 * (1.0 + slope * (x - 0.5 * L) / L)
    s[0] = ne[None, None, :]
    s[1] = s[0]
    s[ETEMP] = 11604.5
    s[EEN] = 1.5 * 1.380649e-23 * s[0] * 11604.5
    o.set_state(s)
    o.evolve(1)
    out = o.get_state()
    E = -V / L
    mu, D, q = -0.03, 0.1, -ECHARGE
    g = n0 * slope / L
    J = q * E * (mu * ne * E - D * g)
    return out[EJH][:, :, 3:-3], np.broadcast_to(J[None, None, 3:-3], out[EJH][:, :, 3:-3].shape)


def test_joule_uniform_field_uniform_density():
    got, want = _run(300.0, 0.0)
    assert np.all(want > 0)   # electrons heated by the field
    assert np.max(np.abs(got - want)) <= 1e-7 * np.max(np.abs(want))


def test_joule_with_density_gradient():
    got, want = _run(300.0, 0.5)
    assert np.max(np.abs(got - want)) <= 1e-7 * np.max(np.abs(want))
    # the diffusion term is resolved: the drift-only value is far off
    E, q = -300.0 / 0.02, -ECHARGE
    drift_only = want + q * E * 0.1 * (1e15 * 0.5 / 0.02)
    assert np.max(np.abs(got - drift_only)) > 1e-3 * np.max(np.abs(want))
