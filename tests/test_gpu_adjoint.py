"""Discrete adjoint of the argonrf2d step against central finite differences
of the GPU forward objective (time-integrated electrode current), for all 15
parameters (12 reaction prefactors, mu_e, D_e, electrode voltage amplitude).

The checked configuration is a 1 mm gap at 5 kV DC, where E/N is about
200 Td and every parameter moves the objective measurably within the run;
the solver tolerance is tightened to 1e-12 so the finite differences are
not limited by it.  Bar: every component within 1e-5 relative, measured
against the largest |dJ/dp_i * p_i|."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

pytestmark = pytest.mark.gpu

CASE = "argonrf2d"
STEPS, EVERY = 30, 10
EXTRA = {
    "geometry.prob_hi": "0.001 0.001",
    "vidyut.voltage_profile": 0,
    "vidyut.voltage_amp_hi": "5000.0 0.0",
    "vidyut.linsolve_reltol": 1e-12,
    "vidyut.linsolve_abstol": 1e-30,
}


def _sim():
    from vidyut3d_amd import Vidyut
    return Vidyut(CASE, inputs=small_inputs(CASE, n=(24, 16), extra=EXTRA))


def _J(sim, p):
    sim.set_params(p)
    sim.init_state()
    return sim.objective_forward(STEPS, EVERY)


def test_objective_matches_currents():
    sim = _sim()
    p = sim.get_params()
    J = _J(sim, p)
    c = sim.currents()
    assert J == c[0, 3] and J != 0.0


def test_adjoint_matches_finite_differences():
    sim = _sim()
    p0 = sim.get_params()
    J0 = _J(sim, p0)
    g = sim.adjoint_gradient()
    assert np.all(np.isfinite(g))
    # the state is back at the end of the forward run
    J0b = sim.currents()[0, 3]
    assert J0b == J0
    fd = np.zeros_like(g)
    for i in range(len(p0)):
        h = 1e-3 * abs(p0[i])
        pp, pm = p0.copy(), p0.copy()
        pp[i] += h
        pm[i] -= h
        fd[i] = (_J(sim, pp) - _J(sim, pm)) / (2 * h)
    scale = np.abs(g * p0)
    ref = float(np.max(np.abs(fd * p0)))
    for i in range(len(p0)):
        err = abs(g[i] - fd[i]) * abs(p0[i]) / ref
        print(f"p{i:2d} adj {g[i]: .6e} fd {fd[i]: .6e} rel-to-max {err:.2e} |g p| {scale[i]:.3e}")
    errs = np.abs(g - fd) * np.abs(p0) / ref
    assert np.all(errs <= 1e-5), errs


def test_adjoint_repeatable_and_checkpoint_independent():
    sim = _sim()
    p0 = sim.get_params()
    _J(sim, p0)
    g1 = sim.adjoint_gradient()
    g2 = sim.adjoint_gradient()
    assert np.array_equal(g1, g2)
    sim.set_params(p0)
    sim.init_state()
    sim.objective_forward(STEPS, 7)
    g3 = sim.adjoint_gradient()
    assert np.array_equal(g1, g3)


def test_autograd_function():
    import torch
    from vidyut3d_amd.adjoint import electrode_current
    sim = _sim()
    p0 = sim.get_params()
    theta = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    J = electrode_current(theta, sim, STEPS, EVERY)
    J.backward()
    g = theta.grad.numpy().copy()
    # directional finite difference along a seeded direction scaled by |p|
    rng = np.random.default_rng(20261203)
    v = rng.standard_normal(len(p0)) * np.abs(p0)
    h = 1e-3
    Jp = _J(sim, p0 + h * v)
    Jm = _J(sim, p0 - h * v)
    dfd = (Jp - Jm) / (2 * h)
    dad = float(g @ v)
    print("directional adj", dad, "fd", dfd)
    assert abs(dad - dfd) <= 1e-5 * abs(dfd)
    assert float(J) == _J(sim, p0)
