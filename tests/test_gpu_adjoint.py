"""Discrete adjoint of the argonrf2d step against central finite differences
of the objective (time-integrated electrode current) computed with the GPU
forward and with the oracle's CPU forward, for all 15 parameters (12
reaction prefactors, mu_e, D_e, electrode voltage amplitude).

Two workloads: a 1 mm gap at 5 kV DC, where E/N is about 200 Td and every
parameter moves the objective within 30 steps; and the RF workload at a
reduced grid (1 cm gap, 500 V at 13.56 MHz, 24^2 cells, 120 steps of
1e-11 s, checkpoints every 20), where only mu_e, D_e and the voltage move
it measurably.

Error measure per component i, in units of the objective:
  err_i = |g_i - fd_i| |p_i| / max(|fd_i p_i|, floor |J|)  <=  1e-5.
The objective carries about 1e-13 of relative noise (solver tolerance
1e-13, FP64 rounding over the run), and a central difference divides it by
2h: about 5e-7 |J| per unit log p_i at h = 1e-6 |p_i| and 5e-9 |J| at
h = 1e-4 |p_i| (measured on the box: up to 4e-7 and 2.6e-9).  The floors
are set to that noise over the 1e-5 bar: 0.1 at h = 1e-6, the step the
configuration specifies, and 1e-3 at h = 1e-4 (truncation O(h^2)).  So every
component with a sensitivity above 0.1% of J is checked to 1e-5 of its own
value, and the others to 1e-8 |J| absolute."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

pytestmark = pytest.mark.gpu

CASE = "argonrf2d"
STEPS, EVERY = 30, 10
TOLS = {"vidyut.linsolve_reltol": 1e-13, "vidyut.linsolve_abstol": 1e-30}
EXTRA = {
    "geometry.prob_hi": "0.001 0.001",
    "vidyut.voltage_profile": 0,
    "vidyut.voltage_amp_hi": "5000.0 0.0",
    **TOLS,
}
RF = {"n": (24, 24), "steps": 120, "every": 20, "extra": dict(TOLS)}
DC = {"n": (24, 16), "steps": STEPS, "every": EVERY, "extra": EXTRA}
FLOORS = {1e-6: 1e-1, 1e-4: 1e-3}


def _sim(w=None):
    from vidyut3d_amd import Vidyut
    w = w or DC
    return Vidyut(CASE, inputs=small_inputs(CASE, n=w["n"], extra=w["extra"]))


def _J(sim, p, w=None):
    w = w or DC
    sim.set_params(p)
    sim.init_state()
    return sim.objective_forward(w["steps"], w["every"])


def _J_oracle(p, w):
    from oracle import Oracle
    o = Oracle(CASE, small_inputs(CASE, n=w["n"], extra=w["extra"]))
    o.set_params(p)
    o.evolve(w["steps"])
    return float(o.currents()[0, 3])


def test_objective_matches_currents():
    sim = _sim()
    p = sim.get_params()
    J = _J(sim, p)
    c = sim.currents()
    assert J == c[0, 3] and J != 0.0


def _check_fd(w, forward):
    sim = _sim(w)
    p0 = sim.get_params()
    J0 = _J(sim, p0, w)
    g = sim.adjoint_gradient()
    assert np.all(np.isfinite(g))
    # the state is back at the end of the forward run
    assert sim.currents()[0, 3] == J0
    Jf = (lambda p: _J(sim, p, w)) if forward == "gpu" else (lambda p: _J_oracle(p, w))
    if forward == "oracle":
        # the two forwards agree before their differences are compared
        assert abs(Jf(p0) - J0) <= 1e-10 * abs(J0)
    worst = {}
    for rel, floor in FLOORS.items():
        errs = np.zeros(len(p0))
        for i in range(len(p0)):
            h = rel * abs(p0[i])
            pp, pm = p0.copy(), p0.copy()
            pp[i] += h
            pm[i] -= h
            fd = (Jf(pp) - Jf(pm)) / (2 * h)
            errs[i] = abs(g[i] - fd) * abs(p0[i]) / max(abs(fd * p0[i]), floor * abs(J0))
            print(f"{forward} h={rel:g} p{i:2d} adj {g[i]: .9e} fd {fd: .9e} "
                  f"sens {g[i] * p0[i] / J0: .3e} err {errs[i]:.2e}", flush=True)
        worst[rel] = float(errs.max())
        assert np.all(errs <= 1e-5), (rel, errs)
    return worst


@pytest.mark.parametrize("forward", ["gpu", "oracle"])
def test_adjoint_matches_finite_differences_dc(forward):
    print(_check_fd(DC, forward))


@pytest.mark.parametrize("forward", ["gpu", "oracle"])
def test_adjoint_matches_finite_differences_rf(forward):
    print(_check_fd(RF, forward))


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
