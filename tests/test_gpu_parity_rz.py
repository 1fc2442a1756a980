"""GPU parity against the oracle for the axisymmetric (RZ) path and the
three-term photoionization, on the reference's own verification cases
compiled unchanged: Laplace_Axisym (cases/laplace_axisym; the potential
against the oracle and the reference's log(r) exact solution) and
Streamer_Axisym_Photoion (cases/streamer_photoion; whole steps with WENO5,
two correctors, the case's potential, species, velocity and photoionization
hooks, with and without the photoionization).  Bar: 1e-10 relative per
field, equal Poisson V-cycle counts."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs, worst_field, efield_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_laplace_axisym_potential():
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    n = 64
    text = small_inputs("laplace_axisym", n=(n, n))
    o = Oracle("laplace_axisym", text)
    g = Vidyut("laplace_axisym", inputs=text)
    g.set_state(o.get_state())
    it_o = o.solve_potential(0.0)
    info = g.solve_potential(0.0)
    a, b = g.get_state(), o.get_state()
    ns = o.nspec
    worst, c = worst_field(a, b, [ns + 2])
    ee = efield_rel_err(a, b, ns + 3, 2)
    r = 0.1 + (np.arange(n) + 0.5) * 0.1 / n
    exact = (20.0 * np.log(r / 0.1) + 10.0 * np.log(0.2 / r)) / np.log(2.0)
    rms = float(np.sqrt(np.mean((a[ns + 2, 0] - exact[None, :]) ** 2)))
    print("laplace_axisym: potential vs oracle", worst, "E", ee, "iters", info, it_o, "RMS vs log(r)", rms,
          flush=True)
    assert worst <= TOL and ee <= TOL
    assert rms < 3e-4
    g.close()


@pytest.mark.parametrize("photo", [1, 0])
def test_streamer_axisym_steps(photo):
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    text = small_inputs("streamer_photoion", n=(64, 48), extra={"vidyut.do_photoionization": photo})
    o = Oracle("streamer_photoion", text)
    g = Vidyut("streamer_photoion", inputs=text)
    g.set_state(o.get_state())
    ns = o.nspec
    comps = list(range(ns)) + [ns + 2, ns + 9] + ([ns + 11] if photo else [])
    for k in range(2):
        io = o.evolve(2)
        ig = g.evolve(2)
        a, b = g.get_state(), o.get_state()
        worst, c = worst_field(a, b, comps)
        ee = efield_rel_err(a, b, ns + 3, 2)
        print("streamer photo", photo, "steps", 2 * (k + 1), "worst", c, worst, "E", ee, "pois iters",
              ig.poisson_iters, int(io[1]), "photo max", float(np.max(np.abs(a[ns + 11]))), flush=True)
        assert worst <= TOL and ee <= TOL
        assert ig.poisson_iters == int(io[1])
    if photo:
        assert np.max(np.abs(g.get_state()[ns + 11])) > 0
    g.close()
