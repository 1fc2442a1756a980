"""GPU parity of the 2D argon LFA case (upwind fluxes): every operator on the
path and whole Evolve steps, HIP library vs the oracle on identical inputs.
Bar: 1e-10 relative per field (FP64)."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs, field_rel_err, worst_field, efield_rel_err

pytestmark = pytest.mark.gpu

CASE = "argon2d"
NS = 4
EEN, ETEMP, POT, EFX, EFY, EFZ, REF = NS, NS + 1, NS + 2, NS + 3, NS + 4, NS + 5, NS + 9
TOL = 1e-10


@pytest.fixture(scope="module")
def pair():
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    text = small_inputs(CASE, n=(64, 48))
    o = Oracle(CASE, text)
    g = Vidyut(CASE, inputs=text)
    return o, g


def _sync_states(o, g):
    s = o.get_state()
    g.set_state(s)
    return s


def test_init_state_matches(pair):
    o, g = pair
    g.init_state()
    sg = g.get_state()
    so = o.get_state()
    for c in range(sg.shape[0]):
        assert field_rel_err(sg[c], so[c]) <= 1e-13, c


def test_time_scales(pair):
    o, g = pair
    _sync_states(o, g)
    a = np.array(g.find_time_scales())
    b = np.array(o.find_time_scales())
    assert np.allclose(a, b, rtol=1e-13, atol=0)


def test_potential(pair):
    o, g = pair
    _sync_states(o, g)
    it_o = o.solve_potential(0.0)
    info = g.solve_potential(0.0)
    assert info.iters == it_o
    a, b = g.get_state(), o.get_state()
    err, c = worst_field(a, b, [POT, REF])
    ee = efield_rel_err(a, b, EFX, 2)
    print("potential worst field", c, err, "E", ee, "iters", it_o)
    assert err <= TOL and ee <= TOL


def test_explicit_source(pair):
    o, g = pair
    s = _sync_states(o, g)
    # give the field something to drift in: solve the potential on both
    o.solve_potential(0.0)
    g.set_state(o.get_state())
    e_o = o.update_explsrc(0, NS, 0.0, 1e-12)
    e_g = g.update_explsrc(0.0, 1e-12)
    for sp in (0, 1, 2):
        err = field_rel_err(e_g[sp], e_o[sp])
        print("expl", sp, err)
        assert err <= 1e-12


def test_reaction_source(pair):
    o, g = pair
    o.solve_potential(0.0)
    g.set_state(o.get_state())
    r_o = o.update_rxnsrc(0.0)
    r_g = g.update_rxnsrc(0.0)
    for sp in range(NS):
        assert field_rel_err(r_g[sp], r_o[sp]) <= 1e-12


@pytest.mark.parametrize("sp", [0, 1, 2])
def test_implicit_solve(pair, sp):
    o, g = pair
    o.solve_potential(0.0)
    s = o.get_state()
    g.set_state(s)
    e = o.update_explsrc(0, NS, 0.0, 1e-12)
    it_o = o.implicit_solve_scalar(sp, 0.0, 1e-12, e)
    info = g.implicit_solve_scalar(sp, 0.0, 1e-12, e)
    assert info.iters == it_o
    err = field_rel_err(g.get_state()[sp], o.get_state()[sp])
    print("implicit", sp, err, it_o)
    assert err <= TOL


def test_evolve_steps():
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    text = small_inputs(CASE, n=(64, 64))
    o = Oracle(CASE, text)
    g = Vidyut(CASE, inputs=text)
    g.set_state(o.get_state())
    for k in range(4):
        io = o.evolve(5)
        ig = g.evolve(5)
        assert abs(ig.time - io[0]) <= 1e-24
        assert ig.poisson_iters == int(io[1])
        a, b = g.get_state(), o.get_state()
        err, c = worst_field(a, b, [0, 1, 2, POT, REF])
        ee = efield_rel_err(a, b, EFX, 2)
        print("steps", 5 * (k + 1), "worst", c, err, "E", ee)
        assert err <= TOL and ee <= TOL
