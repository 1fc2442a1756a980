"""argonrf2d (table chemistry with run-time parameters, RF electrode,
electrode-current tracking) on the GPU against the oracle: whole steps,
the current-density components of the state (ECUR, ICUR, DCUR), the
integrated electrode currents and their time integral, with default and
perturbed parameters."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs, field_rel_err, efield_rel_err

pytestmark = pytest.mark.gpu

CASE = "argonrf2d"
NS = 4
POT, EFX, REF = NS + 2, NS + 3, NS + 9
ECURX, ICURX, DCURX = NS + 13, NS + 16, NS + 19
N = (48, 32)


def _pair(params=None):
    from oracle import Oracle
    from vidyut3d_amd import Vidyut
    inp = small_inputs(CASE, n=N)
    g = Vidyut(CASE, inputs=inp)
    o = Oracle(CASE, inp)
    if params is not None:
        g.set_params(params)
        o.set_params(params)
    g.init_state()
    return g, o


def _check(g, o, steps):
    g.evolve(steps)
    o.evolve(steps)
    a, b = g.get_state(), o.get_state()
    for c in [0, 1, 2, POT, REF]:
        e = field_rel_err(a[c], b[c])
        assert e <= 1e-10, (c, e)
    assert efield_rel_err(a, b, EFX, 2) <= 1e-10
    for base in (ECURX, ICURX, DCURX):
        for d in range(2):
            e = field_rel_err(a[base + d], b[base + d])
            assert e <= 1e-9, (base + d, e)
    cg, co = g.currents(), o.currents()
    assert cg.shape == co.shape == (2, 4)
    for loc in range(2):
        for q in range(4):
            den = max(abs(co[loc, q]), 1e-300)
            assert abs(cg[loc, q] - co[loc, q]) <= 1e-9 * den + 1e-30, (loc, q, cg[loc, q], co[loc, q])


def test_params_roundtrip():
    from vidyut3d_amd import Vidyut
    g = Vidyut(CASE, inputs=small_inputs(CASE, n=N))
    p = g.get_params()
    assert g.num_params() == 15 and len(p) == 15
    assert np.allclose(p[:12], 1.0) and p[12] == -0.03 and p[13] == 0.1 and p[14] == 500.0
    q = p.copy()
    q[3] = 2.5
    q[14] = 420.0
    g.set_params(q)
    assert np.array_equal(g.get_params(), q)


def test_steps_and_currents_default_params():
    g, o = _pair()
    _check(g, o, 12)


def test_steps_and_currents_perturbed_params():
    p = np.array([1.3, 0.7, 1.1, 0.9, 2.0, 0.5, 1.5, 0.8, 1.2, 3.0, 0.6, 1.4, -0.025, 0.12, 650.0])
    g, o = _pair(p)
    _check(g, o, 12)


def test_contexts_keep_their_own_params():
    """Two contexts with different parameters, stepped alternately, each
    match the oracle run with its own parameters."""
    from oracle import Oracle
    from vidyut3d_amd import Vidyut
    p = np.array([1.0, 1.0, 1.0, 2.5, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -0.03, 0.1, 420.0])
    inp = small_inputs(CASE, n=N)
    g1, g2 = Vidyut(CASE, inputs=inp), Vidyut(CASE, inputs=inp)
    g1.set_params(p)
    g1.init_state()
    g2.init_state()
    for _ in range(3):
        g1.evolve(2)
        g2.evolve(2)
    o1, o2 = Oracle(CASE, inp), Oracle(CASE, inp)
    o1.set_params(p)
    o1.evolve(6)
    o2.evolve(6)
    for g, o in ((g1, o1), (g2, o2)):
        a, b = g.get_state(), o.get_state()
        for c in [0, 1, 2, POT]:
            assert field_rel_err(a[c], b[c]) <= 1e-10, c
