"""Error paths of the C ABI on the GPU: a NaN in the state is reported as
VD_E_NAN (the device max-reductions carry it into the norms the host
reads), a solve that runs out of V-cycles as VD_E_NOCONV, and an adjoint
whose checkpoints no longer describe the context as VD_E_BADARG.  After
each error the context stays usable.  The reference aborts in all three
situations (amrex::Abort, MLMG non-convergence)."""
from __future__ import annotations

import numpy as np
import pytest

from tests._common import small_inputs

pytestmark = pytest.mark.gpu

NS = 4
E_ID, ARP = 0, 1


def _argon(extra=None):
    from vidyut3d_amd import Vidyut
    g = Vidyut("argon2d", inputs=small_inputs("argon2d", n=(32, 32), extra=extra))
    g.init_state()
    return g


@pytest.mark.parametrize("comp", [E_ID, ARP])
def test_nan_in_state_is_reported(comp):
    from vidyut3d_amd import VidyutError
    g = _argon()
    good = g.get_state()
    bad = good.copy()
    bad[comp, 0, 7, 11] = np.nan
    g.set_state(bad)
    with pytest.raises(VidyutError) as e:
        g.evolve(1)
    assert e.value.code == 4, str(e.value)        # VD_E_NAN
    assert "non-finite" in str(e.value)
    g.set_state(bad)
    with pytest.raises(VidyutError) as e:
        g.solve_potential(0.0)
    assert e.value.code == 4, str(e.value)
    # the context recovers with a finite state
    g.set_state(good)
    info = g.evolve(1)
    assert info.poisson_iters > 0
    assert np.all(np.isfinite(g.get_state()))


def test_nan_in_a_fine_level_is_reported():
    from vidyut3d_amd import Vidyut, VidyutError
    g = Vidyut("argon2d", inputs=small_inputs("argon2d", n=(32, 32), extra={"amr.max_level": 1}))
    g.init_state()
    s = g.get_level_state(1)
    s[E_ID, 0, 3, 5] = np.nan
    g.set_level_state(1, s)
    with pytest.raises(VidyutError) as e:
        g.evolve(1)
    assert e.value.code == 4, str(e.value)


def test_no_convergence_is_reported():
    from vidyut3d_amd import VidyutError
    g = _argon({"vidyut.linsolve_maxiter": 1, "vidyut.linsolve_reltol": 1e-14,
                "vidyut.linsolve_abstol": 1e-30})
    with pytest.raises(VidyutError) as e:
        g.solve_potential(0.0)
    assert e.value.code == 2, str(e.value)        # VD_E_NOCONV
    assert "converge" in str(e.value)


def test_stale_adjoint_checkpoints_are_refused():
    from vidyut3d_amd import Vidyut, VidyutError
    sim = Vidyut("argonrf2d", inputs=small_inputs("argonrf2d", n=(16, 16)))
    p = sim.get_params()
    sim.init_state()
    sim.objective_forward(6, 3)
    g1 = sim.adjoint_gradient()
    # parameters changed after the forward: the checkpoints are stale
    sim.set_params(p)
    with pytest.raises(VidyutError) as e:
        sim.adjoint_gradient()
    assert e.value.code == 1, str(e.value)        # VD_E_BADARG
    # a state change does the same
    sim.init_state()
    sim.objective_forward(6, 3)
    sim.set_state(sim.get_state())
    with pytest.raises(VidyutError):
        sim.adjoint_gradient()
    # and a new forward makes it valid again, with the same gradient
    sim.set_params(p)
    sim.init_state()
    sim.objective_forward(6, 3)
    g2 = sim.adjoint_gradient()
    assert np.array_equal(g1, g2)
