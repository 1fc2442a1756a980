"""GPU parity of the 3D argon LMEA case (electron energy equation with
energy-dependent rates, inelastic and elastic energy exchange, thermal-flux
electron and energy walls, RF potential): whole Evolve steps of the HIP
library against the oracle at 1e-10 relative per field."""
from __future__ import annotations

import pytest

from tests._common import small_inputs, worst_field, efield_rel_err

pytestmark = pytest.mark.gpu

CASE = "lmea3d"
NS = 6
EEN, ETEMP, POT, EFX, EJH, EIH, REF = NS, NS + 1, NS + 2, NS + 3, NS + 6, NS + 7, NS + 9
TOL = 1e-10


@pytest.mark.parametrize("n", [(24, 16, 16), (20, 12, 18)])
def test_lmea3d_steps_match_oracle(n):
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    text = small_inputs(CASE, n=n)
    o = Oracle(CASE, text)
    g = Vidyut(CASE, inputs=text)
    g.set_state(o.get_state())
    for k in range(3):
        io = o.evolve(4)
        ig = g.evolve(4)
        assert ig.poisson_iters == int(io[1])
        a, b = g.get_state(), o.get_state()
        err, c = worst_field(a, b, list(range(5)) + [EEN, ETEMP, POT, REF, EJH, EIH])
        ee = efield_rel_err(a, b, EFX, 3)
        print(n, "steps", 4 * (k + 1), "worst", c, err, "E", ee)
        assert err <= TOL and ee <= TOL
    g.close()


def test_lmea3d_te_dependent_diffusivity():
    """cases/lmea3d_te: the electron (and energy) diffusivity follows T_e,
    uniform in the first step and varying from the second.  The library
    stops storing 1/D of components it found uniform and stores it again
    when a solve finds it varying (the electrons at step 2), then runs the
    harmonic-mean operator; against the oracle, which stores it always."""
    from oracle import Oracle
    from vidyut3d_amd import Vidyut

    case = "lmea3d_te"
    text = small_inputs(case, n=(24, 16, 20))
    o = Oracle(case, text)
    g = Vidyut(case, inputs=text)
    g.set_state(o.get_state())
    g.kernel_timing(True, ["cell_source"])
    launches = []
    for k in range(4):
        io = o.evolve(1)
        ig = g.evolve(1)
        assert ig.poisson_iters == int(io[1])
        a, b = g.get_state(), o.get_state()
        err, c = worst_field(a, b, list(range(5)) + [EEN, ETEMP, POT, REF, EJH, EIH])
        ee = efield_rel_err(a, b, EFX, 3)
        launches.append(g.kernel_times()["cell_source"][1])
        print(case, "step", k + 1, "worst", c, err, "E", ee, "cell-class launches", launches[-1], flush=True)
        assert err <= TOL and ee <= TOL
    g.close()
    # the cell class counts the cell pass, the Joule heating and, in the
    # step whose prediction failed, the 1/D fill: one step has one more
    per_step = [b - a for a, b in zip([0] + launches, launches)]
    assert max(per_step) == min(per_step) + 1, per_step
