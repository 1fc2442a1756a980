"""GPU parity of static refinement (levels of ratio 2, each the central
half of its parent, no subcycling): whole Evolve steps of the HIP library on
every level against the oracle's multi-level restatement, 1e-10 relative per
field.  Covers the composite potential and species solves, coarse/fine
ghosts, reflux, average_down_faces, FillPatch ghosts for upwind (2D, 3D) and
WENO5 stencils, the electron energy on fine levels and AverageDown."""
from __future__ import annotations

import pytest

from tests._common import worst_field, efield_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10

CASES = [
    # case, n_cell, max_level, steps, nspec, dim
    ("lmea3d", "16 16 16", 2, 3, 6, 3),
    ("argon2d", "32 32", 1, 4, 4, 2),
    ("air3d", "16 16 16", 1, 2, 10, 3),
]


@pytest.mark.parametrize("case,ncell,levels,steps,ns,dim", CASES)
def test_multilevel_steps_match_oracle(case, ncell, levels, steps, ns, dim):
    from oracle import Oracle
    from vidyut3d_amd import Vidyut, read_inputs

    text = read_inputs(case, {"amr.n_cell": ncell, "amr.max_level": levels})
    o = Oracle(case, text)
    g = Vidyut(case, inputs=text)
    assert g.num_levels() == levels + 1 == o.num_levels()
    for lev in range(levels + 1):
        assert g.level_box(lev) == o.level_box(lev)
        g.set_level_state(lev, o.get_level_state(lev))
    POT, EFX, REF = ns + 2, ns + 3, ns + 9
    comps = list(range(ns)) + [ns, ns + 1, POT, REF, ns + 6, ns + 7]
    for k in range(steps):
        io = o.evolve(1)
        ig = g.evolve(1)
        assert ig.poisson_iters == int(io[1])
        for lev in range(levels + 1):
            a, b = g.get_level_state(lev), o.get_level_state(lev)
            err, c = worst_field(a, b, comps)
            ee = efield_rel_err(a, b, EFX, dim)
            print(case, "step", k + 1, "level", lev, "worst", c, err, "E", ee, "pois iters", ig.poisson_iters)
            assert err <= TOL and ee <= TOL, (lev, c, err, ee)
    g.close()
