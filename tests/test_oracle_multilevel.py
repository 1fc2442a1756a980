"""Pins of the oracle's static-refinement restatement (composite solves over
levels, coarse/fine ghosts, reflux, average_down_faces, AverageDown).

* A charge-free potential between Dirichlet plates is linear; the composite
  discretisation (quadratic coarse/fine ghosts, refluxed coarse cells,
  averaged-down covered cells) reproduces linear functions exactly, so every
  level holds E = -V/L.  A wrong ghost weight, a missing reflux or a wrong
  restriction leaves a visible error at the interface.
* Diffusion with zero-flux walls and no field or reactions conserves the
  number of particles; on the composite grid (uncovered cells of every level
  weighted by their volume) that holds only if the coarse faces under the
  fine level carry the fine fluxes (average_down_faces for the explicit part,
  reflux in the implicit solve) and covered cells are averaged down."""
from __future__ import annotations

import numpy as np

NS = 6
EEN, ETEMP, POT, EFX = NS, NS + 1, NS + 2, NS + 3


def _oracle(extra: dict, n=16, levels=2):
    from oracle import Oracle
    from vidyut3d_amd import read_inputs

    ov = {"amr.n_cell": f"{n} {n} {n}", "amr.max_level": levels}
    ov.update(extra)
    return Oracle("lmea3d", read_inputs("lmea3d", ov))


def test_linear_potential_exact_on_all_levels():
    V, L = 300.0, 0.02
    o = _oracle({"vidyut.voltage_profile": 0, "vidyut.voltage_amp_hi": f"{V} 0.0 0.0",
                 "vidyut.do_spacechrg": 0, "vidyut.do_reactions": 0, "vidyut.do_transport": 0})
    o.evolve(1)
    for lev in range(o.num_levels()):
        s = o.get_level_state(lev)
        lo, n = o.level_box(lev)
        dx = L / (16 * 2 ** lev)
        x = (lo[0] + np.arange(n[0]) + 0.5) * dx
        assert np.max(np.abs(s[POT] - (V * x / L)[None, None, :])) <= 1e-8 * V, lev
        assert np.max(np.abs(s[EFX] + V / L)) <= 1e-7 * V / L, lev
        assert np.max(np.abs(s[EFX + 1])) <= 1e-7 * V / L and np.max(np.abs(s[EFX + 2])) <= 1e-7 * V / L


def _composite_total(o, comp):
    """sum over uncovered cells of every level of n * cell volume"""
    L = 0.02
    tot = 0.0
    nl = o.num_levels()
    for lev in range(nl):
        s = o.get_level_state(lev)[comp]
        vol = (L / (16 * 2 ** lev)) ** 3
        mask = np.ones_like(s, bool)
        if lev + 1 < nl:
            clo, cn = o.level_box(lev + 1)
            lo, _ = o.level_box(lev)
            a = [clo[d] // 2 - lo[d] for d in range(3)]
            mask[a[2]:a[2] + cn[2] // 2, a[1]:a[1] + cn[1] // 2, a[0]:a[0] + cn[0] // 2] = False
        tot += float(np.sum(s[mask])) * vol
    return tot


def test_transport_conserves_composite_particle_number():
    # a drifting, diffusing electron bump across the level interfaces (no
    # space charge, no reactions, zero-flux walls, the bump far from them)
    # (periodic in y and z: no walls there for the solver-tolerance
    # transverse field to push particles through)
    o = _oracle({"vidyut.voltage_profile": 0, "vidyut.voltage_amp_hi": "300.0 0.0 0.0", "vidyut.do_spacechrg": 0,
                 "vidyut.do_reactions": 0, "geometry.is_periodic": "0 1 1", "vidyut.pot_bc_lo": "1 0 0",
                 "vidyut.pot_bc_hi": "1 0 0",
                 "vidyut.eden_bc_lo": "2 0 0", "vidyut.eden_bc_hi": "2 0 0", "vidyut.ion_bc_lo": "2 0 0",
                 "vidyut.ion_bc_hi": "2 0 0", "vidyut.neutral_bc_lo": "2 0 0", "vidyut.neutral_bc_hi": "2 0 0",
                 "vidyut.eenrg_bc_lo": "2 0 0", "vidyut.eenrg_bc_hi": "2 0 0", "vidyut.dt": 1e-7})
    for lev in range(o.num_levels()):
        s = o.get_level_state(lev)
        lo, n = o.level_box(lev)
        dx = 0.02 / (16 * 2 ** lev)
        x = (lo[0] + np.arange(n[0]) + 0.5) * dx
        ne = 1e15 * np.exp(-((x - 0.0125) / 0.001) ** 2) + 1e3
        s[0] = ne[None, None, :]
        s[EEN] = 1.5 * 1.380649e-23 * s[0] * 11604.5
        s[ETEMP] = 11604.5
        o.set_level_state(lev, s)
    t0 = _composite_total(o, 0)
    o.evolve(3)
    t1 = _composite_total(o, 0)
    s = o.get_level_state(2)[0]
    assert np.ptp(s) > 0           # the profile changed
    assert abs(t1 - t0) <= 1e-9 * t0, (t0, t1)
