"""Writes the oracle's regression fixtures: small seeded runs of both cases
(initial state, one potential solve, explicit source, and a few whole
steps), stored as compressed .npz next to this script.

The reference cannot be run here (its AMReX dependency is absent), so these
vectors pin the oracle against itself over time: a change to the oracle that
moves any of them fails tests/test_golden.py.  Regenerate only after a
deliberate, separately pinned change:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = {
    # case: (n_cell, steps, dt of the explicit source)
    "argon2d": ("24 16", 5, 1e-12),
    "air3d": ("12 10 14", 3, 5e-13),
}


def generate(case: str) -> dict:
    from oracle import Oracle
    from vidyut3d_amd import read_inputs

    n, steps, dt = CASES[case]
    o = Oracle(case, read_inputs(case, {"amr.n_cell": n}))
    out = {"init": o.get_state().copy()}
    o.solve_potential(0.0)
    out["potential"] = o.get_state().copy()
    out["expl"] = o.update_explsrc(0, o.nspec, 0.0, dt).copy()
    o2 = Oracle(case, read_inputs(case, {"amr.n_cell": n}))
    o2.evolve(steps)
    out["steps"] = o2.get_state().copy()
    return out


def main():
    for case in CASES:
        d = generate(case)
        np.savez_compressed(os.path.join(HERE, f"oracle_{case}.npz"), **d)
        print(case, {k: v.shape for k, v in d.items()})


if __name__ == "__main__":
    main()
