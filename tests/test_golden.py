"""Committed fixtures (tests/golden/oracle_<case>.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them (CPU),
and the GPU library must match them at the parity bar (1e-10 relative per
field) without running the oracle."""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests._common import field_rel_err, efield_rel_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"argon2d": ("24 16", 5, 1e-12), "air3d": ("12 10 14", 3, 5e-13)}
TRANSPORTED = {"argon2d": [0, 1, 2], "air3d": list(range(8))}
DIM = {"argon2d": 2, "air3d": 3}


def _fixture(case):
    return dict(np.load(os.path.join(HERE, f"oracle_{case}.npz")))


def _inputs(case):
    from vidyut3d_amd import read_inputs
    return read_inputs(case, {"amr.n_cell": CASES[case][0]})


@pytest.mark.parametrize("case", list(CASES))
def test_oracle_reproduces_fixture(case):
    from oracle import Oracle
    g = _fixture(case)
    n, steps, dt = CASES[case]
    o = Oracle(case, _inputs(case))
    assert np.array_equal(o.get_state(), g["init"])
    efx, dim = o.nspec + 3, DIM[case]
    # E is one field: its components are measured against its magnitude (a
    # transverse component of a nearly 1D field is rounding noise)
    efield = set(range(efx, efx + 3))
    o.solve_potential(0.0)
    s = o.get_state()
    for c in range(s.shape[0]):
        if c not in efield:
            assert field_rel_err(s[c], g["potential"][c]) <= 1e-12, c
    assert efield_rel_err(s, g["potential"], efx, dim) <= 1e-12
    e = o.update_explsrc(0, o.nspec, 0.0, dt)
    for c in range(e.shape[0]):
        assert field_rel_err(e[c], g["expl"][c]) <= 1e-12, c
    o2 = Oracle(case, _inputs(case))
    o2.evolve(steps)
    s = o2.get_state()
    for c in range(s.shape[0]):
        if c not in efield:
            assert field_rel_err(s[c], g["steps"][c]) <= 1e-12, c
    assert efield_rel_err(s, g["steps"], efx, dim) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_gpu_matches_fixture(case):
    from vidyut3d_amd import Vidyut
    g = _fixture(case)
    n, steps, dt = CASES[case]
    v = Vidyut(case, inputs=_inputs(case))
    v.init_state()
    s0 = v.get_state()
    # exp/log/cos of the initial profile and noise: device and host libm may
    # differ in the last bit
    for c in range(s0.shape[0]):
        assert field_rel_err(s0[c], g["init"][c]) <= 1e-14, c
    ns = v.dims.nspec
    efx = ns + 3
    v.solve_potential(0.0)
    s = v.get_state()
    for c in (ns + 2, ns + 9):     # POT, REF
        assert field_rel_err(s[c], g["potential"][c]) <= 1e-10, c
    assert efield_rel_err(s, g["potential"], efx, v.dims.dim) <= 1e-10
    e = v.update_explsrc(0.0, dt)
    for sp in TRANSPORTED[case]:
        assert field_rel_err(e[sp], g["expl"][sp]) <= 1e-10, sp
    w = Vidyut(case, inputs=_inputs(case))
    w.init_state()
    w.evolve(steps)
    s = w.get_state()
    for c in list(range(ns)) + [ns + 2, ns + 9]:
        assert field_rel_err(s[c], g["steps"][c]) <= 1e-10, c
    assert efield_rel_err(s, g["steps"], efx, v.dims.dim) <= 1e-10
