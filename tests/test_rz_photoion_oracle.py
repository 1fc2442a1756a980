"""Pins the oracle's axisymmetric (RZ) path and its three-term
photoionization against what the mathematics fixes, on the reference's own
verification cases compiled unchanged (cases/laplace_axisym ->
test/verification/Laplace_Axisym, cases/streamer_photoion ->
test/verification/Streamer_Axisym_Photoion):

- the RZ potential converges at second order to the log(r) solution the
  reference's Laplace_Axisym checks (test/verification/Laplace_Axisym/
  get_error.py: phi = (phi2 log(r/rmin) + phi1 log(rmax/r)) / log(rmax/rmin));
- the RZ potential and one photoionization term equal sparse direct solves
  of the finite-volume operators written the way AMReX applies its metric
  terms (a r_c, radial faces b r_face, axial faces and rhs r_c), built here
  independently of the oracle's row-scaled form;
- the explicit source's axisymmetric correction (Source/ScalarSolve.cpp:
  331-378) is -mu n E_r / r: the RZ minus Cartesian difference times r is the
  same in every cell of a uniform state.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests._common import small_inputs


def _oracle(case, n, **extra):
    from oracle import Oracle

    ov = {"vidyut.linsolve_reltol": 1e-13, "vidyut.linsolve_abstol": 1e-30}
    ov.update(extra)
    return Oracle(case, small_inputs(case, n=n, extra=ov))


def test_laplace_axisym_second_order_to_log_solution():
    errs = []
    for n in (16, 32, 64):
        o = _oracle("laplace_axisym", (n, n))
        o.solve_potential(0.0)
        pot = o.get_state()[o.nspec + 2, 0]
        r = 0.1 + (np.arange(n) + 0.5) * 0.1 / n
        exact = (20.0 * np.log(r / 0.1) + 10.0 * np.log(0.2 / r)) / np.log(2.0)
        errs.append(float(np.sqrt(np.mean((pot - exact[None, :]) ** 2))))
        o.close()
    print("RMS errors", errs)
    assert errs[-1] < 3e-4
    for a, b in zip(errs, errs[1:]):
        assert 3.9 < a / b < 4.1


def _rz_operator(nr, nz, r0, dr, dz, a, zlo_dir, zhi_dir, rlo_dir, rhi_dir):
    """AMReX-form RZ operator (metric-weighted, symmetric): row (i, j)
    r_c a u - [r_{i+1/2}(u_E - u) - r_{i-1/2}(u - u_W)]/dr^2
            - r_c [(u_N - u) - (u - u_S)]/dz^2,
    Dirichlet sides (value on the face, maxorder 2) with ghost -u (the value
    part goes to the rhs by the caller), Neumann sides with ghost u."""
    N = nr * nz
    A = sp.lil_matrix((N, N))
    idx = lambda i, j: i + nr * j
    for j in range(nz):
        for i in range(nr):
            q = idx(i, j)
            rc = r0 + (i + 0.5) * dr
            rw, re = r0 + i * dr, r0 + (i + 1) * dr
            A[q, q] += rc * a
            for (nb, rf, dirichlet) in ((i - 1, rw, rlo_dir), (i + 1, re, rhi_dir)):
                A[q, q] += rf / dr**2
                if 0 <= nb < nr:
                    A[q, idx(nb, j)] -= rf / dr**2
                elif dirichlet:
                    A[q, q] += rf / dr**2        # ghost -u
                else:
                    A[q, q] -= rf / dr**2        # ghost u
            for (nb, dirichlet) in ((j - 1, zlo_dir), (j + 1, zhi_dir)):
                A[q, q] += rc / dz**2
                if 0 <= nb < nz:
                    A[q, idx(i, nb)] -= rc / dz**2
                elif dirichlet:
                    A[q, q] += rc / dz**2
                else:
                    A[q, q] -= rc / dz**2
    return A.tocsr()


def test_rz_potential_matches_sparse_direct_solve():
    nr, nz = 16, 12
    o = _oracle("laplace_axisym", (nr, nz))
    o.solve_potential(0.0)
    pot = o.get_state()[o.nspec + 2, 0].reshape(-1)
    r0, dr, dz = 0.1, 0.1 / nr, 0.1 / nz
    A = _rz_operator(nr, nz, r0, dr, dz, 0.0, False, False, True, True)
    rhs = np.zeros(nr * nz)
    for j in range(nz):
        rhs[0 + nr * j] += 2 * 10.0 * r0 / dr**2                 # ghost 2 V_L - u
        rhs[nr - 1 + nr * j] += 2 * 20.0 * (r0 + nr * dr) / dr**2
    u = spla.spsolve(A, rhs)
    err = np.max(np.abs(pot - u)) / np.max(np.abs(u))
    print("RZ potential vs direct", err)
    assert err < 1e-11


def test_photoionization_matches_sparse_direct_solve():
    """One term of the streamer case: (a - div grad) S = rhs in RZ with the
    case's sides (photoion_bc 2 1: Neumann in r, Dirichlet 0 in z), the
    case's a and rhs as the oracle evaluated them."""
    nr, nz = 24, 16
    o = _oracle("streamer_photoion", (nr, nz))
    o.solve_potential(0.0)      # the rhs reads E/N, as after the step's potential solve
    for sph in range(3):
        S, rhs, a = o.photoionization(0.0, sph)
        assert np.all(a == a.flat[0]) and a.flat[0] > 0
        assert np.max(np.abs(rhs)) > 0
        L = 0.0125
        dr, dz = L / nr, L / nz
        A = _rz_operator(nr, nz, 0.0, dr, dz, float(a.flat[0]), True, True, False, False)
        rc = (np.arange(nr) + 0.5) * dr
        b = (rhs[0] * rc[None, :]).reshape(-1)
        u = spla.spsolve(A, b)
        err = np.max(np.abs(S.reshape(-1) - u)) / np.max(np.abs(u))
        print("photoionization term", sph, "a", a.flat[0], "vs direct", err)
        assert err < 1e-10


def test_axisym_correction_is_drift_over_radius():
    """Uniform densities and a uniform radial field: the Cartesian flux
    divergence vanishes away from the sides, so the RZ minus Cartesian
    explicit source is the correction -mu n E_r / r; times r it is the same
    in every interior cell, and it has the sign of -mu E_r."""
    nr, nz = 32, 24
    vals = []
    for coord in (1, 0):
        o = _oracle("streamer_photoion", (nr, nz), **{"geometry.coord_sys": coord,
                                                       "vidyut.do_photoionization": 0})
        s = o.get_state()
        ns = o.nspec
        for c in range(ns):
            s[c] = s[c].mean()
        s[ns + 3] = 1.5e6       # E_r
        s[ns + 4] = 0.0         # E_z
        o.set_state(s)
        vals.append(o.update_explsrc(0, ns, 0.0, 1e-12))
        o.close()
    d = vals[0] - vals[1]
    r = (np.arange(nr) + 0.5) * 0.0125 / nr
    # electrons (mu < 0 in this case, -2.3987 |E|^-0.26): the drift n mu E_r
    # points inward, so the correction adds density; N2+ is immobile here
    rd = d[0, 0, 4:-4, 4:-4] * r[None, 4:-4]
    ref = rd[0, 0]
    print("electrons: r * correction", ref, "spread", np.max(np.abs(rd - ref)) / abs(ref))
    assert ref > 0.0
    assert np.max(np.abs(rd - ref)) <= 1e-12 * abs(ref)
    assert np.all(d[2] == 0.0)
