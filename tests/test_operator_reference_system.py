"""The sparse reference system of the operator entry-point GPU tests, pinned
by a manufactured solution: u = sin(2x + 0.3) cos(y) exp(z/2) on the unit
cube, -div(-grad u) = lap u = -4.75 u, with Dirichlet, Robin
(a u + b du/dn = f) and inhomogeneous Neumann sides whose data are u's own
values and derivatives.  The discrete solution converges to u at second
order (a sign error in the boundary fold, a swapped side or a transposed
side array stops the convergence)."""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from tests.test_gpu_operator_api import DIR, INH, ROB, _grad_faces, _system


def _u(x, y, z):
    return np.sin(2 * x + 0.3) * np.cos(y) * np.exp(z / 2)


def _err(n, maxorder):
    c = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    ux = lambda x, y, z: 2 * np.cos(2 * x + 0.3) * np.cos(y) * np.exp(z / 2)
    uy = lambda x, y, z: -np.sin(2 * x + 0.3) * np.sin(y) * np.exp(z / 2)
    uz = lambda x, y, z: 0.5 * _u(x, y, z)
    P, Q = np.meshgrid(c, c, indexing="ij")      # (slow, fast) axes of a side array
    a, b = 2.0, 0.3
    sides = {(0, 0): (DIR, _u(0, Q, P), 0, 0),
             (0, 1): (ROB, a * _u(1, Q, P) + b * ux(1, Q, P), a, b),
             (1, 0): (INH, uy(Q, 0, P), 0, 0),
             (1, 1): (INH, uy(Q, 1, P), 0, 0),
             (2, 0): (ROB, a * _u(Q, P, 0) - b * uz(Q, P, 0), a, b),
             (2, 1): (DIR, _u(Q, P, 1), 0, 0)}
    bc = [[sides[(d, 0)], sides[(d, 1)]] for d in range(3)]
    h = (1 / n,) * 3
    A, shift = _system((n, n, n), h, 0.0, 1.0, -np.ones((n, n, n)), bc, maxorder)
    s = spla.spsolve(A, (-4.75 * _u(X, Y, Z)).reshape(-1) + shift)
    gx = _grad_faces(s, (n, n, n), h, bc, maxorder, 0)
    xf = np.arange(n + 1) / n
    Zf, Yf, Xf = np.meshgrid(c, c, xf, indexing="ij")
    return np.max(np.abs(s - _u(X, Y, Z).reshape(-1))), np.max(np.abs(gx - ux(Xf, Yf, Zf)))


def test_reference_system_converges_to_the_manufactured_solution():
    for maxorder in (2, 3):
        (e1, g1), (e2, g2) = _err(8, maxorder), _err(16, maxorder)
        assert np.log2(e1 / e2) > 1.85, (maxorder, e1, e2)
        # face gradients: first order at the boundary faces with the linear
        # Dirichlet ghost, close to second with the quadratic one
        assert np.log2(g1 / g2) > (0.95 if maxorder == 2 else 1.5), (maxorder, g1, g2)
