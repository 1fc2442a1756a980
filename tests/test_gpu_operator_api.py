"""The MLMG operator entry points on caller device buffers (vd_poisson_solve,
vd_helmholtz_solve; the reference builds MLABecLaplacian + MLMG per call,
Source/PotentialSolve.cpp:194-207,445-510, Source/ScalarSolve.cpp:573-585,
820-879) against sparse direct solves of the same discrete problems, written
here from the boundary conditions' definitions:

  Dirichlet u = g on the face       ghost = 2 g - u0 (maxorder 2) or
                                    8/3 g - 2 u0 + u1/3 (maxorder 3)
  inhomogeneous Neumann du/dx_d = g ghost = u0 -+ g h (lo / hi)
  Robin a u + b du/dn = f           u_face = (u0 + ghost)/2,
                                    du/dn = (ghost - u0)/h

and the face gradients (E = -grad phi for the potential) from those ghosts.
The context's own steps are unchanged by the calls (bitwise against a fresh
context).  Solver tolerance 1e-13, bar 1e-9 relative."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests._common import HipBuffer, small_inputs

pytestmark = pytest.mark.gpu

DIR, NEU, INH, ROB, PER = 1, 2, 3, 4, 0


def _ghost(kind, maxorder, n, h, g, ra, rb, side):
    """(dep0, dep1, const) of the ghost beyond the boundary cell of a side."""
    if kind == DIR:
        if min(n + 1, maxorder) >= 3:
            return -2.0, 1.0 / 3.0, 8.0 / 3.0 * g
        return -1.0, 0.0, 2.0 * g
    if kind == NEU:
        return 1.0, 0.0, 0.0 * g
    if kind == INH:
        return 1.0, 0.0, (g * h if side else -g * h)
    if kind == ROB:
        den = 0.5 * ra + rb / h
        return -(0.5 * ra - rb / h) / den, 0.0, g / den
    raise ValueError(kind)


def _system(shape, h, alpha, beta, bcell, bc, maxorder):
    """Matrix and rhs shift of alpha u - beta div(b grad u) (harmonic face b,
    boundary faces the boundary cell's b) on one component; shape (nz, ny,
    nx), bc[d] = ((kind, values, ra, rb), (kind, values, ra, rb)) with values
    over the side's cells in the array order of the other axes."""
    nz, ny, nx = shape
    n = (nx, ny, nz)
    dim = sum(1 for e in n if e > 1) if nz == 1 else 3
    N = nx * ny * nz
    idx = lambda i, j, k: i + nx * (j + ny * k)
    A = sp.lil_matrix((N, N))
    shift = np.zeros(N)
    b = bcell.reshape(-1)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                A[r, r] += alpha
                I = [i, j, k]
                for d in range(dim):
                    for s, step in ((0, -1), (1, 1)):
                        J = list(I)
                        J[d] += step
                        h2 = h[d] ** 2
                        if 0 <= J[d] < n[d]:
                            q = idx(*J)
                            bf = 2 * b[r] * b[q] / (b[r] + b[q])
                            A[r, q] += -beta * bf / h2
                            A[r, r] += beta * bf / h2
                            continue
                        kind, vals, ra, rb = bc[d][s]
                        if kind == PER:
                            J[d] %= n[d]
                            q = idx(*J)
                            bf = 2 * b[r] * b[q] / (b[r] + b[q])
                            A[r, q] += -beta * bf / h2
                            A[r, r] += beta * bf / h2
                            continue
                        t = [a for a in range(3) if a != d]
                        gval = vals[I[t[1]], I[t[0]]] if vals is not None else 0.0
                        d0, d1, c0 = _ghost(kind, maxorder, n[d], h[d], gval, ra, rb, s)
                        bf = b[r]
                        # -beta bf (ghost - u0)/h^2 with ghost = d0 u0 + d1 u1 + c0: the c0 part moves to the rhs
                        A[r, r] += beta * bf / h2 - beta * bf * d0 / h2
                        if d1:
                            J1 = list(I)
                            J1[d] += -step
                            A[r, idx(*J1)] += -beta * bf * d1 / h2
                        shift[r] += beta * bf * c0 / h2
    return A.tocsr(), shift


def _grad_faces(u, shape, h, bc, maxorder, d):
    nz, ny, nx = shape
    n = (nx, ny, nz)
    u = u.reshape(shape)
    ax = 2 - d
    m = n[d]
    lo = np.take(u, [0], axis=ax)
    hi = np.take(u, [m - 1], axis=ax)
    inner = np.diff(u, axis=ax) / h[d]
    faces = []
    for s, x0, x1 in ((0, lo, np.take(u, [min(1, m - 1)], axis=ax)), (1, hi, np.take(u, [max(m - 2, 0)], axis=ax))):
        kind, vals, ra, rb = bc[d][s]
        if kind == PER:
            faces.append((np.take(u, [0], axis=ax) - np.take(u, [m - 1], axis=ax)) / h[d])
            continue
        gv = np.zeros_like(x0) if vals is None else np.expand_dims(vals, ax)
        d0, d1, _ = _ghost(kind, maxorder, m, h[d], 0.0, ra, rb, s)
        c0 = _ghost(kind, maxorder, m, h[d], 1.0, ra, rb, s)[2] * gv
        ghost = d0 * x0 + d1 * x1 + c0
        faces.append((x0 - ghost) / h[d] if s == 0 else (ghost - x0) / h[d])
    return np.concatenate([faces[0], inner, faces[1]], axis=ax)


def _vd_bc(bc, dev_vals):
    from vidyut3d_amd import BC
    B = BC()
    for d in range(3):
        for s in range(2):
            if d < len(bc):
                kind, vals, ra, rb = bc[d][s]
                B.kind[d][s] = kind
                B.value[d][s] = dev_vals[(d, s)].ptr if (d, s) in dev_vals else None
                B.robin_a[d][s] = ra
                B.robin_b[d][s] = rb
    return B


def test_poisson_solve_matches_direct_solve_3d():
    from vidyut3d_amd import SolveInfo, Vidyut
    nx, ny, nz = 16, 12, 20
    g = Vidyut("air3d", inputs=small_inputs("air3d", n=(nx, ny, nz)))
    L = 0.02
    h = (L / nx, L / ny, L / nz)
    rng = np.random.default_rng(5)
    shape = (nz, ny, nx)
    rhs = rng.standard_normal(shape) * 1e7
    sides = {
        (0, 0): (DIR, rng.standard_normal((nz, ny)) * 50.0, 0.0, 0.0),
        (0, 1): (ROB, rng.standard_normal((nz, ny)), 2.0, 0.003),
        (1, 0): (INH, rng.standard_normal((nz, nx)) * 1e3, 0.0, 0.0),
        (1, 1): (NEU, None, 0.0, 0.0),
        (2, 0): (DIR, rng.standard_normal((ny, nx)) * 20.0, 0.0, 0.0),
        (2, 1): (DIR, None, 0.0, 0.0),
    }
    bc = [[sides[(d, 0)], sides[(d, 1)]] for d in range(3)]
    dev_vals = {k: HipBuffer(v[1].shape, v[1]) for k, v in sides.items() if v[1] is not None}
    A, shift = _system(shape, h, 0.0, 1.0, -np.ones(shape), bc, 2)
    exact = spla.spsolve(A, rhs.reshape(-1) + shift)
    d_rhs = HipBuffer(shape, rhs)
    d_phi = HipBuffer(shape, 0.0)
    efs = [HipBuffer((nz + (d == 2), ny + (d == 1), nx + (d == 0)), np.nan) for d in range(3)]
    faces = (ctypes.c_void_p * 3)(*[e.ptr.value for e in efs])
    info = SolveInfo()
    B = _vd_bc(bc, dev_vals)
    g._chk(g.lib.vd_poisson_solve(g.h, d_rhs.ptr, d_phi.ptr, ctypes.byref(B), 1e-13, 1e-30, faces,
                                  ctypes.byref(info)))
    phi = d_phi.download().reshape(-1)
    err = np.max(np.abs(phi - exact)) / np.max(np.abs(exact))
    print("poisson iters", info.iters, "err", err)
    assert info.iters > 0 and err <= 1e-9
    for d in range(3):
        ef = efs[d].download()
        ex = -_grad_faces(exact, shape, h, bc, 2, d)
        e = np.max(np.abs(ef - ex)) / np.max(np.abs(ex))
        print("E face", d, e)
        assert e <= 1e-9


def test_helmholtz_solve_matches_direct_solve_2d():
    from vidyut3d_amd import SolveInfo, Vidyut
    nx, ny = 24, 16      # argon2d: y periodic
    g = Vidyut("argon2d", inputs=small_inputs("argon2d", n=(nx, ny)))
    L = 0.01
    h = (L / nx, L / ny, 1.0)
    shape = (1, ny, nx)
    rng = np.random.default_rng(7)
    nc = 2
    alpha, beta = 2.0e4, 1.0      # diffusion (b/h^2 ~ 6e5) dominates the diagonal
    bcell = np.exp(rng.uniform(-1, 1, (nc,) + shape)) * 0.1
    rhs = rng.standard_normal((nc,) + shape) * 1e20
    sides = {(0, 0): (DIR, rng.uniform(1, 2, (nc, 1, ny)) * 1e15, 0.0, 0.0), (0, 1): (INH, rng.standard_normal((nc, 1, ny)) * 1e17, 0.0, 0.0),
             (1, 0): (PER, None, 0.0, 0.0), (1, 1): (PER, None, 0.0, 0.0)}
    d_sol = HipBuffer((nc,) + shape, 0.0)
    d_rhs = HipBuffer((nc,) + shape, rhs)
    d_b = HipBuffer((nc,) + shape, bcell)
    dev_vals = {k: HipBuffer(v[1].shape, v[1]) for k, v in sides.items() if v[1] is not None}
    gfs = [HipBuffer((nc, 1, ny + (d == 1), nx + (d == 0)), np.nan) for d in range(2)]
    faces = (ctypes.c_void_p * 3)(gfs[0].ptr.value, gfs[1].ptr.value, None)
    bc_all = [[sides[(0, 0)], sides[(0, 1)]], [sides[(1, 0)], sides[(1, 1)]]]
    info = SolveInfo()
    B = _vd_bc(bc_all, dev_vals)
    g._chk(g.lib.vd_helmholtz_solve(g.h, nc, d_rhs.ptr, d_sol.ptr, alpha, beta, d_b.ptr, ctypes.byref(B),
                                    1e-13, 1e-30, faces, ctypes.byref(info)))
    sol = d_sol.download()
    for c in range(nc):
        bcc = [[(k, None if v is None else v[c], ra, rb) for (k, v, ra, rb) in bc_all[d]] for d in range(2)]
        A, shift = _system(shape, h, alpha, beta, bcell[c], bcc, 3)
        exact = spla.spsolve(A, rhs[c].reshape(-1) + shift)
        err = np.max(np.abs(sol[c].reshape(-1) - exact)) / np.max(np.abs(exact))
        print("helmholtz comp", c, "iters", info.iters, "err", err)
        assert err <= 1e-9
        for d in range(2):
            gd = gfs[d].download()[c]
            ex = _grad_faces(exact, shape, h, bcc, 3, d)
            e = np.max(np.abs(gd - ex)) / np.max(np.abs(ex))
            print("grad face", c, d, e)
            assert e <= 1e-9


def test_operator_calls_leave_the_context_unchanged():
    from vidyut3d_amd import SolveInfo, Vidyut
    text = small_inputs("argon2d", n=(32, 32))
    ref = Vidyut("argon2d", inputs=text)
    ref.init_state()
    ref.evolve(3)
    g = Vidyut("argon2d", inputs=text)
    g.init_state()
    g.evolve(1)
    shape = (1, 32, 32)
    rhs = HipBuffer(shape, 1.0)
    rhs3 = HipBuffer((3,) + shape, 1.0)
    sol = HipBuffer((3,) + shape, 0.0)
    bb = HipBuffer((3,) + shape, 0.1)
    B = _vd_bc([[(DIR, None, 0, 0), (DIR, None, 0, 0)], [(PER, None, 0, 0), (PER, None, 0, 0)]], {})
    info = SolveInfo()
    g._chk(g.lib.vd_poisson_solve(g.h, rhs.ptr, sol.ptr, ctypes.byref(B), 1e-10, 1e-30, None, ctypes.byref(info)))
    g._chk(g.lib.vd_helmholtz_solve(g.h, 3, rhs3.ptr, sol.ptr, 1e9, 1.0, bb.ptr, ctypes.byref(B), 1e-10, 1e-30,
                                    None, ctypes.byref(info)))
    g.evolve(2)
    assert np.array_equal(g.get_state(), ref.get_state())
