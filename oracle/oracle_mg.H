// TEST INFRASTRUCTURE ONLY -- the oracle.  Nothing in the product links,
// loads or calls this code; tests, the smoke check and the CPU-baseline leg
// of bench.py use it as the checker.
//
// Geometric multigrid for the cell-centred operator that AMReX's
// MLABecLaplacian defines and vidyut3d configures
// (Source/PotentialSolve.cpp:194-207,445-500;
//  Source/ScalarSolve.cpp:573-585,819-875):
//
//   (A a x - B div(b grad x)) = f
//
// Written the plain way: arrays carry one ghost layer, boundary conditions
// are imposed by filling the ghosts before every stencil sweep, exactly as
// MLLinOp::applyBC does.  MLMG internals (cycle, smoother colouring,
// restriction, prolongation, bottom solver, stopping test) are not in the
// reference tree (AMReX is an un-vendored submodule); the algorithm restated
// here is the one the product implements, specified as:
//   * V-cycle, nu1 pre / nu2 post red-black Gauss-Seidel sweeps (red = cells
//     with (i+j+k) even, swept first), exact point update
//     x += (f - Lx) * (1/diag) with diag including the ghost's dependence on
//     x (the residual times the reciprocal, as AMReX's MLABecLaplacian
//     smoother writes phi += omega/g_m_d * res);
//   * restriction = average of the 2^D children, prolongation = cell-centred
//     linear interpolation (3/4, 1/4 per direction; coarse ghosts from the
//     homogeneous boundary model), coarse a = average, coarse face b =
//     average of the 2^(D-1) fine faces;
//   * coarsening by 2 while every extent stays even and >= 2, a periodic
//     extent stays even, and the level count stays within
//     max_coarsening_level;
//   * bottom = red-black sweeps until |r|_inf <= 1e-4 |r0|_inf (max 200);
//     before that, the first coarse level with every extent <= 64 is solved
//     directly (separation of variables, OMG::direct) for every component
//     whose operator separates there;
//   * MLMG stopping test: target = max(abstol, reltol*max(|f|_inf, |r0|_inf)),
//     residual recomputed from the solution after every cycle;
//   * a component whose operator is diagonally dominant (Jacobi bound
//     q <= 1e-2, OMG::smooth_comp) iterates nu2 sweeps on the finest level
//     instead of V-cycles.
// Parity of these internals against AMReX itself is unpinned; the converged
// solution is pinned by dense solves and closed forms in the tests.
#ifndef VD_ORACLE_MG_H_
#define VD_ORACLE_MG_H_

#include <vector>
#include <map>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <algorithm>

namespace oracle {

// O_CF: a fine level's coarse/fine side; homogeneous ghost
// (2/3) x0 - (1/5) x1 (the parent part is folded into the rhs).
enum OBC { O_PER = 0, O_DIR = 1, O_NEU = 2, O_INHNEU = 3, O_ROBIN = 4, O_CF = 5 };
constexpr double kCfW = 8.0 / 15.0, kCfD0 = 2.0 / 3.0, kCfD1 = -1.0 / 5.0;

// One level: nc components, cell arrays with one ghost layer.
struct OLevel
{
    int n[3];
    double dx[3];
    int dim;
    int nc;
    std::vector<double> a;        // cell, 1 comp, no ghosts (A-term coefficient)
    std::vector<double> bf[3];    // face coefficients, nc comps, no ghosts, nodal in d
    std::vector<double> x, f, r;  // ghosted, nc comps
    long long gsz() const
    {
        return (long long)(n[0] + 2) * (dim > 1 ? n[1] + 2 : 1) * (dim > 2 ? n[2] + 2 : 1);
    }
    long long csz() const { return (long long)n[0] * n[1] * n[2]; }
    long long g(int i, int j, int k, int c) const
    {
        long long sx = n[0] + 2, sy = (dim > 1 ? n[1] + 2 : 1);
        int gj = (dim > 1) ? j + 1 : 0;
        int gk = (dim > 2) ? k + 1 : 0;
        return (long long)(i + 1) + sx * (gj + sy * (long long)gk) + (long long)c * gsz();
    }
    long long cidx(int i, int j, int k) const
    {
        return (long long)i + (long long)n[0] * (j + (long long)n[1] * k);
    }
    // face index for direction d: extents n with n[d]+1
    long long fidx(int d, int i, int j, int k, int c) const
    {
        int e0 = n[0] + (d == 0), e1 = n[1] + (d == 1), e2 = n[2] + (d == 2);
        return (long long)i + (long long)e0 * (j + (long long)e1 * k) +
               (long long)c * ((long long)e0 * e1 * e2);
    }
};

// Axisymmetric (RZ, 2D, x = r) metric, the product's row-scaled form of
// AMReX's metric terms (a r_c, radial faces b r_face, axial faces and rhs
// r_c): every operator row of cell i is divided by its centre radius r_c,
// so its radial faces carry r_{i-1/2}/r_c and r_{i+1/2}/r_c.
inline void rz_metric(double rlo, double dx, int i, double& rm, double& rp)
{
    const double rc = rlo + (i + 0.5) * dx;
    rm = (rc - 0.5 * dx) / rc;
    rp = (rc + 0.5 * dx) / rc;
}

struct OMG
{
    int dim = 3;
    int rz = 0;             // axisymmetric metric in x (rz_metric)
    double rz_lo = 0.0;     // radius of the domain's low x face
    double alpha = 0.0, beta = 1.0;
    int bc_lo[3], bc_hi[3];
    int maxorder = 3;
    int nu1 = 2, nu2 = 2, maxiter = 100, max_coarsening = 10;
    // Smoothing iterations instead of V-cycles for a component whose
    // finest-level operator is diagonally dominant enough: with every face
    // coefficient <= bmax_c, the Jacobi (and Gauss-Seidel) iteration matrix
    // has infinity norm at most q_c = 2 B bmax_c sum_d 1/dx_d^2 / (A a); for
    // q_c <= 1e-2 an iteration is e = nu2 red-black sweeps from zero on the
    // residual, x += e, with the same stopping test.  Empty: V-cycles only.
    // (The algorithm the product implements; AMReX's MLMG iterates V-cycles,
    // and both converge to the same tolerance.)
    std::vector<double> smooth_bmax;
    bool smooth_comp(int c) const
    {
        const double aa = alpha * (L[0].a.empty() ? 0.0 : L[0].a[0]);
        if (c >= (int)smooth_bmax.size() || !(aa > 0.0)) return false;
        double hs = 0.0;
        for (int d = 0; d < dim; ++d) hs += 1.0 / (L[0].dx[d] * L[0].dx[d]);
        return 2.0 * beta * smooth_bmax[c] * hs / aa <= 1e-2;
    }
    // Robin coefficients per (side, dir, comp): constant over a face here.
    std::vector<OLevel> L;

    void build(int dim_, const int n0[3], const double dx0[3], int nc)
    {
        dim = dim_;
        L.clear();
        OLevel l0;
        for (int d = 0; d < 3; ++d)
        {
            l0.n[d] = (d < dim) ? n0[d] : 1;
            l0.dx[d] = dx0[d];
        }
        l0.dim = dim;
        l0.nc = nc;
        L.push_back(l0);
        while ((int)L.size() - 1 < max_coarsening)
        {
            const OLevel& c = L.back();
            bool ok = true;
            for (int d = 0; d < dim; ++d)
            {
                if (c.n[d] % 2 != 0 || c.n[d] / 2 < 2) ok = false;
                // a periodic extent stays even (an odd ring has two
                // neighbours of one colour)
                if (bc_lo[d] == O_PER && (c.n[d] / 2) % 2 != 0) ok = false;
            }
            if (!ok) break;
            OLevel nl;
            for (int d = 0; d < 3; ++d)
            {
                nl.n[d] = (d < dim) ? c.n[d] / 2 : 1;
                nl.dx[d] = (d < dim) ? c.dx[d] * 2.0 : c.dx[d];
            }
            nl.dim = dim;
            nl.nc = nc;
            L.push_back(nl);
        }
        for (auto& l : L)
        {
            l.a.assign(l.csz(), 0.0);
            for (int d = 0; d < dim; ++d)
            {
                long long fs = (long long)(l.n[0] + (d == 0)) * (l.n[1] + (d == 1)) * (l.n[2] + (d == 2));
                l.bf[d].assign(fs * nc, 0.0);
            }
            l.x.assign(l.gsz() * nc, 0.0);
            l.f.assign(l.gsz() * nc, 0.0);
            l.r.assign(l.gsz() * nc, 0.0);
        }
    }

    // Coarse coefficients from level 0 ones.
    void coarsen_coeffs()
    {
        for (size_t lv = 1; lv < L.size(); ++lv)
        {
            OLevel& c = L[lv];
            const OLevel& fl = L[lv - 1];
            int nd = 1 << dim;
            for (int k = 0; k < c.n[2]; ++k)
                for (int j = 0; j < c.n[1]; ++j)
                    for (int i = 0; i < c.n[0]; ++i)
                    {
                        double s = 0.0;
                        for (int kk = 0; kk < (dim > 2 ? 2 : 1); ++kk)
                            for (int jj = 0; jj < (dim > 1 ? 2 : 1); ++jj)
                                for (int ii = 0; ii < 2; ++ii)
                                    s += fl.a[fl.cidx(2 * i + ii, 2 * j + jj, dim > 2 ? 2 * k + kk : 0)];
                        c.a[c.cidx(i, j, k)] = s / nd;
                    }
            for (int d = 0; d < dim; ++d)
            {
                int e[3] = {c.n[0] + (d == 0), c.n[1] + (d == 1), c.n[2] + (d == 2)};
                int nt = 1 << (dim - 1);
                for (int comp = 0; comp < c.nc; ++comp)
                    for (int k = 0; k < e[2]; ++k)
                        for (int j = 0; j < e[1]; ++j)
                            for (int i = 0; i < e[0]; ++i)
                            {
                                int fi = (d == 0) ? 2 * i : 2 * i;
                                int fj = (dim > 1) ? 2 * j : 0;
                                int fk = (dim > 2) ? 2 * k : 0;
                                double s = 0.0;
                                // tangential children
                                for (int t2 = 0; t2 < ((dim > 2 && d != 2) ? 2 : 1); ++t2)
                                    for (int t1 = 0; t1 < ((dim > 1 && d != 1) ? 2 : 1); ++t1)
                                        for (int t0 = 0; t0 < (d != 0 ? 2 : 1); ++t0)
                                            s += fl.bf[d][fl.fidx(d, fi + t0, fj + t1, fk + t2, comp)];
                                c.bf[d][c.fidx(d, i, j, k, comp)] = s / nt;
                            }
            }
        }
    }

    int eff_order(const OLevel& l, int d) const { return std::min(l.n[d] + 1, maxorder); }

    // Fill the ghost layer of v (homogeneous BCs) on level l.
    void fill_ghosts(OLevel& l, std::vector<double>& v) const
    {
        for (int c = 0; c < l.nc; ++c)
            for (int d = 0; d < dim; ++d)
            {
                int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                int n1 = (dim > t1) ? l.n[t1] : 1, n2 = (dim > t2) ? l.n[t2] : 1;
                for (int b = 0; b < n2; ++b)
                    for (int a = 0; a < n1; ++a)
                        for (int side = 0; side < 2; ++side)
                        {
                            int I[3] = {0, 0, 0};
                            I[t1] = a;
                            I[t2] = b;
                            int bct = side ? bc_hi[d] : bc_lo[d];
                            int gi = side ? l.n[d] : -1;
                            int i0 = side ? l.n[d] - 1 : 0;
                            int i1 = side ? l.n[d] - 2 : 1;
                            auto at = [&](int q) {
                                int J[3] = {I[0], I[1], I[2]};
                                J[d] = q;
                                return l.g(J[0], J[1], J[2], c);
                            };
                            double val;
                            if (bct == O_PER)
                            {
                                val = v[at(side ? 0 : l.n[d] - 1)];
                            } else if (bct == O_DIR)
                            {
                                int m = eff_order(l, d);
                                if (m >= 3)
                                    val = -2.0 * v[at(i0)] + (1.0 / 3.0) * v[at(i1)];
                                else
                                    val = -v[at(i0)];
                            } else if (bct == O_CF)
                            {
                                val = kCfD0 * v[at(i0)] + kCfD1 * v[at(i1)];
                            } else if (bct == O_ROBIN)
                            {
                                double ra = robin_a(side, d, c), rb = robin_b(side, d, c);
                                double h = l.dx[d];
                                val = -v[at(i0)] * (0.5 * ra - rb / h) / (0.5 * ra + rb / h);
                            } else
                            {
                                val = v[at(i0)];
                            }
                            v[at(gi)] = val;
                        }
            }
    }

    // d(ghost)/d(x0) for the boundary face of direction d, side s.
    double ghost_dep(const OLevel& l, int side, int d, int c) const
    {
        int bct = side ? bc_hi[d] : bc_lo[d];
        if (bct == O_DIR) return eff_order(l, d) >= 3 ? -2.0 : -1.0;
        if (bct == O_CF) return kCfD0;
        if (bct == O_ROBIN)
        {
            double ra = robin_a(side, d, c), rb = robin_b(side, d, c), h = l.dx[d];
            return -(0.5 * ra - rb / h) / (0.5 * ra + rb / h);
        }
        if (bct == O_PER) return 0.0;
        return 1.0; // Neumann: ghost = x0
    }

    std::vector<double> rob_a, rob_b; // [side][d][c]
    double robin_a(int s, int d, int c) const
    {
        return rob_a.empty() ? 0.0 : rob_a[(s * 3 + d) * L[0].nc + c];
    }
    double robin_b(int s, int d, int c) const
    {
        return rob_b.empty() ? 1.0 : rob_b[(s * 3 + d) * L[0].nc + c];
    }

    double Lx_at(const OLevel& l, const std::vector<double>& v, int i, int j, int k, int c) const
    {
        double s = alpha * l.a[l.cidx(i, j, k)] * v[l.g(i, j, k, c)];
        double xc = v[l.g(i, j, k, c)];
        for (int d = 0; d < dim; ++d)
        {
            int I[3] = {i, j, k};
            int Ip[3] = {i, j, k}, Im[3] = {i, j, k};
            Ip[d] += 1;
            Im[d] -= 1;
            int Fp[3] = {i, j, k};
            Fp[d] += 1;
            double bp = l.bf[d][l.fidx(d, Fp[0], Fp[1], Fp[2], c)];
            double bm = l.bf[d][l.fidx(d, I[0], I[1], I[2], c)];
            if (d == 0 && rz)
            {
                double rm, rp;
                rz_metric(rz_lo, l.dx[0], i, rm, rp);
                bm *= rm;
                bp *= rp;
            }
            double h2 = l.dx[d] * l.dx[d];
            s -= beta * (bp * (v[l.g(Ip[0], Ip[1], Ip[2], c)] - xc) -
                         bm * (xc - v[l.g(Im[0], Im[1], Im[2], c)])) / h2;
        }
        return s;
    }

    double diag_at(const OLevel& l, int i, int j, int k, int c) const
    {
        double dg = alpha * l.a[l.cidx(i, j, k)];
        for (int d = 0; d < dim; ++d)
        {
            int I[3] = {i, j, k};
            int Fp[3] = {i, j, k};
            Fp[d] += 1;
            double bp = l.bf[d][l.fidx(d, Fp[0], Fp[1], Fp[2], c)];
            double bm = l.bf[d][l.fidx(d, I[0], I[1], I[2], c)];
            if (d == 0 && rz)
            {
                double rm, rp;
                rz_metric(rz_lo, l.dx[0], i, rm, rp);
                bm *= rm;
                bp *= rp;
            }
            double h2 = l.dx[d] * l.dx[d];
            double fm = 1.0, fp = 1.0;
            if (I[d] == 0 && bc_lo[d] != O_PER) fm = 1.0 - ghost_dep(l, 0, d, c);
            if (I[d] == l.n[d] - 1 && bc_hi[d] != O_PER) fp = 1.0 - ghost_dep(l, 1, d, c);
            dg += beta * (bp * fp + bm * fm) / h2;
        }
        return dg;
    }

    void gsrb(OLevel& l, std::vector<double>& v, const std::vector<double>& rhs, int sweeps)
    {
        for (int s = 0; s < sweeps; ++s)
            for (int color = 0; color < 2; ++color)
            {
                fill_ghosts(l, v);
                for (int c = 0; c < l.nc; ++c)
#pragma omp parallel for collapse(2)
                    for (int k = 0; k < l.n[2]; ++k)
                        for (int j = 0; j < l.n[1]; ++j)
                            for (int i = 0; i < l.n[0]; ++i)
                            {
                                if (((i + j + k) & 1) != color) continue;
                                double lx = Lx_at(l, v, i, j, k, c);
                                double dg = diag_at(l, i, j, k, c);
                                // AMReX's form: phi += (omega/g_m_d) * res, omega = 1
                                v[l.g(i, j, k, c)] += (rhs[l.g(i, j, k, c)] - lx) * (1.0 / dg);
                            }
            }
    }

    // r = rhs - L v ; returns max |r| per component in nrm
    void residual(OLevel& l, std::vector<double>& v, const std::vector<double>& rhs,
                  std::vector<double>& res, std::vector<double>* nrm)
    {
        fill_ghosts(l, v);
        if (nrm) nrm->assign(l.nc, 0.0);
        for (int c = 0; c < l.nc; ++c)
        {
            double mx = 0.0;
#pragma omp parallel for collapse(2) reduction(max : mx)
            for (int k = 0; k < l.n[2]; ++k)
                for (int j = 0; j < l.n[1]; ++j)
                    for (int i = 0; i < l.n[0]; ++i)
                    {
                        double rr = rhs[l.g(i, j, k, c)] - Lx_at(l, v, i, j, k, c);
                        res[l.g(i, j, k, c)] = rr;
                        mx = std::max(mx, std::fabs(rr));
                    }
            if (nrm) (*nrm)[c] = mx;
        }
    }

    void restrict_to(const OLevel& fl, const std::vector<double>& fr, OLevel& cl, std::vector<double>& cr)
    {
        int nd = 1 << dim;
        for (int c = 0; c < cl.nc; ++c)
#pragma omp parallel for collapse(2)
            for (int k = 0; k < cl.n[2]; ++k)
                for (int j = 0; j < cl.n[1]; ++j)
                    for (int i = 0; i < cl.n[0]; ++i)
                    {
                        double s = 0.0;
                        for (int kk = 0; kk < (dim > 2 ? 2 : 1); ++kk)
                            for (int jj = 0; jj < (dim > 1 ? 2 : 1); ++jj)
                                for (int ii = 0; ii < 2; ++ii)
                                    s += fr[fl.g(2 * i + ii, dim > 1 ? 2 * j + jj : 0, dim > 2 ? 2 * k + kk : 0, c)];
                        cr[cl.g(i, j, k, c)] = s / nd;
                    }
    }

    // Value of the coarse correction at (I,J,K) where indices may lie one
    // cell outside the level: out-of-range directions are resolved one at a
    // time (x, then y, then z) with the homogeneous ghost model
    // ghost = dep0*x_b + dep1*x_b2, or the periodic image.
    double cval(const OLevel& l, const std::vector<double>& v, int I, int J, int K, int c) const
    {
        int P[3] = {I, J, K};
        for (int d = 0; d < dim; ++d)
        {
            if (P[d] >= 0 && P[d] < l.n[d]) continue;
            const int side = (P[d] < 0) ? 0 : 1;
            const int bct = side ? bc_hi[d] : bc_lo[d];
            if (bct == O_PER)
            {
                int Q[3] = {P[0], P[1], P[2]};
                Q[d] = side ? 0 : l.n[d] - 1;
                return cval(l, v, Q[0], Q[1], Q[2], c);
            }
            double dep0, dep1;
            if (bct == O_DIR)
            {
                if (eff_order(l, d) >= 3)
                {
                    dep0 = -2.0;
                    dep1 = 1.0 / 3.0;
                } else
                {
                    dep0 = -1.0;
                    dep1 = 0.0;
                }
            } else if (bct == O_CF)
            {
                dep0 = kCfD0;
                dep1 = kCfD1;
            } else if (bct == O_ROBIN)
            {
                dep0 = ghost_dep(l, side, d, c);
                dep1 = 0.0;
            } else
            {
                dep0 = 1.0;
                dep1 = 0.0;
            }
            int Q0[3] = {P[0], P[1], P[2]}, Q1[3] = {P[0], P[1], P[2]};
            Q0[d] = side ? l.n[d] - 1 : 0;
            Q1[d] = side ? l.n[d] - 2 : 1;
            if (Q1[d] < 0 || Q1[d] >= l.n[d]) Q1[d] = Q0[d];
            return dep0 * cval(l, v, Q0[0], Q0[1], Q0[2], c) + dep1 * cval(l, v, Q1[0], Q1[1], Q1[2], c);
        }
        return v[l.g(P[0], P[1], P[2], c)];
    }

    // Cell-centred (bi/tri)linear prolongation: weights 3/4, 1/4 per
    // direction toward the nearer coarse neighbours, added in the order
    // (c, b, a) = z, y, x offset bits.
    void prolong_add_linear(const OLevel& cl, const std::vector<double>& ce, OLevel& fl, std::vector<double>& fe)
    {
        for (int c = 0; c < fl.nc; ++c)
#pragma omp parallel for collapse(2)
            for (int k = 0; k < fl.n[2]; ++k)
                for (int j = 0; j < fl.n[1]; ++j)
                    for (int i = 0; i < fl.n[0]; ++i)
                    {
                        const int I = i / 2, J = (dim > 1) ? j / 2 : 0, K = (dim > 2) ? k / 2 : 0;
                        const int sx = (i & 1) ? 1 : -1, sy = (j & 1) ? 1 : -1, sz = (k & 1) ? 1 : -1;
                        double acc = 0.0;
                        for (int cc = 0; cc < (dim > 2 ? 2 : 1); ++cc)
                            for (int bb = 0; bb < (dim > 1 ? 2 : 1); ++bb)
                                for (int aa = 0; aa < 2; ++aa)
                                {
                                    double w = (aa ? 0.25 : 0.75);
                                    if (dim > 1) w *= (bb ? 0.25 : 0.75);
                                    if (dim > 2) w *= (cc ? 0.25 : 0.75);
                                    acc += w * cval(cl, ce, I + aa * sx, J + bb * sy, K + cc * sz, c);
                                }
                        fe[fl.g(i, j, k, c)] += acc;
                    }
    }
    int linear_prolong = 1;

    void prolong_add(const OLevel& cl, const std::vector<double>& ce, OLevel& fl, std::vector<double>& fe)
    {
        if (linear_prolong)
        {
            prolong_add_linear(cl, ce, fl, fe);
            return;
        }
        for (int c = 0; c < fl.nc; ++c)
#pragma omp parallel for collapse(2)
            for (int k = 0; k < fl.n[2]; ++k)
                for (int j = 0; j < fl.n[1]; ++j)
                    for (int i = 0; i < fl.n[0]; ++i)
                        fe[fl.g(i, j, k, c)] += ce[cl.g(i / 2, dim > 1 ? j / 2 : 0, dim > 2 ? k / 2 : 0, c)];
    }

    void zero_cells(OLevel& l, std::vector<double>& v) { std::fill(v.begin(), v.end(), 0.0); }

    void bottom(OLevel& l)
    {
        zero_cells(l, l.x);
        std::vector<double> n0, n;
        residual(l, l.x, l.f, l.r, &n0);
        for (int it = 0; it < 200; ++it)
        {
            bool done = true;
            for (int c = 0; c < l.nc; ++c)
                if (!(n0[c] == 0.0 || (it > 0 && n[c] <= 1e-4 * n0[c]))) done = false;
            if (it > 0 && done) break;
            if (it == 0)
            {
                bool allzero = true;
                for (int c = 0; c < l.nc; ++c)
                    if (n0[c] != 0.0) allzero = false;
                if (allzero) break;
            }
            gsrb(l, l.x, l.f, 1);
            residual(l, l.x, l.f, l.r, &n);
        }
    }

    // Direct bottom solve by separation of variables (the algorithm the
    // product implements; AMReX's MLMG coarsens further and ends in a
    // BiCGStab bottom solve): the V-cycle stops at the first coarse level
    // whose extents are all <= direct_max_n, and there a component whose
    // operator has a uniform a, a uniform face b and sides without a
    // second-cell ghost term is solved exactly.  Its operator is the
    // Kronecker sum A a I + sum_d (B b / dx_d^2) That_d of the symmetric 1D
    // matrices That_d = tridiag(-1, 2, -1) (end diagonals 2 - dep0, or the
    // periodic wrap), so with That_d = Q_d diag(mu_d) Q_d^T
    //   x = (Q_x Q_y Q_z) diag(1 / (A a + sum_d B b mu_d / dx_d^2)) (Q^T...) f,
    // a mode whose eigenvalue sum is below 1e-12 of the largest getting 0.
    // 0: off.
    int direct_max_n = 128;
    bool direct_finest = false;   // the finest level qualifies too
    int direct_level() const
    {
        if (direct_max_n <= 0) return -1;
        for (int q = direct_finest ? 0 : 1; q < (int)L.size(); ++q)
        {
            bool ok = true;
            for (int d = 0; d < dim; ++d) ok = ok && L[q].n[d] <= direct_max_n;
            if (ok) return q;
        }
        return -1;
    }
    bool separable(const OLevel& l, int c, double& a0, double& b0) const
    {
        if (rz) return false;   // the radial metric breaks the Kronecker form
        a0 = l.a[0];
        for (double v : l.a)
            if (v != a0) return false;
        b0 = l.bf[0][l.fidx(0, 0, 0, 0, c)];
        for (int d = 0; d < dim; ++d)
        {
            const long long fs = (long long)(l.n[0] + (d == 0)) * (l.n[1] + (d == 1)) * (l.n[2] + (d == 2));
            for (long long q = 0; q < fs; ++q)
                if (l.bf[d][q + (long long)c * fs] != b0) return false;
            for (int side = 0; side < 2; ++side)
            {
                const int bct = side ? bc_hi[d] : bc_lo[d];
                if (bct == O_CF || (bct == O_DIR && eff_order(l, d) >= 3)) return false;
            }
        }
        return true;
    }
    // Symmetric eigenproblem by cyclic Jacobi rotations: columns of V are
    // the eigenvectors of S (n x n, row-major), ev the eigenvalues.
    static void jacobi_eig(int n, std::vector<double> S, std::vector<double>& V, std::vector<double>& ev)
    {
        V.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
        auto at = [&](int r, int q) -> double& { return S[(size_t)r * n + q]; };
        double norm2 = 0.0;
        for (double v : S) norm2 += v * v;
        for (int sweep = 0; sweep < 100; ++sweep)
        {
            double off = 0.0;
            for (int r = 0; r < n; ++r)
                for (int q = 0; q < n; ++q)
                    if (r != q) off += at(r, q) * at(r, q);
            if (!(off > 1e-32 * norm2)) break;
            for (int p = 0; p < n - 1; ++p)
                for (int q = p + 1; q < n; ++q)
                {
                    if (at(p, q) == 0.0) continue;
                    const double th = (at(q, q) - at(p, p)) / (2.0 * at(p, q));
                    const double t = std::copysign(1.0, th) / (std::fabs(th) + std::sqrt(1.0 + th * th));
                    const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
                    for (int k = 0; k < n; ++k)
                    {
                        const double u = at(k, p), w = at(k, q);
                        at(k, p) = cs * u - sn * w;
                        at(k, q) = sn * u + cs * w;
                    }
                    for (int k = 0; k < n; ++k)
                    {
                        const double u = at(p, k), w = at(q, k);
                        at(p, k) = cs * u - sn * w;
                        at(q, k) = sn * u + cs * w;
                    }
                    for (int k = 0; k < n; ++k)
                    {
                        const double u = V[(size_t)k * n + p], w = V[(size_t)k * n + q];
                        V[(size_t)k * n + p] = cs * u - sn * w;
                        V[(size_t)k * n + q] = sn * u + cs * w;
                    }
                }
        }
        ev.resize(n);
        for (int i = 0; i < n; ++i) ev[i] = at(i, i);
    }
    // x = L^{-1} f for component c of level l (interior values).
    void direct(OLevel& l, int c, double a0, double b0)
    {
        std::vector<double> V[3], mu[3];
        for (int d = 0; d < dim; ++d)
        {
            const int n = l.n[d];
            std::vector<double> S((size_t)n * n, 0.0);
            const bool per = bc_lo[d] == O_PER;
            for (int i = 0; i < n; ++i)
            {
                S[(size_t)i * n + i] += 2.0;
                if (i > 0) S[(size_t)i * n + i - 1] -= 1.0;
                else if (per) S[(size_t)i * n + n - 1] -= 1.0;
                else S[(size_t)i * n + i] -= ghost_dep(l, 0, d, c);
                if (i < n - 1) S[(size_t)i * n + i + 1] -= 1.0;
                else if (per) S[(size_t)i * n] -= 1.0;
                else S[(size_t)i * n + i] -= ghost_dep(l, 1, d, c);
            }
            // the eigenbasis depends only on (n, the two side entries):
            // computed once per key
            static std::map<std::tuple<int, double, double>, std::pair<std::vector<double>, std::vector<double>>> cache;
            const auto key = std::make_tuple(n, per ? 1e300 : S[0], per ? 1e300 : S[(size_t)n * n - 1]);
            auto it = cache.find(key);
            if (it == cache.end())
            {
                jacobi_eig(n, S, V[d], mu[d]);
                cache.emplace(key, std::make_pair(V[d], mu[d]));
            } else
            {
                V[d] = it->second.first;
                mu[d] = it->second.second;
            }
            for (double& m : mu[d]) m *= beta * b0 / (l.dx[d] * l.dx[d]);
        }
        const int n0 = l.n[0], n1 = l.n[1], n2 = l.n[2];
        std::vector<double> u((size_t)n0 * n1 * n2), w(u.size());
        auto id = [&](int i, int j, int k) { return (size_t)i + (size_t)n0 * ((size_t)j + (size_t)n1 * k); };
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i) u[id(i, j, k)] = l.f[l.g(i, j, k, c)];
        // product along direction d with M (forward: Q^T, back: Q)
        auto along = [&](int d, bool back) {
            const int n = l.n[d];
            const std::vector<double>& Q = V[d];
            // every output is its own sum in m order: threads over (k, j)
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = 0; k < n2; ++k)
                for (int j = 0; j < n1; ++j)
                    for (int i = 0; i < n0; ++i)
                    {
                        const int I[3] = {i, j, k};
                        const int o = I[d];
                        double acc = 0.0;
                        for (int m = 0; m < n; ++m)
                        {
                            int J[3] = {i, j, k};
                            J[d] = m;
                            acc += (back ? Q[(size_t)o * n + m] : Q[(size_t)m * n + o]) * u[id(J[0], J[1], J[2])];
                        }
                        w[id(i, j, k)] = acc;
                    }
            std::swap(u, w);
        };
        for (int d = 0; d < dim; ++d) along(d, false);
        double top = std::fabs(alpha * a0);
        for (int d = 0; d < dim; ++d)
        {
            double mx = 0.0;
            for (double m : mu[d]) mx = std::max(mx, std::fabs(m));
            top += mx;
        }
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i)
                {
                    const int I[3] = {i, j, k};
                    double den = alpha * a0;
                    for (int d = 0; d < dim; ++d) den += mu[d][I[d]];
                    u[id(i, j, k)] = (std::fabs(den) > 1e-12 * top) ? u[id(i, j, k)] / den : 0.0;
                }
        for (int d = dim - 1; d >= 0; --d) along(d, true);
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i) l.x[l.g(i, j, k, c)] = u[id(i, j, k)];
    }

    // V-cycle: solve L_lv x_lv = f_lv approximately from x_lv = 0.
    void vcycle(int lv)
    {
        OLevel& l = L[lv];
        if (lv == direct_level())
        {
            std::vector<double> a0(l.nc), b0(l.nc);
            std::vector<char> sep(l.nc);
            bool all = true;
            for (int c = 0; c < l.nc; ++c)
            {
                sep[c] = separable(l, c, a0[c], b0[c]);
                all = all && sep[c];
            }
            // the other components recurse as usual (components do not couple)
            if (!all) vcycle_recurse(lv);
            for (int c = 0; c < l.nc; ++c)
                if (sep[c]) direct(l, c, a0[c], b0[c]);
            return;
        }
        vcycle_recurse(lv);
    }
    void vcycle_recurse(int lv)
    {
        OLevel& l = L[lv];
        if (lv == (int)L.size() - 1)
        {
            bottom(l);
            return;
        }
        zero_cells(l, l.x);
        gsrb(l, l.x, l.f, nu1);
        residual(l, l.x, l.f, l.r, nullptr);
        restrict_to(l, l.r, L[lv + 1], L[lv + 1].f);
        vcycle(lv + 1);
        prolong_add(L[lv + 1], L[lv + 1].x, l, l.x);
        gsrb(l, l.x, l.f, nu2);
    }

    // Solve L sol = rhs (both nc comps, no ghosts, layout cidx + c*N) with
    // the inhomogeneous boundary contribution bterm already folded into rhs
    // by the caller (rhs_fold = rhs - L_inh(0)).  bnorm: |rhs|_inf per comp of
    // the unfolded rhs.  Returns iterations (max over comps) or -1.
    int solve(std::vector<double>& sol, const std::vector<double>& rhs_fold,
              const std::vector<double>& bnorm, double reltol, double abstol,
              std::vector<int>* iters_per_comp = nullptr)
    {
        OLevel& l0 = L[0];
        const long long N = l0.csz();
        // x <- sol, per component independently (each comp has its own
        // stopping test, as separate MLMG solves would)
        int maxit_used = 0;
        if (iters_per_comp) iters_per_comp->assign(l0.nc, 0);
        std::vector<double> xs(l0.gsz() * l0.nc, 0.0), fs(l0.gsz() * l0.nc, 0.0), rs(l0.gsz() * l0.nc, 0.0);
        for (int c = 0; c < l0.nc; ++c)
            for (int k = 0; k < l0.n[2]; ++k)
                for (int j = 0; j < l0.n[1]; ++j)
                    for (int i = 0; i < l0.n[0]; ++i)
                    {
                        xs[l0.g(i, j, k, c)] = sol[l0.cidx(i, j, k) + c * N];
                        fs[l0.g(i, j, k, c)] = rhs_fold[l0.cidx(i, j, k) + c * N];
                    }
        std::vector<double> rn0;
        residual(l0, xs, fs, rs, &rn0);
        std::vector<double> target(l0.nc);
        std::vector<int> active(l0.nc, 1);
        for (int c = 0; c < l0.nc; ++c)
        {
            target[c] = std::max(abstol, reltol * std::max(bnorm[c], rn0[c]));
            if (rn0[c] <= target[c]) active[c] = 0;
        }
        std::vector<double> rn = rn0;
        for (int it = 1; it <= maxiter; ++it)
        {
            bool any = false;
            for (int c = 0; c < l0.nc; ++c) any = any || active[c];
            if (!any) break;
            // V-cycle on the residual (inactive comps get zero residual so
            // their correction is exactly zero), then the smoothing
            // iteration of the components that take it
            for (int pass = 0; pass < 2; ++pass)
            {
                const bool smooth = pass == 1;
                bool any_here = false;
                for (int c = 0; c < l0.nc; ++c) any_here = any_here || (active[c] && smooth_comp(c) == smooth);
                if (!any_here) continue;
                for (int c = 0; c < l0.nc; ++c)
                {
                    const bool on = active[c] && smooth_comp(c) == smooth;
                    for (long long q = 0; q < l0.gsz(); ++q)
                        l0.f[q + c * l0.gsz()] = on ? rs[q + c * l0.gsz()] : 0.0;
                }
                if (smooth)
                {
                    zero_cells(l0, l0.x);
                    gsrb(l0, l0.x, l0.f, nu2);
                } else
                    vcycle(0);
                for (int c = 0; c < l0.nc; ++c)
                    if (active[c] && smooth_comp(c) == smooth)
                        for (int k = 0; k < l0.n[2]; ++k)
                            for (int j = 0; j < l0.n[1]; ++j)
                                for (int i = 0; i < l0.n[0]; ++i)
                                    xs[l0.g(i, j, k, c)] += l0.x[l0.g(i, j, k, c)];
            }
            residual(l0, xs, fs, rs, &rn);
            for (int c = 0; c < l0.nc; ++c)
                if (active[c])
                {
                    if (iters_per_comp) (*iters_per_comp)[c] = it;
                    maxit_used = it;
                    if (rn[c] <= target[c]) active[c] = 0;
                }
        }
        for (int c = 0; c < l0.nc; ++c)
            if (active[c]) return -1;
        for (int c = 0; c < l0.nc; ++c)
            for (int k = 0; k < l0.n[2]; ++k)
                for (int j = 0; j < l0.n[1]; ++j)
                    for (int i = 0; i < l0.n[0]; ++i)
                        sol[l0.cidx(i, j, k) + c * N] = xs[l0.g(i, j, k, c)];
        return maxit_used;
    }
};

} // namespace oracle
#endif
