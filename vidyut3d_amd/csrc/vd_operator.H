// Kernels of the entry points on caller buffers: the operator entry points
// (vd_poisson_solve, vd_helmholtz_solve; the MLMG operator of the reference
// applied to caller buffers, with boundary data given per side instead of by
// the case's boundary hooks: setLevelBC, Source/PotentialSolve.cpp:194-207,
// 445-448; Source/ScalarSolve.cpp:573-585,820-822) and the ghost fill of a
// caller's ghosted FAB (vd_fill_ghosts).
#ifndef VD_OPERATOR_H_
#define VD_OPERATOR_H_

namespace vd {

// Boundary data of side (d, s) folded into the rhs of nc components: the
// ghost of the side's boundary cell is dep0 x0 + dep1 x1 + gc (the
// homogeneous part is the solver's side model), and the constant part gc
// enters the rhs as  rhs += beta * b_face * gc / h^2  (the fold of
// side_bc_kernel).  gc: Dirichlet 2 g (maxorder 2) or 8/3 g (maxorder 3,
// n >= 2); inhomogeneous Neumann -g h (lo) / +g h (hi) with g = du/dx_d;
// Robin f / (a/2 + b/h).  b_face: bconst, or the boundary cell's own
// coefficient (1/iD) for cell coefficients.  gout: gc per face, for the face
// gradients.
struct BcFoldArgs
{
    Grid g;
    int d, side, nc, kind, maxorder;
    const double* value;     // nc * n1 * n2 (null: zero)
    double ra, rb;           // Robin a, b
    double beta, bconst;
    const double* iD;        // nc * N inverse cell coefficients, or null (bconst)
    double* rhs;             // nc * N
    double* gout;            // nc * n1 * n2
    int n1, n2;
};

__global__ void __launch_bounds__(256) bc_fold_kernel(BcFoldArgs A)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int c = blockIdx.z;
    if (a >= A.n1 || b >= A.n2 || c >= A.nc) return;
    const Grid& g = A.g;
    const int d = A.d;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int C[3] = {0, 0, 0};
    C[t1] = a;
    C[t2] = b;
    C[d] = A.side ? g.n[d] - 1 : 0;
    const long long slab = (long long)A.n1 * A.n2;
    const long long so = (long long)c * slab + a + (long long)A.n1 * b;
    const double gv = A.value ? A.value[so] : 0.0;
    const double h = g.dx[d];
    double gc = 0.0;
    if (A.kind == BK_DIR)
    {
        const int m = min(g.n[d] + 1, A.maxorder);
        gc = (m >= 3) ? (8.0 / 3.0) * gv : 2.0 * gv;
    } else if (A.kind == BK_INHNEU)
        gc = A.side ? gv * h : -gv * h;
    else if (A.kind == BK_ROBIN)
        gc = gv / (0.5 * A.ra + A.rb / h);
    A.gout[so] = gc;
    if (A.kind == BK_DIR || A.kind == BK_INHNEU || A.kind == BK_ROBIN)
    {
        const long long cell = g.idx(C[0], C[1], C[2]);
        double bface = A.iD ? 1.0 / A.iD[(long long)c * g.N + cell] : A.bconst;
        if (d == 0 && g.coord == 1)
        {
            double rm, rp;
            rz_factors(g.problo[0], g.dx[0], C[0], rm, rp);
            bface *= A.side ? rp : rm;
        }
        A.rhs[(long long)c * g.N + cell] -= -A.beta * bface * gc / (h * h);
    }
}

// scale * du/dx_d on the faces of direction d ((n_d + 1) faces along d,
// x fastest), nc components: interior faces (u_R - u_L)/h, periodic faces
// through the image, domain faces through the side's ghost
// dep0 x0 + dep1 x1 + gc (the face gradient of getGradSolution,
// Source/PotentialSolve.cpp:500-510, Source/ScalarSolve.cpp:873-891).
struct FaceGradArgs
{
    Grid g;
    int d, nc;
    const double* u;         // nc * N
    double dep0[2], dep1[2];
    const double* gc[2];     // nc * n1 * n2 per side (null: periodic)
    int n1, n2;
    double scale;
    double* out;             // nc * nfaces
};

__global__ void __launch_bounds__(256) face_grad_kernel(FaceGradArgs A)
{
    const Grid& g = A.g;
    const int d = A.d;
    const int e[3] = {g.n[0] + (d == 0), g.n[1] + (d == 1), g.n[2] + (d == 2)};
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int kc = blockIdx.z;
    const int k = kc % e[2], c = kc / e[2];
    if (i >= e[0] || j >= e[1] || c >= A.nc) return;
    const int F[3] = {i, j, k};
    const int fd = F[d], n = g.n[d];
    const double* u = A.u + (long long)c * g.N;
    const double h = g.dx[d];
    auto at = [&](int q) {
        int I[3] = {F[0], F[1], F[2]};
        I[d] = q;
        return u[g.idx(I[0], I[1], I[2])];
    };
    double grad;
    if (fd > 0 && fd < n)
        grad = (at(fd) - at(fd - 1)) / h;
    else if (g.periodic[d])
        grad = (at(0) - at(n - 1)) / h;
    else
    {
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        const int s = (fd == 0) ? 0 : 1;
        const long long so = (long long)c * A.n1 * A.n2 + F[t1] + (long long)A.n1 * F[t2];
        const double gcv = A.gc[s] ? A.gc[s][so] : 0.0;
        const double x0 = s ? at(n - 1) : at(0);
        const double x1 = (n > 1) ? (s ? at(n - 2) : at(1)) : x0;
        const double ghost = A.dep0[s] * x0 + A.dep1[s] * x1 + gcv;
        grad = s ? (ghost - x0) / h : (x0 - ghost) / h;
    }
    A.out[(long long)c * e[0] * e[1] * e[2] + (long long)i + (long long)e[0] * ((long long)j + (long long)e[1] * k)] =
        A.scale * grad;
}

// Ghost cells of a caller's ghosted FAB (e[d] = n[d] + 2 gd[d] cells per
// direction, x fastest, component outermost, NVAR components), the fill of
// FillPatchSingleLevel with the reference's bcspec (Source/GridWork.cpp:135-175,
// Source/Vidyut.cpp:118-130): per direction the periodic image (int_dir) or
// the nearest valid cell (foextrap), composed over the directions, so an
// edge or corner ghost takes the valid cell whose every index is wrapped or
// clamped.  Valid cells are only read.
struct GhostFillArgs
{
    double* fab;
    int n[3], e[3], gd[3], per[3];
    int g, comp0;
    long long ncell;   // cells of one component of the ghosted box
};

__global__ void __launch_bounds__(64) fill_ghosts_kernel(GhostFillArgs A)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int kc = blockIdx.z;
    const int k = kc % A.e[2], c = kc / A.e[2];
    if (a >= A.e[0]) return;
    const int I[3] = {a - A.gd[0], b - A.gd[1], k - A.gd[2]};
    bool ghost = false;
    int J[3];
    for (int d = 0; d < 3; ++d)
    {
        int q = I[d];
        if (q < 0 || q >= A.n[d])
        {
            ghost = true;
            if (A.per[d]) q = (q < 0) ? q + A.n[d] : q - A.n[d];
            else q = (q < 0) ? 0 : A.n[d] - 1;
        }
        J[d] = q + A.gd[d];
    }
    if (!ghost) return;
    double* f = A.fab + (long long)(A.comp0 + c) * A.ncell;
    f[(long long)a + (long long)A.e[0] * ((long long)b + (long long)A.e[1] * k)] =
        f[(long long)J[0] + (long long)A.e[0] * ((long long)J[1] + (long long)A.e[1] * J[2])];
}

// iD = 1 / b over nc * N cells (the solver's level-0 inverse coefficients).
__global__ void __launch_bounds__(256) inv_coef_kernel(long long n, const double* __restrict__ b, double* __restrict__ iD)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) iD[q] = 1.0 / b[q];
}

} // namespace vd
#endif
