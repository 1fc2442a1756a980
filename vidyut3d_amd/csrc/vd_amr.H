// Static nested refinement (ratio 2, no subcycling) on gfx950: the
// cross-level operations of vidyut3d's multi-level step
// (Source/GridWork.cpp:112-202 FillPatchTwoLevels/AverageDown,
// Source/ScalarSolve.cpp:125-131 average_down_faces) and of the composite
// MLMG solves (Source/PotentialSolve.cpp:445-500, Source/ScalarSolve.cpp:
// 573-585 build one MLABecLaplacian over all levels).
//
// A fine level is one box strictly inside its parent level's box.  Arrays
// of a level are laid out like the single-level state (x fastest, component
// outermost, no ghosts); a LevelMap locates the fine box in the parent box.
//
// Composite discretisation (AMReX's MLMG internals are not in the reference
// tree; this is the specification the oracle restates):
//   * coarse/fine ghost of a fine boundary cell B0 (B1 the next one inward):
//       g = (8/15) u_int + (2/3) x(B0) - (1/5) x(B1),
//     the quadratic through the parent value (a coarse half-cell outside the
//     interface), B0 and B1; u_int is the parent value C moved to the ghost's
//     tangential position with central slopes:
//       u_int = u_C + 0.125 * sum_t s_t (u_{C+e_t} - u_{C-e_t}),  s_t = +-1;
//   * composite residual: on covered coarse cells the mean of the 2^D fine
//     residuals; on coarse cells across an interface face the coarse face
//     gradient replaced by the mean of the fine face gradients (reflux);
//   * FillPatch ghosts of a fine level: parent value plus MC-limited slopes
//     per direction times +-1/4 (conservative linear interpolation);
//   * AverageDown: covered coarse cell = mean of its 2^D children; covered
//     coarse face = mean of its 2^(D-1) fine faces.
#ifndef VD_AMR_H_
#define VD_AMR_H_

#include "vd_grid.H"
#include "vd_reduce.H"

namespace vd {

struct LevelMap
{
    int dim;
    int n[3];        // fine box extents
    int plo[3];      // parent region origin in the parent box (local parent index)
    int pn[3];       // parent box extents
    long long N, pN; // cells of the fine box and of the parent box
};

constexpr double kCfW = 8.0 / 15.0;   // weight of the parent value in the C/F ghost
constexpr double kCfD0 = 2.0 / 3.0;   // ... of the boundary cell
constexpr double kCfD1 = -1.0 / 5.0;  // ... of the next cell inward

__device__ __forceinline__ long long amr_pidx(const LevelMap& m, int I, int J, int K)
{
    return (long long)I + (long long)m.pn[0] * ((long long)J + (long long)m.pn[1] * K);
}
__device__ __forceinline__ long long amr_fidx(const LevelMap& m, int i, int j, int k)
{
    return (long long)i + (long long)m.n[0] * ((long long)j + (long long)m.n[1] * k);
}

// Components of up to kMaxComp arrays: comp c of a field lives at
// base + map[c] * stride.
struct CompSet
{
    int nc;
    int map[kMaxComp];
};

// Slab layout of the C/F data of a fine level: component outermost, then
// the sides in the order (d = 0, s = 0), (0, 1), (1, 0), ..., each n1*n2
// values (a along t1, b along t2, t1 < t2 the tangential directions).
struct CfSlabs
{
    long long off[3][2];  // offset of side (d, s) within one component
    int n1[3], n2[3];
    long long tot;        // values of one component over all sides
};

// w * u_int of every fine ghost position next to the box, for nc components
// of the parent array (parent + map[c]*pstride).  One thread per (a, b) of a
// side; all sides in one launch (each writes its own slab): grid.y covers
// the largest n2, grid.z = (side, component) with side = 2 d + s.
__global__ void __launch_bounds__(256) amr_cf_interp_kernel(LevelMap m, const double* __restrict__ parent,
                                                          long long pstride, CompSet cs, CfSlabs sl,
                                                          double* __restrict__ slabs, double w)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int side = (int)blockIdx.z / cs.nc;
    const int c = (int)blockIdx.z - side * cs.nc;
    const int d = side >> 1, s = side & 1;
    const int n1 = sl.n1[d], n2 = sl.n2[d];
    if (a >= n1 || b >= n2) return;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int P[3];
    P[d] = s ? m.plo[d] + m.n[d] / 2 : m.plo[d] - 1;
    int sg[3] = {0, 0, 0};
    P[t1] = m.plo[t1] + (a >> 1);
    sg[t1] = (a & 1) ? 1 : -1;
    if (m.dim > 2)
    {
        P[t2] = m.plo[t2] + (b >> 1);
        sg[t2] = (b & 1) ? 1 : -1;
    } else
    {
        P[t2] = 0;
    }
    const double* u = parent + (long long)cs.map[c] * pstride;
    const long long pc = amr_pidx(m, P[0], P[1], P[2]);
    double sum = 0.0;
    {
        const long long st = (t1 == 0) ? 1 : (t1 == 1 ? m.pn[0] : (long long)m.pn[0] * m.pn[1]);
        sum += double(sg[t1]) * (u[pc + st] - u[pc - st]);
    }
    if (m.dim > 2)
    {
        const long long st = (t2 == 0) ? 1 : (t2 == 1 ? m.pn[0] : (long long)m.pn[0] * m.pn[1]);
        sum += double(sg[t2]) * (u[pc + st] - u[pc - st]);
    }
    slabs[(long long)c * sl.tot + sl.off[d][s] + a + (long long)n1 * b] = w * (u[pc] + 0.125 * sum);
}

// rhs[c][B0] -= -beta * b_f * (8/15) u_int / h^2 on side (d, s): the parent
// part of the C/F ghost moved to the right-hand side (the cycle uses the
// homogeneous part).  b_f: bconst, or 1/iD[c][B0] when iD is given.  Both
// sides of direction d in one launch (grid.z = (s, component)): they never
// share a boundary cell, while sides of different directions share edge
// cells and keep separate launches in (d, s) order.
__global__ void __launch_bounds__(256) amr_cf_fold_kernel(LevelMap m, int d, const double* __restrict__ slabs,
                                                        CfSlabs sl, int nc, double* __restrict__ rhs, double beta,
                                                        double bconst, const double* __restrict__ iD, double h)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int s = (int)blockIdx.z / nc;
    const int c = (int)blockIdx.z - s * nc;
    const int n1 = sl.n1[d], n2 = sl.n2[d];
    if (a >= n1 || b >= n2 || c >= nc) return;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int I[3] = {0, 0, 0};
    I[d] = s ? m.n[d] - 1 : 0;
    I[t1] = a;
    I[t2] = (m.dim > 2) ? b : 0;
    const long long o = (long long)c * m.N + amr_fidx(m, I[0], I[1], I[2]);
    const double ui = slabs[(long long)c * sl.tot + sl.off[d][s] + a + (long long)n1 * b];
    const double bf = iD ? 1.0 / iD[o] : bconst;
    rhs[o] -= -beta * bf * (kCfW * ui) / (h * h);
}

// dst[c][B0] = src[c][B0] on the boundary cells of every side of the box,
// the cells amr_cf_fold_kernel changes (the same index rule), for nc
// components; a cell on an edge is written by each of its sides with the
// same value.  Restores a folded rhs to its raw values there.
__global__ void __launch_bounds__(256) amr_shell_copy_kernel(LevelMap m, CfSlabs sl, int nc,
                                                           const double* __restrict__ src, double* __restrict__ dst)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int side = (int)blockIdx.z / nc;
    const int c = (int)blockIdx.z - side * nc;
    const int d = side >> 1, s = side & 1;
    if (a >= sl.n1[d] || b >= sl.n2[d]) return;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int I[3] = {0, 0, 0};
    I[d] = s ? m.n[d] - 1 : 0;
    I[t1] = a;
    I[t2] = (m.dim > 2) ? b : 0;
    const long long o = (long long)c * m.N + amr_fidx(m, I[0], I[1], I[2]);
    dst[o] = src[o];
}

// Mean of the 2^D children of every parent cell under the fine box:
// coarse[cc][P] = mean fine[fc][children], cc = cmap[c], fc = fmap[c].
__global__ void __launch_bounds__(256) amr_average_kernel(LevelMap m, const double* __restrict__ fine, long long fstride,
                                                        CompSet fm, double* __restrict__ coarse, long long cstride,
                                                        CompSet cm)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    const int hz = (m.dim > 2) ? m.n[2] / 2 : 1;
    const int c = blockIdx.z / hz;
    const int K = blockIdx.z - c * hz;
    if (I >= m.n[0] / 2 || J >= ((m.dim > 1) ? m.n[1] / 2 : 1)) return;
    const double* f = fine + (long long)fm.map[c] * fstride;
    double sum = 0.0;
#pragma unroll
    for (int kk = 0; kk < (m.dim > 2 ? 2 : 1); ++kk)
#pragma unroll
        for (int jj = 0; jj < (m.dim > 1 ? 2 : 1); ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
                sum += f[amr_fidx(m, 2 * I + ii, (m.dim > 1) ? 2 * J + jj : 0, (m.dim > 2) ? 2 * K + kk : 0)];
    const double nd = (m.dim > 2) ? 8.0 : 4.0;
    coarse[(long long)cm.map[c] * cstride + amr_pidx(m, m.plo[0] + I, (m.dim > 1) ? m.plo[1] + J : 0,
                                                     (m.dim > 2) ? m.plo[2] + K : 0)] = sum / nd;
}

// Reflux of side (d, s): for the parent cell C just outside the interface
// face, r[C] += sgn * beta * (mean_fine_grad - coarse_grad) / hc, where the
// gradients point in +d, sgn = +1 when the fine box lies on C's high side.
// Fine face gradient b_f (x_in - g)/hf (lo side) or b_f (g - x_in)/hf (hi
// side) with the C/F ghost g from u_int (slabs, or 0 when slabs is null) and
// the fine values; coarse gradient from the parent array (0 when null).
// Coefficients: Poisson constant bconst; species b_f = 1/iD_f[B0],
// b_c = 2/(iD_c[C] + iD_c[C']) (harmonic mean of the parent cells).
struct RefluxArgs
{
    LevelMap m;
    const double* fine;     // fine values, comp c at fine + fmap[c]*fstride
    long long fstride;
    CompSet fmap;
    const double* slabs;    // C/F u_int (nullable: zero)
    CfSlabs sl;
    const double* coarse;   // parent values (nullable: zero), comp at coarse + cmap[c]*cstride
    long long cstride;
    CompSet cmap;
    double* r;              // parent residual, comp c at r + c*pN
    double beta, bconst;
    const double* iDf;      // fine inverse coefficients (nc * N), nullable
    const double* iDc;      // parent inverse coefficients (nc * pN), nullable
    double hf[3], hc[3];    // fine and parent cell sizes per direction
};

// All sides in one launch (grid.z = (side, component), side = 2 d + s): a
// side adds only into the parent cells just outside its own face.
__global__ void __launch_bounds__(256) amr_reflux_kernel(RefluxArgs A)
{
    const LevelMap& m = A.m;
    const int side = (int)blockIdx.z / A.fmap.nc;
    const int c = (int)blockIdx.z - side * A.fmap.nc;
    const int d = side >> 1, s = side & 1;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int A1 = blockIdx.x * blockDim.x + threadIdx.x;    // parent tangential index t1
    const int B1 = blockIdx.y;                               // parent tangential index t2
    const int m1 = m.n[t1] / 2, m2 = (m.dim > 2) ? m.n[t2] / 2 : 1;
    if (A1 >= m1 || B1 >= m2) return;
    const int n1 = A.sl.n1[d];
    // fine boundary cells B0/B1 and the children faces
    double gsum = 0.0;
    const double* f = A.fine + (long long)A.fmap.map[c] * A.fstride;
    const long long fst = (d == 0) ? 1 : (d == 1 ? m.n[0] : (long long)m.n[0] * m.n[1]);
    for (int q2 = 0; q2 < ((m.dim > 2) ? 2 : 1); ++q2)
        for (int q1 = 0; q1 < 2; ++q1)
        {
            const int a = 2 * A1 + q1, b = (m.dim > 2) ? 2 * B1 + q2 : 0;
            int I[3] = {0, 0, 0};
            I[d] = s ? m.n[d] - 1 : 0;
            I[t1] = a;
            I[t2] = b;
            const long long o0 = amr_fidx(m, I[0], I[1], I[2]);
            const long long o1 = s ? o0 - fst : o0 + fst;
            const double x0 = f[o0], x1 = (m.n[d] > 1) ? f[o1] : x0;
            const double ui = A.slabs ? A.slabs[(long long)c * A.sl.tot + A.sl.off[d][s] + a + (long long)n1 * b] : 0.0;
            const double g = kCfW * ui + kCfD0 * x0 + kCfD1 * x1;
            const double bf = A.iDf ? 1.0 / A.iDf[(long long)c * m.N + o0] : A.bconst;
            gsum += s ? bf * (g - x0) / A.hf[d] : bf * (x0 - g) / A.hf[d];
        }
    const double gmean = gsum / ((m.dim > 2) ? 4.0 : 2.0);
    int P[3], Pin[3];
    P[d] = s ? m.plo[d] + m.n[d] / 2 : m.plo[d] - 1;
    Pin[d] = s ? P[d] - 1 : P[d] + 1;
    P[t1] = Pin[t1] = m.plo[t1] + A1;
    P[t2] = Pin[t2] = (m.dim > 2) ? m.plo[t2] + B1 : 0;
    const long long pc = amr_pidx(m, P[0], P[1], P[2]), pi = amr_pidx(m, Pin[0], Pin[1], Pin[2]);
    double gc = 0.0;
    if (A.coarse)
    {
        const double* u = A.coarse + (long long)A.cmap.map[c] * A.cstride;
        const double bc = A.iDc ? 2.0 / (A.iDc[(long long)c * m.pN + pc] + A.iDc[(long long)c * m.pN + pi]) : A.bconst;
        gc = s ? bc * (u[pc] - u[pi]) / A.hc[d] : bc * (u[pi] - u[pc]) / A.hc[d];
    }
    const double sgn = s ? -1.0 : 1.0;
    A.r[(long long)c * m.pN + pc] += sgn * A.beta * (gmean - gc) / A.hc[d];
}

// fine[c] += P(coarse[c]) over the whole fine box: cell-centred
// (bi/tri)linear interpolation, weights 3/4, 1/4 per direction toward the
// nearer parents, summed in (z, y, x) offset-bit order (the multigrid
// prolongation).  One thread per parent cell of the region.
// DIM is the level's dimension (m.dim) at compile time, so the
// neighbourhood stays in registers (a run-time bound on these loops put it
// in scratch memory).
template <int DIM>
__global__ void __launch_bounds__(256) amr_prolong_kernel(LevelMap m, const double* __restrict__ coarse,
                                                        double* __restrict__ fine, unsigned active)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    const int hz = (DIM > 2) ? m.n[2] / 2 : 1;
    const int c = blockIdx.z / hz;
    const int K = blockIdx.z - c * hz;
    if (!((active >> c) & 1u)) return;
    if (I >= m.n[0] / 2 || J >= ((DIM > 1) ? m.n[1] / 2 : 1)) return;
    const double* u = coarse + (long long)c * m.pN;
    const int QI = m.plo[0] + I, QJ = (DIM > 1) ? m.plo[1] + J : 0, QK = (DIM > 2) ? m.plo[2] + K : 0;
    const long long cy = m.pn[0], cz = (long long)m.pn[0] * m.pn[1];
    const long long o = amr_pidx(m, QI, QJ, QK);
    double nb[3][3][3];
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x)
            {
                const int dz = (DIM > 2) ? z - 1 : 0, dy = (DIM > 1) ? y - 1 : 0;
                nb[z][y][x] = ((DIM > 2 || z == 1) && (DIM > 1 || y == 1)) ? u[o + (x - 1) + cy * dy + cz * dz] : 0.0;
            }
#pragma unroll
    for (int kk = 0; kk < (DIM > 2 ? 2 : 1); ++kk)
#pragma unroll
        for (int jj = 0; jj < (DIM > 1 ? 2 : 1); ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
            {
                const int sx = ii ? 1 : -1, sy = jj ? 1 : -1, sz = kk ? 1 : -1;
                double acc = 0.0;
#pragma unroll
                for (int cc = 0; cc < (DIM > 2 ? 2 : 1); ++cc)
#pragma unroll
                    for (int bb = 0; bb < (DIM > 1 ? 2 : 1); ++bb)
#pragma unroll
                        for (int aa = 0; aa < 2; ++aa)
                        {
                            double w = aa ? 0.25 : 0.75;
                            if (DIM > 1) w *= bb ? 0.25 : 0.75;
                            if (DIM > 2) w *= cc ? 0.25 : 0.75;
                            acc += w * nb[(DIM > 2) ? 1 + cc * sz : 1][(DIM > 1) ? 1 + bb * sy : 1][1 + aa * sx];
                        }
                const long long fo = (long long)c * m.N +
                                     amr_fidx(m, 2 * I + ii, (DIM > 1) ? 2 * J + jj : 0, (DIM > 2) ? 2 * K + kk : 0);
                fine[fo] += acc;
            }
}

// Ghosted copy of a fine level's state for its stencil kernels (FillPatch
// with FillPatchTwoLevels, Source/GridWork.cpp:135-175): components
// [0, ncg) on the box grown by ng; inside the box a copy, outside the parent
// value plus MC-limited slopes times +-1/4 per direction (cell_cons_interp
// form).  Layout: x fastest over n+2ng, component outermost.
__global__ void __launch_bounds__(256) amr_fill_ghosted_kernel(LevelMap m, int ng, int ncg, const double* __restrict__ fine,
                                                             const double* __restrict__ parent, double* __restrict__ out)
{
    const int gx = m.n[0] + 2 * ng, gy = (m.dim > 1) ? m.n[1] + 2 * ng : 1, gz = (m.dim > 2) ? m.n[2] + 2 * ng : 1;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    if (x >= gx || y >= gy || z >= gz) return;
    const long long gN = (long long)gx * gy * gz;
    const long long go = (long long)x + (long long)gx * ((long long)y + (long long)gy * z);
    const int i = x - ng, j = (m.dim > 1) ? y - ng : 0, k = (m.dim > 2) ? z - ng : 0;
    const bool inside = (i >= 0 && i < m.n[0]) && (j >= 0 && j < m.n[1]) && (k >= 0 && k < m.n[2]);
    if (inside)
    {
        const long long fo = amr_fidx(m, i, j, k);
        for (int c = 0; c < ncg; ++c) out[(long long)c * gN + go] = fine[(long long)c * m.N + fo];
        return;
    }
    // parent cell (floor division: box origins are even) and child signs
    const int F[3] = {i, j, k};
    int P[3], sg[3];
    for (int d = 0; d < 3; ++d)
    {
        if (d < m.dim)
        {
            P[d] = m.plo[d] + (F[d] >> 1);
            sg[d] = (F[d] & 1) ? 1 : -1;
        } else
        {
            P[d] = 0;
            sg[d] = 0;
        }
    }
    const long long pc = amr_pidx(m, P[0], P[1], P[2]);
    for (int c = 0; c < ncg; ++c)
    {
        const double* u = parent + (long long)c * m.pN;
        const double uc = u[pc];
        double acc = 0.0;
        for (int d = 0; d < m.dim; ++d)
        {
            const long long st = (d == 0) ? 1 : (d == 1 ? m.pn[0] : (long long)m.pn[0] * m.pn[1]);
            const double dl = uc - u[pc - st], dr = u[pc + st] - uc;
            const double dc = 0.5 * (dl + dr);
            double sl = 0.0;
            if (dl * dr > 0.0) sl = copysign(fmin(fabs(dc), fmin(2.0 * fabs(dl), 2.0 * fabs(dr))), dc);
            acc += double(sg[d]) * sl;
        }
        out[(long long)c * gN + go] = uc + 0.25 * acc;
    }
}

// Covered coarse faces of direction d = mean of their 2^(D-1) fine faces
// (average_down_faces), nc components of face arrays (nodal in d).
__global__ void __launch_bounds__(256) amr_avgdown_faces_kernel(LevelMap m, int d, int nc, const void* __restrict__ ffv,
                                                              void* __restrict__ cfv, int is_float)
{
    const int em0 = m.n[0] / 2 + (d == 0), em1 = (m.dim > 1) ? m.n[1] / 2 + (d == 1) : 1;
    const int em2 = (m.dim > 2) ? m.n[2] / 2 + (d == 2) : 1;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    const int c = blockIdx.z / em2;
    const int K = blockIdx.z - c * em2;
    if (I >= em0 || J >= em1 || c >= nc) return;
    const long long fe0 = m.n[0] + (d == 0), fe1 = m.n[1] + (d == 1), fe2 = m.n[2] + (d == 2);
    const long long ce0 = m.pn[0] + (d == 0), ce1 = m.pn[1] + (d == 1), ce2 = m.pn[2] + (d == 2);
    const long long fnf = fe0 * fe1 * fe2, cnf = ce0 * ce1 * ce2;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    double sum = 0.0;
    for (int q2 = 0; q2 < ((m.dim > 2) ? 2 : 1); ++q2)
        for (int q1 = 0; q1 < ((m.dim > 1) ? 2 : 1); ++q1)
        {
            int F[3] = {2 * I, 2 * J, 2 * K};
            F[t1] += q1;
            if (m.dim > 2) F[t2] += q2;
            if (m.dim < 3) F[2] = 0;
            if (m.dim < 2) F[1] = 0;
            const long long fo = (long long)c * fnf + F[0] + fe0 * (F[1] + fe1 * F[2]);
            sum += is_float ? double(static_cast<const float*>(ffv)[fo]) : static_cast<const double*>(ffv)[fo];
        }
    const double mean = sum / ((m.dim > 2) ? 4.0 : (m.dim > 1 ? 2.0 : 1.0));
    const long long co = (long long)c * cnf + (m.plo[0] + I) + ce0 * ((long long)((m.dim > 1) ? m.plo[1] + J : 0) +
                                                                      ce1 * ((m.dim > 2) ? m.plo[2] + K : 0));
    if (is_float) static_cast<float*>(cfv)[co] = float(mean);
    else static_cast<double*>(cfv)[co] = mean;
}

// Per-workgroup maxima of |v| over nc components of N cells (grid.y = nc).
__global__ void __launch_bounds__(256) amr_absmax_kernel(const double* __restrict__ v, long long N, unsigned active,
                                                       double* __restrict__ partials, long long nb)
{
    __shared__ double scratch[16];
    const int c = blockIdx.y;
    double mx = 0.0;
    if ((active >> c) & 1u)
        for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += (long long)gridDim.x * blockDim.x)
            mx = vmax(mx, fabs(v[(long long)c * N + q]));
    block_max_store(mx, partials + (long long)c * nb + blockIdx.x, scratch);
}

} // namespace vd
#endif
