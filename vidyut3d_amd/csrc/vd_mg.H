// Geometric multigrid on gfx950 for the cell-centred operator
//   L x = alpha_a x - beta * sum_d d/dx_d ( b_d dx_d x )
// (AMReX's MLABecLaplacian form, which vidyut3d uses for the potential
// (Source/PotentialSolve.cpp:194-207) and the species diffusion solves
// (Source/ScalarSolve.cpp:573-585)), batched over up to kMaxComp independent
// components with per-component stopping tests.
//
// Algorithm (kept identical in the CPU oracle, so the two differ only by
// floating-point rounding):
//   * inhomogeneous boundary data folded into the rhs once per solve; the
//     cycle itself works with homogeneous boundary ghosts
//     ghost = dep0*x0 + dep1*x1 (Neumann 1,0; Dirichlet order-2 -1,0;
//     Dirichlet order-3 -2,1/3; Robin -(a/2-b/h)/(a/2+b/h),0; periodic wrap);
//   * correction scheme: r = f - L x, e = Vcycle(r), x += e, until
//     |r|_inf <= max(abstol, reltol*max(|f|_inf, |r0|_inf)) per component;
//   * V-cycle: nu1 pre / nu2 post red-black Gauss-Seidel sweeps (red =
//     (i+j+k) even first) with the exact point update x += (f-Lx)/diag,
//     residual fused with restriction (average of 2^D children), cell-centred
//     linear prolongation (weights 3/4, 1/4 per direction, coarse ghosts from
//     the same homogeneous model), bottom = red-black sweeps in one workgroup until
//     |r|_inf <= 1e-4 |r0|_inf (max 200 sweeps);
//   * coefficients: a uniform; face b either a constant, the harmonic mean
//     2/(1/b_L + 1/b_R) of cell values (finest level), or stored face
//     arrays averaged from the finer level (2^(D-1) faces).
//
// Roofline: every sweep/residual is a streaming stencil over HBM; bytes per
// cell per component: GSRB sweep 3 doubles (x read+write, f read), residual
// +restrict 2 doubles (+1/2^D write), prolongation 2 doubles + 1/2^D.
#ifndef VD_MG_H_
#define VD_MG_H_

#include "vd_grid.H"
#include "vd_reduce.H"
#include "vd_ktimer.H"
#include "vd_amr.H"
#include <rocblas/rocblas.h>
#include "vd_mg_direct.H"
#include <vector>
#include <string>
#include <stdexcept>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace vd {

enum CoefKind : int { CK_CONST = 0, CK_CELLINV = 1, CK_FACE = 2 };

template <class T>
struct MGView
{
    int n[3];
    long long Nl;          // cells of this level
    double h2inv[3];       // 1/dx^2
    double alpha_a;        // A*a (uniform)
    double beta;           // B
    double bconst;         // constant face coefficient (CK_CONST)
    const double* bcomp;   // CK_CONST per-component constants (nullptr: bconst)
    const T* iD;           // CK_CELLINV: nc comps of 1/b (cell)
    const T* bf[3];        // CK_FACE: nc comps of faces, nodal in d
    long long nf[3];       // faces per comp for each direction
    int rz;                // axisymmetric (RZ) metric in x (2D), see rz_factors
    double rz_lo, rz_dx;   // radius of the level's low x face, its cell width
};


// Stencil evaluation at one cell: returns L x and the diagonal.
// Branch-free: every neighbour load is issued unconditionally from an
// in-range index (the cell itself or its periodic image at a domain side),
// then boundary ghosts are selected, so the loads of a wavefront overlap
// instead of waiting on each other behind per-cell branches.  Accumulation
// order (the same in the oracle):
//   acc = alpha_a x;  per direction d: acc -= h_d (b_p (x_p - x) - b_m (x - x_m))
//   diag = alpha_a;   per direction d: diag += h_d (b_p f_p + b_m f_m)
// with x_m/x_p the homogeneous ghost dep0*x0 + dep1*x1 at a non-periodic
// side (f = 1 - dep0 there, 1 otherwise).
//
// The arithmetic of one operator row from gathered values: xmr/xpr are the
// raw neighbours (any in-range value where lob/hib mark a non-periodic
// domain side, whose ghost comes from the side's model instead), bm/bp the
// face coefficients.  Every kernel that applies the operator goes through
// here, so they agree bit for bit.
// NGT1: every extent is known to exceed 1 (the n > 1 choices resolve at
// compile time to the same values).
// Fused multiply-add of the stencil arithmetic.  Every product-sum of the
// operator row is an explicit fma (or a product that only feeds one as its
// addend), so no kernel leaves the compiler a contraction choice: the
// per-colour, residual, coarse and fused kernels, however their code is
// arranged, round every row identically.
__device__ __forceinline__ double vfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float vfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Gauss-Seidel point update x + (f - Lx)/diag, written the way AMReX's
// MLABecLaplacian smoother writes it (phi += omega/g_m_d * res, omega = 1):
// the residual times the reciprocal of the diagonal, one fma.  Every
// smoother here goes through it, with idg = 1/diag of the same diagonal.
template <class T>
__device__ __forceinline__ T gs_update(T x, T res, T idg)
{
    return vfma(res, idg, x);
}

template <int DIM, class T, bool NGT1 = false>
__device__ __forceinline__ void stencil_core(const MGView<T>& v, const SideBC (*bc)[2], T xc, const T* xmr,
                                             const T* xpr, const bool* lob, const bool* hib, const T* bm,
                                             const T* bp, T& Lx, T& diag)
{
    T acc = T(0), dg = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
        const int n = v.n[d];
        const SideBC& s0 = bc[d][0];
        const SideBC& s1 = bc[d][1];
        // next interior cell of a boundary cell: the opposite raw neighbour
        const T x1m = (NGT1 || n > 1) ? xpr[d] : xc, x1p = (NGT1 || n > 1) ? xmr[d] : xc;
        const T xm = lob[d] ? vfma(T(s0.dep0), xc, T(s0.dep1) * x1m) : xmr[d];
        const T xp = hib[d] ? vfma(T(s1.dep0), xc, T(s1.dep1) * x1p) : xpr[d];
        const T fm = lob[d] ? T(1) - T(s0.dep0) : T(1);
        const T fp = hib[d] ? T(1) - T(s1.dep0) : T(1);
        const T h = T(v.beta) * T(v.h2inv[d]);
        // t = bp (xp - x) - bm (x - xm), acc = alpha x - sum_d h_d t_d
        const T t = vfma(bp[d], xp - xc, -(bm[d] * (xc - xm)));
        const T s = vfma(bp[d], fp, bm[d] * fm);
        if (d == 0)
        {
            acc = vfma(T(v.alpha_a), xc, -(h * t));
            dg = vfma(h, s, T(v.alpha_a));
        } else
        {
            acc = vfma(-h, t, acc);
            dg = vfma(h, s, dg);
        }
    }
    Lx = acc;
    diag = dg;
}

// stencil_b takes the component's own x (xb); c selects the coefficients.
template <int DIM, int KIND, class T>
__device__ __forceinline__ void stencil_b(const MGView<T>& v, const SideBC (*bc)[2], int periodic_mask,
                                          const T* __restrict__ xb, int c, int i, int j, int k, T& Lx, T& diag)
{
    const long long base = (long long)c * v.Nl;
    const int sy = v.n[0], sz = v.n[0] * v.n[1];
    const int ic = i + sy * j + sz * k;
    const T* __restrict__ ib = nullptr;
    if constexpr (KIND == CK_CELLINV) ib = v.iD + base;
    // gather: indices first, then all loads
    int im[DIM], ip[DIM];
    bool lob[DIM], hib[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
        const int I = (d == 0) ? i : (d == 1 ? j : k);
        const int n = v.n[d];
        const int st = (d == 0) ? 1 : (d == 1 ? sy : sz);
        const bool per = (periodic_mask >> d) & 1;
        lob[d] = (I == 0) && !per;
        hib[d] = (I == n - 1) && !per;
        im[d] = (I > 0) ? ic - st : (per ? ic + (n - 1) * st : ic);
        ip[d] = (I < n - 1) ? ic + st : (per ? ic - (n - 1) * st : ic);
    }
    const T xc = xb[ic];
    T xmr[DIM], xpr[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
        xmr[d] = xb[im[d]];
        xpr[d] = xb[ip[d]];
    }
    T iDc = T(0), iDm[DIM], iDp[DIM];
    if constexpr (KIND == CK_CELLINV)
    {
        iDc = ib[ic];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
        {
            iDm[d] = ib[im[d]];
            iDp[d] = ib[ip[d]];
        }
    }
    T bm[DIM], bp[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
        if constexpr (KIND == CK_CONST)
        {
            bm[d] = bp[d] = T(v.bcomp ? v.bcomp[c] : v.bconst);
        } else if constexpr (KIND == CK_CELLINV)
        {
            bm[d] = T(2) / (iDc + iDm[d]);
            bp[d] = T(2) / (iDc + iDp[d]);
        } else
        {
            const int e0 = v.n[0] + (d == 0), e1 = v.n[1] + (d == 1);
            const int flo = i + e0 * (j + e1 * k);
            const int fhi = flo + ((d == 0) ? 1 : (d == 1 ? e0 : e0 * e1));
            const T* bfb = v.bf[d] + (long long)c * v.nf[d];
            bm[d] = bfb[flo];
            bp[d] = bfb[fhi];
        }
    }
    if (v.rz)
    {
        double rm, rp;
        rz_factors(v.rz_lo, v.rz_dx, i, rm, rp);
        bm[0] *= T(rm);
        bp[0] *= T(rp);
    }
    stencil_core<DIM, T>(v, bc, xc, xmr, xpr, lob, hib, bm, bp, Lx, diag);
}

template <int DIM, int KIND, class T>
__device__ __forceinline__ void stencil(const MGView<T>& v, const SideBC (*bc)[2], int periodic_mask,
                                        const T* __restrict__ x, int c, int i, int j, int k, T& Lx, T& diag)
{
    stencil_b<DIM, KIND, T>(v, bc, periodic_mask, x + (long long)c * v.Nl, c, i, j, k, Lx, diag);
}

// Diagonal of the operator row at one cell (no x reads).
template <int DIM, int KIND, class T>
__device__ inline T stencil_diag(const MGView<T>& v, const SideBC (*bc)[2], int periodic_mask, int c, int i, int j,
                                 int k)
{
    const long long base = (long long)c * v.Nl;
    const long long sy = v.n[0], sz = (long long)v.n[0] * v.n[1];
    const long long ic = (long long)i + sy * j + sz * k;
    T dg = T(v.alpha_a);
    T iDc = T(0);
    if constexpr (KIND == CK_CELLINV) iDc = v.iD[base + ic];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
    {
        const int I = (d == 0) ? i : (d == 1 ? j : k);
        const int n = v.n[d];
        const long long st = (d == 0) ? 1 : (d == 1 ? sy : sz);
        const bool per = (periodic_mask >> d) & 1;
        T fm = T(1), fp = T(1);
        long long im = ic, ip = ic;
        if (I > 0) im = ic - st;
        else if (per) im = ic + (long long)(n - 1) * st;
        else fm = T(1) - T(bc[d][0].dep0);
        if (I < n - 1) ip = ic + st;
        else if (per) ip = ic - (long long)(n - 1) * st;
        else fp = T(1) - T(bc[d][1].dep0);
        T bm, bp;
        if constexpr (KIND == CK_CONST)
        {
            bm = bp = T(v.bcomp ? v.bcomp[c] : v.bconst);
        } else if constexpr (KIND == CK_CELLINV)
        {
            bm = T(2) / (iDc + v.iD[base + im]);
            bp = T(2) / (iDc + v.iD[base + ip]);
        } else
        {
            const long long e0 = v.n[0] + (d == 0), e1 = v.n[1] + (d == 1);
            const long long flo = (long long)i + e0 * ((long long)j + e1 * (long long)k);
            const long long fhi = flo + ((d == 0) ? 1 : (d == 1 ? e0 : e0 * e1));
            bm = v.bf[d][(long long)c * v.nf[d] + flo];
            bp = v.bf[d][(long long)c * v.nf[d] + fhi];
        }
        if (d == 0 && v.rz)
        {
            double rm, rp;
            rz_factors(v.rz_lo, v.rz_dx, i, rm, rp);
            bm *= T(rm);
            bp *= T(rp);
        }
        const T h = T(v.beta) * T(v.h2inv[d]);
        dg = vfma(h, vfma(bp, fp, bm * fm), dg);   // stencil_core's rounding
    }
    return dg;
}

// One colour of a red-black sweep.  zero_first: the sweep starts from x = 0
// (the first pre-smoothing sweep of a V-cycle level): colour-0 cells get
// f/diag (what x += (f - L 0)/diag gives exactly) and their colour-1
// partners are set to zero, so no separate clearing pass is needed.
template <int DIM, int KIND, class T>
__global__ void __launch_bounds__(256) mg_gsrb_kernel(MGView<T> v, const LevelBC* __restrict__ bcs, int per_mask,
                                                      T* __restrict__ x, const T* __restrict__ f, int color,
                                                      unsigned active, int zero_first)
{
    const dim3 bI_ = xcd_block();
    // grid z enumerates (k, component): one component per thread
    const int c = bI_.z / v.n[2];
    const int k = bI_.z - c * v.n[2];
    if (!((active >> c) & 1u)) return;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    if (j >= v.n[1]) return;
    const int i = 2 * (bI_.x * blockDim.x + threadIdx.x) + ((j + k + color) & 1);
    const int ip = i ^ 1;  // the other cell of this thread's pair
    const long long row = (long long)c * v.Nl + (long long)v.n[0] * ((long long)j + (long long)v.n[1] * k);
    if (zero_first)
    {
        if (ip < v.n[0]) x[row + ip] = T(0);
        if (i >= v.n[0]) return;
        T dg = stencil_diag<DIM, KIND, T>(v, bcs->s[c], per_mask, c, i, j, k);
        x[row + i] = gs_update(T(0), f[row + i] - T(0), T(1) / dg);
        return;
    }
    if (i >= v.n[0]) return;
    const T fv = f[row + i];
    T Lx, dg;
    stencil<DIM, KIND, T>(v, bcs->s[c], per_mask, x, c, i, j, k, Lx, dg);
    x[row + i] = gs_update(x[row + i], fv - Lx, T(1) / dg);
}

// Coarse rhs = average of the 2^D fine residuals f - L x (fused).
template <int DIM, int KIND, class T>
__global__ void __launch_bounds__(256) mg_resid_restrict_kernel(MGView<T> vf, const LevelBC* __restrict__ bcs,
                                                                int per_mask, const T* __restrict__ x,
                                                                const T* __restrict__ f, int nc0, int nc1, int nc2,
                                                                long long Nc, T* __restrict__ fc, unsigned active)
{
    const dim3 bI_ = xcd_block();
    const int c = bI_.z / nc2;
    const int k = bI_.z - c * nc2;
    if (!((active >> c) & 1u)) return;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    if (i >= nc0 || j >= nc1) return;
    const long long icc = (long long)i + (long long)nc0 * ((long long)j + (long long)nc1 * k);
    T s = T(0);
#pragma unroll
    for (int kk = 0; kk < (DIM > 2 ? 2 : 1); ++kk)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
            {
                const int fi = 2 * i + ii, fj = 2 * j + jj, fk = (DIM > 2) ? 2 * k + kk : 0;
                T Lx, dg;
                stencil<DIM, KIND, T>(vf, bcs->s[c], per_mask, x, c, fi, fj, fk, Lx, dg);
                const long long o =
                    (long long)c * vf.Nl + (long long)fi + (long long)vf.n[0] * ((long long)fj + (long long)vf.n[1] * fk);
                s += f[o] - Lx;
            }
    fc[(long long)c * Nc + icc] = s / T(DIM > 2 ? 8 : 4);
}

// Coarse correction value at (I,J,K) with indices up to one cell outside
// the level: out-of-range directions are resolved in the order x, y, z with
// the homogeneous ghost model ghost = dep0*x_b + dep1*x_b2 (periodic: image).
template <int DIM, class T>
__device__ inline T cval_z(const MGView<T>& v, const SideBC (*bc)[2], int pm, const T* xb, int I, int J, int K)
{
    if (DIM > 2 && (K < 0 || K >= v.n[2]))
    {
        const int side = (K < 0) ? 0 : 1;
        const int n = v.n[2];
        const long long r = (long long)I + (long long)v.n[0] * J;
        const long long sz = (long long)v.n[0] * v.n[1];
        if ((pm >> 2) & 1) return xb[r + sz * (side ? 0 : n - 1)];
        const int kb = side ? n - 1 : 0;
        int kb2 = side ? n - 2 : 1;
        if (kb2 < 0 || kb2 >= n) kb2 = kb;
        const SideBC& s = bc[2][side];
        return vfma(T(s.dep0), xb[r + sz * kb], T(s.dep1) * xb[r + sz * kb2]);
    }
    return xb[(long long)I + (long long)v.n[0] * ((long long)J + (long long)v.n[1] * K)];
}
template <int DIM, class T>
__device__ inline T cval_y(const MGView<T>& v, const SideBC (*bc)[2], int pm, const T* xb, int I, int J, int K)
{
    if (DIM > 1 && (J < 0 || J >= v.n[1]))
    {
        const int side = (J < 0) ? 0 : 1;
        const int n = v.n[1];
        if ((pm >> 1) & 1) return cval_z<DIM, T>(v, bc, pm, xb, I, side ? 0 : n - 1, K);
        const int jb = side ? n - 1 : 0;
        int jb2 = side ? n - 2 : 1;
        if (jb2 < 0 || jb2 >= n) jb2 = jb;
        const SideBC& s = bc[1][side];
        return vfma(T(s.dep0), cval_z<DIM, T>(v, bc, pm, xb, I, jb, K), T(s.dep1) * cval_z<DIM, T>(v, bc, pm, xb, I, jb2, K));
    }
    return cval_z<DIM, T>(v, bc, pm, xb, I, J, K);
}
template <int DIM, class T>
__device__ inline T cval_x(const MGView<T>& v, const SideBC (*bc)[2], int pm, const T* xb, int I, int J, int K)
{
    if (I < 0 || I >= v.n[0])
    {
        const int side = (I < 0) ? 0 : 1;
        const int n = v.n[0];
        if (pm & 1) return cval_y<DIM, T>(v, bc, pm, xb, side ? 0 : n - 1, J, K);
        const int ib = side ? n - 1 : 0;
        int ib2 = side ? n - 2 : 1;
        if (ib2 < 0 || ib2 >= n) ib2 = ib;
        const SideBC& s = bc[0][side];
        return vfma(T(s.dep0), cval_y<DIM, T>(v, bc, pm, xb, ib, J, K), T(s.dep1) * cval_y<DIM, T>(v, bc, pm, xb, ib2, J, K));
    }
    return cval_y<DIM, T>(v, bc, pm, xb, I, J, K);
}

// Cell-centred (bi/tri)linear interpolation of the coarse correction at
// fine cell (i,j,k): weights 3/4, 1/4 per direction toward the nearer coarse
// neighbours, summed in (z, y, x) offset-bit order.
template <int DIM, class T>
__device__ inline T prolong_value(const MGView<T>& vc, const SideBC (*bc)[2], int pm, const T* xcb, int i, int j, int k)
{
    const int I = i >> 1, J = (DIM > 1) ? (j >> 1) : 0, K = (DIM > 2) ? (k >> 1) : 0;
    const int sx = (i & 1) ? 1 : -1, sy = (j & 1) ? 1 : -1, sz = (k & 1) ? 1 : -1;
    T acc = T(0);
    bool inner = (I + sx >= 0) && (I + sx < vc.n[0]);
    if (DIM > 1) inner = inner && (J + sy >= 0) && (J + sy < vc.n[1]);
    if (DIM > 2) inner = inner && (K + sz >= 0) && (K + sz < vc.n[2]);
    if (inner)
    {
        // all eight coarse neighbours inside the level: direct loads
        const int cy = vc.n[0], cz = vc.n[0] * vc.n[1];
        const int o = I + cy * J + cz * K;
#pragma unroll
        for (int cc = 0; cc < (DIM > 2 ? 2 : 1); ++cc)
#pragma unroll
            for (int bb = 0; bb < (DIM > 1 ? 2 : 1); ++bb)
#pragma unroll
                for (int aa = 0; aa < 2; ++aa)
                {
                    T w = aa ? T(0.25) : T(0.75);
                    if (DIM > 1) w *= bb ? T(0.25) : T(0.75);
                    if (DIM > 2) w *= cc ? T(0.25) : T(0.75);
                    acc += w * xcb[o + aa * sx + bb * sy * cy + cc * sz * cz];
                }
        return acc;
    }
#pragma unroll
    for (int cc = 0; cc < (DIM > 2 ? 2 : 1); ++cc)
#pragma unroll
        for (int bb = 0; bb < (DIM > 1 ? 2 : 1); ++bb)
#pragma unroll
            for (int aa = 0; aa < 2; ++aa)
            {
                T w = aa ? T(0.25) : T(0.75);
                if (DIM > 1) w *= bb ? T(0.25) : T(0.75);
                if (DIM > 2) w *= cc ? T(0.25) : T(0.75);
                acc += w * cval_x<DIM, T>(vc, bc, pm, xcb, I + aa * sx, J + bb * sy, K + cc * sz);
            }
    return acc;
}

// x_f += P x_c (linear prolongation): one thread per coarse cell adds the
// interpolated correction to its 2^D children, reading the 3^D coarse
// neighbourhood once.  Each child's value is the same weighted sum, in the
// same order, as prolong_value().
template <int DIM, class T>
__global__ void __launch_bounds__(256) mg_prolong_kernel(int nf0, int nf1, int nf2, long long Nf, T* __restrict__ xf,
                                                         MGView<T> vc, const LevelBC* __restrict__ bcc, int pm,
                                                         const T* __restrict__ xc, unsigned active)
{
    const dim3 bI_ = xcd_block();
    const int c = bI_.z / vc.n[2];
    const int K = bI_.z - c * vc.n[2];
    if (!((active >> c) & 1u)) return;
    const int I = bI_.x * blockDim.x + threadIdx.x;
    const int J = bI_.y * blockDim.y + threadIdx.y;
    if (I >= vc.n[0] || J >= vc.n[1]) return;
    const T* xcb = xc + (long long)c * vc.Nl;
    T nb[DIM > 2 ? 3 : 1][DIM > 1 ? 3 : 1][3];
    bool inner = (I >= 1) && (I + 1 < vc.n[0]);
    if (DIM > 1) inner = inner && (J >= 1) && (J + 1 < vc.n[1]);
    if (DIM > 2) inner = inner && (K >= 1) && (K + 1 < vc.n[2]);
    const int cy = vc.n[0], cz = vc.n[0] * vc.n[1];
    if (inner)
    {
        // straight-line block: all neighbourhood loads in flight together
        const int o = I + cy * J + cz * K;
#pragma unroll
        for (int z = 0; z < (DIM > 2 ? 3 : 1); ++z)
#pragma unroll
            for (int y = 0; y < (DIM > 1 ? 3 : 1); ++y)
#pragma unroll
                for (int x = 0; x < 3; ++x)
                {
                    const int dz = (DIM > 2) ? z - 1 : 0, dy = (DIM > 1) ? y - 1 : 0, dxo = x - 1;
                    nb[z][y][x] = xcb[o + dxo + cy * dy + cz * dz];
                }
    } else
    {
#pragma unroll
        for (int z = 0; z < (DIM > 2 ? 3 : 1); ++z)
#pragma unroll
            for (int y = 0; y < (DIM > 1 ? 3 : 1); ++y)
#pragma unroll
                for (int x = 0; x < 3; ++x)
                {
                    const int dz = (DIM > 2) ? z - 1 : 0, dy = (DIM > 1) ? y - 1 : 0, dxo = x - 1;
                    nb[z][y][x] = cval_x<DIM, T>(vc, bcc->s[c], pm, xcb, I + dxo, J + dy, K + dz);
                }
    }
#pragma unroll
    for (int kk = 0; kk < (DIM > 2 ? 2 : 1); ++kk)
#pragma unroll
        for (int jj = 0; jj < (DIM > 1 ? 2 : 1); ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
            {
                const int i = 2 * I + ii, j = (DIM > 1) ? 2 * J + jj : 0, k = (DIM > 2) ? 2 * K + kk : 0;
                if (i >= nf0 || j >= nf1 || k >= nf2) continue;
                const int sx = ii ? 1 : -1, sy = jj ? 1 : -1, sz = kk ? 1 : -1;
                T acc = T(0);
#pragma unroll
                for (int cc = 0; cc < (DIM > 2 ? 2 : 1); ++cc)
#pragma unroll
                    for (int bb = 0; bb < (DIM > 1 ? 2 : 1); ++bb)
#pragma unroll
                        for (int aa = 0; aa < 2; ++aa)
                        {
                            T w = aa ? T(0.25) : T(0.75);
                            if (DIM > 1) w *= bb ? T(0.25) : T(0.75);
                            if (DIM > 2) w *= cc ? T(0.25) : T(0.75);
                            acc += w * nb[(DIM > 2) ? 1 + cc * sz : 0][(DIM > 1) ? 1 + bb * sy : 0][1 + aa * sx];
                        }
                const long long o = (long long)i + (long long)nf0 * ((long long)j + (long long)nf1 * k);
                xf[(long long)c * Nf + o] += acc;
            }
}

// x += e (solution update after a cycle); grid y enumerates components.
template <class T, class S>
__global__ void __launch_bounds__(256) mg_add_kernel(long long N, S* __restrict__ x, const T* __restrict__ e,
                                                     unsigned active, long long xstride_comp, const int* __restrict__ compmap)
{
    const dim3 bI_ = xcd_block();
    const int c = bI_.y;
    if (!((active >> c) & 1u)) return;
    const long long q = (long long)bI_.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const long long xo = (long long)compmap[c] * xstride_comp + q;
    x[xo] = x[xo] + S(e[(long long)c * N + q]);
}

template <class T>
__global__ void __launch_bounds__(256) mg_zero_kernel(long long N, T* __restrict__ x, unsigned active)
{
    const int c = blockIdx.y;
    if (!((active >> c) & 1u)) return;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < N) x[(long long)c * N + q] = T(0);
}

// Bottom solve in one workgroup: e = 0, red-black sweeps until every active
// component has |r|_inf <= 1e-4 |r0|_inf (or r0 = 0), at most 200 sweeps.
template <int DIM, int KIND, class T>
__global__ void __launch_bounds__(1024) mg_bottom_kernel(MGView<T> v, const LevelBC* __restrict__ bcs, int per_mask,
                                                         T* __restrict__ x, const T* __restrict__ f,
                                                         unsigned active, int ncomp)
{
    __shared__ double red[32][kMaxComp];
    __shared__ double n0[kMaxComp], nn[kMaxComp];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6, nw = (nt + 63) >> 6;
    const long long Nl = v.Nl;
    for (long long q = tid; q < Nl; q += nt)
        for (int c = 0; c < ncomp; ++c)
            if ((active >> c) & 1u) x[(long long)c * Nl + q] = T(0);
    __syncthreads();
    auto resnorm = [&](double* out) {
        double loc[kMaxComp];
        for (int c = 0; c < ncomp; ++c) loc[c] = 0.0;
        for (long long q = tid; q < Nl; q += nt)
        {
            const int i = int(q % v.n[0]);
            const int j = int((q / v.n[0]) % v.n[1]);
            const int k = int(q / ((long long)v.n[0] * v.n[1]));
            for (int c = 0; c < ncomp; ++c)
            {
                if (!((active >> c) & 1u)) continue;
                T Lx, dg;
                stencil<DIM, KIND, T>(v, bcs->s[c], per_mask, x, c, i, j, k, Lx, dg);
                loc[c] = vmax(loc[c], fabs((double)(f[(long long)c * Nl + q] - Lx)));
            }
        }
        for (int c = 0; c < ncomp; ++c)
        {
            double w = wave_max(loc[c]);
            if (lane == 0) red[wid][c] = w;
        }
        __syncthreads();
        if (tid < ncomp)
        {
            double w = 0.0;
            for (int q = 0; q < nw; ++q) w = vmax(w, red[q][tid]);
            out[tid] = w;
        }
        __syncthreads();
    };
    auto sweep = [&]() {
        for (int color = 0; color < 2; ++color)
        {
            for (long long q = tid; q < Nl; q += nt)
            {
                const int i = int(q % v.n[0]);
                const int j = int((q / v.n[0]) % v.n[1]);
                const int k = int(q / ((long long)v.n[0] * v.n[1]));
                if (((i + j + k) & 1) != color) continue;
                for (int c = 0; c < ncomp; ++c)
                {
                    if (!((active >> c) & 1u)) continue;
                    T Lx, dg;
                    stencil<DIM, KIND, T>(v, bcs->s[c], per_mask, x, c, i, j, k, Lx, dg);
                    const long long o = (long long)c * Nl + q;
                    x[o] = gs_update(x[o], f[o] - Lx, T(1) / dg);
                }
            }
            __threadfence_block();
            __syncthreads();
        }
    };
    resnorm(n0);
    for (int it = 0; it < 200; ++it)
    {
        bool done = true;
        for (int c = 0; c < ncomp; ++c)
        {
            if (!((active >> c) & 1u)) continue;
            if (it == 0)
            {
                if (n0[c] != 0.0) done = false;
            } else if (!(n0[c] == 0.0 || nn[c] <= 1e-4 * n0[c]))
            {
                done = false;
            }
        }
        if (done) break;
        sweep();
        resnorm(nn);
    }
}

// ---------------------------------------------------------------------------
// Coarse part of a V-cycle, one workgroup per component, levels l0..L-1
// resident in LDS: the workgroup loads the rhs of level l0, goes down
// (pre-smoothing, residual + restriction), through the bottom solve and back
// up (prolongation, post-smoothing) with workgroup barriers between phases,
// and writes the correction of level l0 back.  The arithmetic of every phase
// is exactly that of the per-level kernels above.
constexpr int kMaxLev = 12;

template <class T>
struct CoarseArgs
{
    int nlev;
    int per_mask;
    int nu1, nu2;
    int ncomp;
    MGView<T> v[kMaxLev];
    T* x[kMaxLev];    // level l0 entry used (correction out); deeper levels live in LDS
    T* f[kMaxLev];    // level l0 entry used (rhs in)
    const LevelBC* bc[kMaxLev];
};

// LDS words (of T) the coarse cycle needs for levels of the given sizes
inline long long coarse_lds_words(const long long* N, int nlev)
{
    long long w = 0;
    for (int l = 0; l < nlev; ++l) w += 2 * N[l];
    return w;
}

__device__ inline void decode_cell(unsigned q, unsigned n0, unsigned n1, int& i, int& j, int& k)
{
    const unsigned r = q / n0;
    i = int(q - r * n0);
    k = int(r / n1);
    j = int(r - (unsigned)k * n1);
}

// Phases of the coarse cycle run either on the whole workgroup (W = false,
// __syncthreads between phases) or, for levels of at most 64 cells, on its
// first wave alone (W = true).  A wave's LDS operations complete in issue
// order, so between wave-only phases a compiler barrier with
// wavefront-scope fences orders the lanes' LDS accesses.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool W>
__device__ inline void ls_sync()
{
    if constexpr (W) wave_sync();
    else __syncthreads();
}
template <bool W>
__device__ inline unsigned ls_tid()
{
    return W ? (threadIdx.x & 63u) : threadIdx.x;
}
template <bool W>
__device__ inline unsigned ls_nt()
{
    return W ? 64u : blockDim.x;
}
constexpr long long kWaveCells = 64;   // largest level the first wave handles alone

// One colour of a red-black sweep over an LDS-resident level: one thread per
// x-pair of cells, as in mg_gsrb_kernel.
template <int DIM, int KIND, class T, bool W>
__device__ void ls_sweep(const MGView<T>& v, const SideBC (*bc)[2], int pm, int c, T* x, const T* f, int color,
                         bool zero_first)
{
    const unsigned n0 = (unsigned)v.n[0], nh = (n0 + 1) >> 1, n1 = (unsigned)v.n[1];
    const unsigned np = nh * n1 * (unsigned)v.n[2];
    for (unsigned p = ls_tid<W>(); p < np; p += ls_nt<W>())
    {
        int i2, j, k;
        decode_cell(p, nh, n1, i2, j, k);
        const int i = 2 * i2 + ((j + k + color) & 1);
        const int ip = i ^ 1;
        const int row = v.n[0] * (j + v.n[1] * k);
        if (zero_first)
        {
            if (ip < v.n[0]) x[row + ip] = T(0);
            if (i >= v.n[0]) continue;
            T dg = stencil_diag<DIM, KIND, T>(v, bc, pm, c, i, j, k);
            x[row + i] = gs_update(T(0), f[row + i] - T(0), T(1) / dg);
            continue;
        }
        if (i >= v.n[0]) continue;
        T Lx, dg;
        stencil_b<DIM, KIND, T>(v, bc, pm, x, c, i, j, k, Lx, dg);
        x[row + i] = gs_update(x[row + i], f[row + i] - Lx, T(1) / dg);
    }
    ls_sync<W>();
}

// max |f - L x| over an LDS-resident level, the same value in every thread
// (of the wave, W)
template <int DIM, int KIND, class T, bool W>
__device__ double ls_resnorm(const MGView<T>& v, const SideBC (*bc)[2], int pm, int c, const T* x, const T* f,
                             double* red)
{
    double m = 0.0;
    for (unsigned q = ls_tid<W>(); q < (unsigned)v.Nl; q += ls_nt<W>())
    {
        int i, j, k;
        decode_cell(q, (unsigned)v.n[0], (unsigned)v.n[1], i, j, k);
        T Lx, dg;
        stencil_b<DIM, KIND, T>(v, bc, pm, x, c, i, j, k, Lx, dg);
        m = vmax(m, fabs((double)(f[q] - Lx)));
    }
    m = wave_max(m);
    if constexpr (W) return m;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) red[wid] = m;
    __syncthreads();
    double w = 0.0;
    for (int q = 0; q < nw; ++q) w = vmax(w, red[q]);
    __syncthreads();
    return w;
}

// Bottom solve: x = 0, red-black sweeps until |r| <= 1e-4 |r0| (or r0 = 0),
// at most 200 sweeps.
template <int DIM, int KIND, class T, bool W>
__device__ void ls_bottom(const MGView<T>& v, const SideBC (*bc)[2], int pm, int c, T* x, const T* f, double* red)
{
    for (unsigned q = ls_tid<W>(); q < (unsigned)v.Nl; q += ls_nt<W>()) x[q] = T(0);
    ls_sync<W>();
    const double r0 = ls_resnorm<DIM, KIND, T, W>(v, bc, pm, c, x, f, red);
    if (r0 == 0.0) return;
    for (int it = 0; it < 200; ++it)
    {
        ls_sweep<DIM, KIND, T, W>(v, bc, pm, c, x, f, 0, false);
        ls_sweep<DIM, KIND, T, W>(v, bc, pm, c, x, f, 1, false);
        const double rn = ls_resnorm<DIM, KIND, T, W>(v, bc, pm, c, x, f, red);
        if (rn <= 1e-4 * r0) break;
    }
}

// LDS layout of the coarse cycle: level l holds x_l then f_l, levels back to
// back
template <class T>
__device__ inline T* ls_level(const CoarseArgs<T>* Ap, T* lds, int l)
{
    unsigned o = 0;
    for (int q = 0; q < l; ++q) o += 2u * (unsigned)Ap->v[q].Nl;
    return lds + o;
}

// Down the cycle at level l: nu1 sweeps from zero, residual restricted into
// f[l+1].
template <int DIM, int KIND, class T, bool W>
__device__ void ls_down(const CoarseArgs<T>* Ap, const SideBC (*sbc)[3][2], T* lds, int l, int c)
{
    const int pm = Ap->per_mask, nu1 = Ap->nu1;
    const MGView<T> v = Ap->v[l];
    const MGView<T> vc = Ap->v[l + 1];
    const SideBC(*bc)[2] = sbc[l];
    T* x = ls_level(Ap, lds, l);
    const T* f = x + (unsigned)v.Nl;
    T* fc = ls_level(Ap, lds, l + 1) + (unsigned)vc.Nl;
    for (int s = 0; s < nu1; ++s)
    {
        ls_sweep<DIM, KIND, T, W>(v, bc, pm, c, x, f, 0, s == 0);
        ls_sweep<DIM, KIND, T, W>(v, bc, pm, c, x, f, 1, false);
    }
    if (nu1 == 0)
    {
        for (unsigned q = ls_tid<W>(); q < (unsigned)v.Nl; q += ls_nt<W>()) x[q] = T(0);
        ls_sync<W>();
    }
    for (unsigned q = ls_tid<W>(); q < (unsigned)vc.Nl; q += ls_nt<W>())
    {
        int i, j, k;
        decode_cell(q, (unsigned)vc.n[0], (unsigned)vc.n[1], i, j, k);
        T sum = T(0);
        for (int kk = 0; kk < (DIM > 2 ? 2 : 1); ++kk)
            for (int jj = 0; jj < 2; ++jj)
                for (int ii = 0; ii < 2; ++ii)
                {
                    const int fi = 2 * i + ii, fj = 2 * j + jj, fk = (DIM > 2) ? 2 * k + kk : 0;
                    T Lx, dg;
                    stencil_b<DIM, KIND, T>(v, bc, pm, x, c, fi, fj, fk, Lx, dg);
                    sum += f[fi + v.n[0] * (fj + v.n[1] * fk)] - Lx;
                }
        fc[q] = sum / T(DIM > 2 ? 8 : 4);
    }
    ls_sync<W>();
}

// Up the cycle at level l: x += P x[l+1], nu2 sweeps.
template <int DIM, int KIND, class T, bool W>
__device__ void ls_up(const CoarseArgs<T>* Ap, const SideBC (*sbc)[3][2], T* lds, int l, int c)
{
    const int pm = Ap->per_mask, nu2 = Ap->nu2;
    const MGView<T> v = Ap->v[l];
    const MGView<T> vc = Ap->v[l + 1];
    const SideBC(*bc)[2] = sbc[l];
    const SideBC(*bcc)[2] = sbc[l + 1];
    T* x = ls_level(Ap, lds, l);
    const T* f = x + (unsigned)v.Nl;
    const T* xc = ls_level(Ap, lds, l + 1);
    for (unsigned q = ls_tid<W>(); q < (unsigned)v.Nl; q += ls_nt<W>())
    {
        int i, j, k;
        decode_cell(q, (unsigned)v.n[0], (unsigned)v.n[1], i, j, k);
        x[q] += prolong_value<DIM, T>(vc, bcc, pm, xc, i, j, k);
    }
    ls_sync<W>();
    for (int s = 0; s < nu2; ++s)
    {
        ls_sweep<DIM, KIND, T, W>(v, bc, pm, c, x, f, 0, false);
        ls_sweep<DIM, KIND, T, W>(v, bc, pm, c, x, f, 1, false);
    }
}

// 256 threads: enough for the levels it takes (at most kFuseCells cells, so
// at most 512 red-black pairs) and room for the registers of its phases
// without spilling
constexpr int kCoarseThreads = 256;
template <int DIM, int KIND, class T>
__global__ void __launch_bounds__(kCoarseThreads) mg_coarse_cycle_kernel(const CoarseArgs<T>* __restrict__ Ap, unsigned act)
{
    extern __shared__ unsigned char vd_coarse_lds[];
    __shared__ double red[16];
    // side models of every level for this component, read once: the
    // sweeps evaluate them in every phase
    __shared__ SideBC sbc[kMaxLev][3][2];
    const int c = blockIdx.x;
    if (!((act >> c) & 1u)) return;
    T* lds = reinterpret_cast<T*>(vd_coarse_lds);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int L = Ap->nlev, pm = Ap->per_mask;
    for (int q = tid; q < L * 6; q += nt)
    {
        const int l = q / 6, r = q - 6 * (q / 6);
        sbc[l][r >> 1][r & 1] = Ap->bc[l]->s[c][r >> 1][r & 1];
    }
    {
        const unsigned N0 = (unsigned)Ap->v[0].Nl;
        const T* __restrict__ fg = Ap->f[0] + (long long)c * N0;
        T* f0 = lds + N0;
        for (unsigned q = tid; q < N0; q += nt) f0[q] = fg[q];
        __syncthreads();
    }
    // levels lw.. (at most kWaveCells cells) run on the first wave alone
    int lw = L;
    for (int l = L - 1; l >= 0 && Ap->v[l].Nl <= kWaveCells; --l) lw = l;
    for (int l = 0; l < lw && l < L - 1; ++l) ls_down<DIM, KIND, T, false>(Ap, sbc, lds, l, c);
    if (lw < L)
    {
        if (tid < 64)
        {
            for (int l = lw; l < L - 1; ++l) ls_down<DIM, KIND, T, true>(Ap, sbc, lds, l, c);
            const MGView<T> v = Ap->v[L - 1];
            T* x = ls_level(Ap, lds, L - 1);
            ls_bottom<DIM, KIND, T, true>(v, sbc[L - 1], pm, c, x, x + (unsigned)v.Nl, red);
            for (int l = L - 2; l >= lw; --l) ls_up<DIM, KIND, T, true>(Ap, sbc, lds, l, c);
        }
        __syncthreads();
    } else
    {
        const MGView<T> v = Ap->v[L - 1];
        T* x = ls_level(Ap, lds, L - 1);
        ls_bottom<DIM, KIND, T, false>(v, sbc[L - 1], pm, c, x, x + (unsigned)v.Nl, red);
    }
    for (int l = (lw < L ? lw : L - 1) - 1; l >= 0; --l) ls_up<DIM, KIND, T, false>(Ap, sbc, lds, l, c);
    {
        const unsigned N0 = (unsigned)Ap->v[0].Nl;
        T* __restrict__ xg = Ap->x[0] + (long long)c * N0;
        for (unsigned q = tid; q < N0; q += nt) xg[q] = lds[q];
    }
}

} // namespace vd
#include "vd_mg_fused.H"
namespace vd {

// ---------------------------------------------------------------------------
// Host side: hierarchy, buffers and the solve loop.

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// Synchronising device-to-host read of a few words (norms, reductions)
// through a pinned staging buffer: a copy into pageable memory takes the
// runtime's staged path, which lengthens every read the solver loop waits
// on.  One buffer per host thread, kept until the process ends.
// Coarsening rule of the multigrid hierarchy: halve while every extent
// stays even and >= 2 and a periodic extent stays even (an odd periodic ring
// of three or more cells puts two neighbours in one colour, and a red-black
// sweep would read and write the same cells at once).
inline bool mg_can_coarsen(const int n[3], int dim, int per_mask)
{
    for (int d = 0; d < dim; ++d)
    {
        if (n[d] % 2 != 0 || n[d] / 2 < 2) return false;
        if (((per_mask >> d) & 1) && (n[d] / 2) % 2 != 0) return false;
    }
    return true;
}

// The bottom solve is at most 200 red-black sweeps on the coarsest level; a
// direction of 32 cells still loses about 85% of its smoothest error mode per
// bottom solve, one of 125 cells only 10%, and the V-cycle would stall.
constexpr int kMaxCoarsestExtent = 32;

// Extents of the coarsest level a grid coarsens to.
inline void mg_coarsest_extents(const int n0[3], int dim, int per_mask, int max_coarsening, int out[3])
{
    int n[3] = {n0[0], dim > 1 ? n0[1] : 1, dim > 2 ? n0[2] : 1};
    for (int l = 0; l < max_coarsening && mg_can_coarsen(n, dim, per_mask); ++l)
        for (int d = 0; d < dim; ++d) n[d] /= 2;
    for (int d = 0; d < 3; ++d) out[d] = n[d];
}

// Norms and maxima read back from the device carry a NaN of any field they
// reduce (vmax); a non-finite one is reported rather than compared.
// mask: the components whose values the last launches computed (the slots of
// the others may hold anything, e.g. never-written memory).
inline void require_finite(const double* v, int n, const char* what, unsigned mask = ~0u)
{
    for (int i = 0; i < n; ++i)
        if (((mask >> i) & 1u) && !std::isfinite(v[i]))
            throw std::runtime_error(std::string("non-finite ") + what + " (NaN or Inf in the state or right-hand side)");
}

inline void d2h_sync(void* dst, const void* src, size_t bytes, hipStream_t stream, const char* what)
{
    static thread_local void* pin = nullptr;
    static thread_local size_t cap = 0;
    if (bytes > cap)
    {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes, 4096);
        hip_check(hipHostMalloc(&pin, want, hipHostMallocDefault), "pinned staging buffer");
        cap = want;
    }
    hip_check(hipMemcpyAsync(pin, src, bytes, hipMemcpyDeviceToHost, stream), what);
    hip_check(hipStreamSynchronize(stream), what);
    std::memcpy(dst, pin, bytes);
}

template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        free();
        n = count;
        if (count) hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    }
    void free()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o)
        {
            free();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
};

struct MGLevelHost
{
    int n[3];
    double dx[3];
    long long N;
    long long nf[3];
};

struct SideSpec
{
    int kind;       // BcKind
    double ra, rb;  // Robin a, b (constant over the side)
};

// One solver instance (fixed geometry, kind and component count).
template <int DIM, class T>
class MGSolver
{
public:
    // axisymmetric metric (rz_factors) with the domain's low radius
    int rz = 0;
    double rz_lo = 0.0;
    int nc = 1;
    int kind = CK_CONST;
    int maxorder = 2;
    double alpha_a = 0.0, beta = 1.0, bconst = -1.0;
    // components whose face coefficient is spatially uniform run on the
    // constant-coefficient kernels with b = bcomp[c]
    unsigned uniform_mask = 0;
    DevBuf<double> bcomp_dev;
    std::vector<double> bcomp_h;
    int nu1 = 2, nu2 = 2, maxiter = 100;
    int per_mask = 0;
    std::vector<MGLevelHost> lv;
    std::vector<DevBuf<T>> ex, fx;        // per level correction and rhs (level 0: e0, r0)
    std::vector<DevBuf<T>> bface[3];      // per level >= 1 (CK_CELLINV/CK_FACE): face coefficients
    DevBuf<double> f0;                     // folded rhs (level 0, FP64)
    DevBuf<double> iD0;                    // level-0 inverse cell coefficients (CK_CELLINV, FP64)
    DevBuf<T> iD0t;                        // iD0 in the cycle precision (T = float only)
    DevBuf<LevelBC> bcdev;                 // per level
    DevBuf<unsigned long long> nrm;        // nc words
    DevBuf<int> compmap;                   // component -> state component (solution update)
    DevBuf<int> idmap;                     // identity component map
    DevBuf<double> partials;               // per-workgroup residual maxima (level 0)
    std::vector<SideSpec> sides;           // [c][d][s]
    hipStream_t stream = nullptr;
    KTimer* kt = nullptr;                  // optional per-launch timing
    int kbase = K_PGSRB0;                  // first of this solver's 5 classes
    int kfused = K_PPRE0;                  // fused pre class (post: kfused + 1)
    MGSolver() = default;
    MGSolver(const MGSolver&) = delete;
    MGSolver& operator=(const MGSolver&) = delete;
    ~MGSolver()
    {
        drop_coarse_graphs();
        if (blas) (void)rocblas_destroy_handle(blas);
    }
    int kc(int off, int l) const { return l == 0 ? kbase + off : kbase + 4; }
    // fused plane-marching smoothing passes (3D, constant coefficients,
    // nu1 = nu2 = 2); ex_alt receives the post pass and is swapped with ex
    bool fused_on = true;
    long long fused_min_cells = 1LL << 22; // smallest level x active components that runs fused
    std::vector<DevBuf<T>> ex_alt;
    DevBuf<T> xghost;                      // ghosted coarse correction (prolongation)

    void build(const int n0[3], const double dx0[3], int ncomp, int coef_kind, int max_coarsening, int per_mask_)
    {
        nc = ncomp;
        kind = coef_kind;
        per_mask = per_mask_;
        lv.clear();
        MGLevelHost l0;
        for (int d = 0; d < 3; ++d)
        {
            l0.n[d] = (d < DIM) ? n0[d] : 1;
            l0.dx[d] = dx0[d];
        }
        lv.push_back(l0);
        while ((int)lv.size() - 1 < max_coarsening)
        {
            const MGLevelHost& c = lv.back();
            if (!mg_can_coarsen(c.n, DIM, per_mask)) break;
            MGLevelHost nl;
            for (int d = 0; d < 3; ++d)
            {
                nl.n[d] = (d < DIM) ? c.n[d] / 2 : 1;
                nl.dx[d] = (d < DIM) ? c.dx[d] * 2.0 : c.dx[d];
            }
            lv.push_back(nl);
        }
        const int L = (int)lv.size();
        ex.clear();
        fx.clear();
        ex.resize(L);
        fx.resize(L);
        for (int d = 0; d < 3; ++d)
        {
            bface[d].clear();
            bface[d].resize(L);
        }
        for (int l = 0; l < L; ++l)
        {
            auto& h = lv[l];
            h.N = (long long)h.n[0] * h.n[1] * h.n[2];
            for (int d = 0; d < 3; ++d)
                h.nf[d] = (long long)(h.n[0] + (d == 0)) * (h.n[1] + (d == 1)) * (h.n[2] + (d == 2));
            ex[l].alloc(h.N * nc);
            fx[l].alloc(h.N * nc);
            if (kind != CK_CONST && l >= 1)
                for (int d = 0; d < DIM; ++d) bface[d][l].alloc(h.nf[d] * nc);
        }
        // fused passes: the post pass's output buffers and the ghosted
        // coarse correction (largest: level 1)
        ex_alt.clear();
        ex_alt.resize(L);
        if (DIM == 3 && L > 1)
        {
            for (int l = 0; l + 1 < L; ++l) ex_alt[l].alloc(lv[l].N * nc);
            xghost.alloc((size_t)(lv[1].n[0] + 2) * (lv[1].n[1] + 2) * (lv[1].n[2] + 2) * nc);
        }
        f0.alloc(lv[0].N * nc);
        if (kind == CK_CELLINV) iD0.alloc(lv[0].N * nc);
        if (kind == CK_CELLINV && !std::is_same<T, double>::value) iD0t.alloc(lv[0].N * nc);
        bcdev.alloc(L);
        {
            Launch3 L3 = cell_launch(lv[0].n);
            partials.alloc((size_t)L3.grid.x * L3.grid.y * L3.grid.z * nc);
        }
        nrm.alloc(kMaxComp);
        compmap.alloc(kMaxComp);
        idmap.alloc(kMaxComp);
        {
            std::vector<int> h(kMaxComp);
            for (int c = 0; c < kMaxComp; ++c) h[c] = c;
            hip_check(hipMemcpy(idmap.p, h.data(), sizeof(int) * kMaxComp, hipMemcpyHostToDevice), "idmap");
        }
        bcomp_dev.alloc(kMaxComp);
        bcomp_h.assign(kMaxComp, 0.0);
        sides.assign(nc * 3 * 2, SideSpec{BK_NEU, 0.0, 1.0});
    }

    void set_compmap(const std::vector<int>& m)
    {
        std::vector<int> h(kMaxComp, 0);
        for (size_t i = 0; i < m.size(); ++i) h[i] = m[i];
        hip_check(hipMemcpy(compmap.p, h.data(), sizeof(int) * kMaxComp, hipMemcpyHostToDevice), "compmap");
    }

    SideSpec& side(int c, int d, int s) { return sides[(c * 3 + d) * 2 + s]; }

    // Homogeneous ghost model of side (d, s) of component c on level l.
    SideBC side_model(int l, int c, int d, int s) const
    {
        SideBC b{BK_NEU, 1.0, 0.0};
        if (d >= DIM) return b;
        const SideSpec& sp = sides[(c * 3 + d) * 2 + s];
        b.kind = sp.kind;
        if (sp.kind == BK_DIR)
        {
            int m = std::min(lv[l].n[d] + 1, maxorder);
            if (m >= 3)
            {
                b.dep0 = -2.0;
                b.dep1 = 1.0 / 3.0;
            } else
            {
                b.dep0 = -1.0;
                b.dep1 = 0.0;
            }
        } else if (sp.kind == BK_CF)
        {
            b.dep0 = kCfD0;
            b.dep1 = kCfD1;
        } else if (sp.kind == BK_ROBIN)
        {
            double hh = lv[l].dx[d];
            b.dep0 = -(0.5 * sp.ra - sp.rb / hh) / (0.5 * sp.ra + sp.rb / hh);
            b.dep1 = 0.0;
        } else
        {
            b.dep0 = 1.0; // Neumann / inhomogeneous Neumann / periodic (unused)
            b.dep1 = 0.0;
        }
        return b;
    }

    // Homogeneous ghost model for each level from the side specs.
    void upload_bc()
    {
        std::vector<LevelBC> h(lv.size());
        for (size_t l = 0; l < lv.size(); ++l)
        {
            std::memset(&h[l], 0, sizeof(LevelBC));
            for (int c = 0; c < nc; ++c)
                for (int d = 0; d < 3; ++d)
                    for (int s = 0; s < 2; ++s) h[l].s[c][d][s] = side_model((int)l, c, d, s);
        }
        hip_check(hipMemcpyAsync(bcdev.p, h.data(), sizeof(LevelBC) * lv.size(), hipMemcpyHostToDevice, stream),
                  "upload_bc");
    }

    MGView<T> view(int l) const
    {
        MGView<T> v{};
        for (int d = 0; d < 3; ++d)
        {
            v.n[d] = lv[l].n[d];
            v.h2inv[d] = 1.0 / (lv[l].dx[d] * lv[l].dx[d]);
            v.nf[d] = lv[l].nf[d];
            v.bf[d] = (kind == CK_CONST || l == 0) ? nullptr : bface[d][l].p;
        }
        v.Nl = lv[l].N;
        v.alpha_a = alpha_a;
        v.beta = beta;
        v.bconst = bconst;
        v.bcomp = uniform_mask ? bcomp_dev.p : nullptr;
        v.rz = rz;
        v.rz_lo = rz_lo;
        v.rz_dx = lv[l].dx[0];
        if constexpr (std::is_same<T, double>::value)
            v.iD = (kind == CK_CELLINV && l == 0) ? iD0.p : nullptr;
        else
            v.iD = (kind == CK_CELLINV && l == 0) ? iD0t.p : nullptr;
        return v;
    }

    // FP64 view of the finest level (residual and solution update stay FP64)
    MGView<double> view0d() const
    {
        MGView<double> v{};
        for (int d = 0; d < 3; ++d)
        {
            v.n[d] = lv[0].n[d];
            v.h2inv[d] = 1.0 / (lv[0].dx[d] * lv[0].dx[d]);
            v.nf[d] = lv[0].nf[d];
            v.bf[d] = nullptr;
        }
        v.Nl = lv[0].N;
        v.alpha_a = alpha_a;
        v.beta = beta;
        v.bconst = bconst;
        v.bcomp = uniform_mask ? bcomp_dev.p : nullptr;
        v.iD = (kind == CK_CELLINV) ? iD0.p : nullptr;
        v.rz = rz;
        v.rz_lo = rz_lo;
        v.rz_dx = lv[0].dx[0];
        return v;
    }

    // kind used on level l
    int level_kind(int l) const
    {
        if (kind == CK_CONST) return CK_CONST;
        return l == 0 ? CK_CELLINV : CK_FACE;
    }

    // Run f(kind_tag, mask) for the uniform components (constant kernels)
    // and for the others (the level's own coefficient kind).
    template <class F>
    void dispatch(int l, unsigned act, F&& f)
    {
        const unsigned u = act & uniform_mask, o = act & ~uniform_mask;
        if (u) f(std::integral_constant<int, CK_CONST>{}, u);
        if (o)
        {
            switch (level_kind(l))
            {
            case CK_CONST: f(std::integral_constant<int, CK_CONST>{}, o); break;
            case CK_CELLINV: f(std::integral_constant<int, CK_CELLINV>{}, o); break;
            default: f(std::integral_constant<int, CK_FACE>{}, o); break;
            }
        }
    }

    // Declare components uniform (b = vals[c] for c in mask).
    void set_uniform(unsigned mask, const std::vector<double>& vals)
    {
        uniform_mask = mask;
        for (int c = 0; c < nc; ++c) bcomp_h[c] = ((mask >> c) & 1u) ? vals[c] : 0.0;
        if (mask)
            hip_check(hipMemcpyAsync(bcomp_dev.p, bcomp_h.data(), sizeof(double) * kMaxComp, hipMemcpyHostToDevice,
                                     stream),
                      "bcomp");
    }


    template <int K>
    void launch_gsrb_k(int l, T* x, const T* f, unsigned act, int sweeps, bool zero_first)
    {
        auto v = view(l);
        int nh[3] = {(v.n[0] + 1) / 2, v.n[1], v.n[2] * nc};
        Launch3 L3 = cell_launch(nh);
        for (int s = 0; s < sweeps; ++s)
            for (int color = 0; color < 2; ++color)
            {
                KScope ks(kt, kc(0, l));
                mg_gsrb_kernel<DIM, K, T><<<L3.grid, L3.block, 0, stream>>>(
                    v, bcdev.p + l, per_mask, x, f, color, act, (zero_first && s == 0 && color == 0) ? 1 : 0);
            }
    }
    void gsrb(int l, T* x, const T* f, unsigned act, int sweeps, bool zero_first = false)
    {
        dispatch(l, act, [&](auto K, unsigned m) { launch_gsrb_k<decltype(K)::value>(l, x, f, m, sweeps, zero_first); });
    }

    // first level handled by the LDS-resident coarse cycle: levels >= 1 with
    // at most kFuseCells cells whose x and rhs (down to the coarsest level)
    // fit the workgroup's LDS; lv.size() when none
    static constexpr long long kFuseCells = 1024;
    static long long lds_limit_bytes()
    {
        static long long lim = -1;
        if (lim < 0)
        {
            int dev = 0, v = 0;
            hip_check(hipGetDevice(&dev), "hipGetDevice");
            hip_check(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                      "max shared memory per block");
            lim = (long long)v - 4096;   // static LDS of the kernel (reduction words, side models)
        }
        return lim;
    }
    long long coarse_lds_bytes(int l0) const
    {
        long long N[kMaxLev];
        const int nl = (int)lv.size() - l0;
        for (int q = 0; q < nl && q < kMaxLev; ++q) N[q] = lv[l0 + q].N;
        return coarse_lds_words(N, nl < kMaxLev ? nl : kMaxLev) * (long long)sizeof(T);
    }
    int fuse_level() const
    {
        for (int l = 1; l < (int)lv.size(); ++l)
            if (lv[l].N <= kFuseCells && (int)lv.size() - l <= kMaxLev && coarse_lds_bytes(l) <= lds_limit_bytes())
                return l;
        return (int)lv.size();
    }
    // level table of the coarse cycle: rebuilt on the host each call and
    // uploaded only when it differs from the copy on the device (it changes
    // with the operator's coefficients, e.g. a new dt, not per V-cycle)
    DevBuf<CoarseArgs<T>> cargs;
    CoarseArgs<T> cargs_last;
    bool cargs_valid = false;

    template <int K>
    void coarse_cycle_k(int l0, unsigned act)
    {
        upload_coarse_table(l0);
        const long long lds = coarse_lds_bytes(l0);
        if (lds > lds_limit_bytes()) throw std::runtime_error("coarse cycle: levels exceed the LDS budget");
        static bool attr_set = false;
        if (!attr_set)
        {
            hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(&mg_coarse_cycle_kernel<DIM, K, T>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit_bytes()),
                      "coarse cycle LDS attribute");
            attr_set = true;
        }
        KScope ks(kt, kbase + 4);
        mg_coarse_cycle_kernel<DIM, K, T><<<nc, kCoarseThreads, (size_t)lds, stream>>>(cargs.p, act);
    }
    void upload_coarse_table(int l0)
    {
        CoarseArgs<T> A;
        std::memset(&A, 0, sizeof(A));
        A.nlev = (int)lv.size() - l0;
        A.per_mask = per_mask;
        A.nu1 = nu1;
        A.nu2 = nu2;
        A.ncomp = nc;
        for (int q = 0; q < A.nlev; ++q)
        {
            A.v[q] = view(l0 + q);
            A.x[q] = ex[l0 + q].p;
            A.f[q] = fx[l0 + q].p;
            A.bc[q] = bcdev.p + l0 + q;
        }
        if (!cargs.p) cargs.alloc(1);
        if (!cargs_valid || std::memcmp(&A, &cargs_last, sizeof(A)) != 0)
        {
            std::memcpy(&cargs_last, &A, sizeof(A));
            cargs_valid = true;
            // the source is the member copy, unchanged until the next upload,
            // which this stream orders after the kernels reading this one
            hip_check(hipMemcpyAsync(cargs.p, &cargs_last, sizeof(A), hipMemcpyHostToDevice, stream),
                      "coarse table upload");
        }
    }

    template <int K>
    void resid_restrict_k(int l, const T* x, const T* f, unsigned act)
    {
        auto v = view(l);
        const auto& c = lv[l + 1];
        int ce[3] = {c.n[0], c.n[1], c.n[2] * nc};
        Launch3 L3 = cell_launch(ce);
        KScope ks(kt, kc(2, l));
        mg_resid_restrict_kernel<DIM, K, T><<<L3.grid, L3.block, 0, stream>>>(
            v, bcdev.p + l, per_mask, x, f, c.n[0], c.n[1], c.n[2], c.N, fx[l + 1].p, act);
    }
    void resid_restrict(int l, const T* x, const T* f, unsigned act)
    {
        dispatch(l, act, [&](auto K, unsigned m) { resid_restrict_k<decltype(K)::value>(l, x, f, m); });
    }

    template <int K>
    void bottom_k(int l, unsigned act)
    {
        auto v = view(l);
        KScope ks(kt, kbase + 4);
        mg_bottom_kernel<DIM, K, T><<<1, 1024, 0, stream>>>(v, bcdev.p + l, per_mask, ex[l].p, fx[l].p, act, nc);
    }
    void bottom(int l, unsigned act)
    {
        dispatch(l, act, [&](auto K, unsigned m) { bottom_k<decltype(K)::value>(l, m); });
    }

    void zero(int l, unsigned act)
    {
        long long N = lv[l].N;
        KScope ks(kt, kbase + 4);
        mg_zero_kernel<T><<<dim3((unsigned)((N + 255) / 256), nc, 1), 256, 0, stream>>>(N, ex[l].p, act);
    }

    void prolong(int l, unsigned act)
    {
        const auto& f = lv[l];
        const auto& c = lv[l + 1];
        int ce[3] = {c.n[0], c.n[1], c.n[2] * nc};
        Launch3 L3 = cell_launch(ce);
        KScope ks(kt, kc(3, l));
        if constexpr (DIM == 3)
        {
            // ghosted coarse correction, then direct loads (the boundary
            // neighbourhood of mg_prolong_kernel is a long dependent chain)
            const int gc[3] = {c.n[0] + 2, c.n[1] + 2, c.n[2] + 2};
            if (xghost.n >= (size_t)gc[0] * gc[1] * gc[2] * nc)
            {
                int ge[3] = {gc[0], gc[1], gc[2] * nc};
                Launch3 G = cell_launch(ge);
                mg_coarse_ghost_kernel<T><<<G.grid, G.block, 0, stream>>>(view(l + 1), bcdev.p + l + 1, per_mask,
                                                                         ex[l + 1].p, xghost.p, act);
                mg_prolong_ghost_kernel<T><<<L3.grid, L3.block, 0, stream>>>(f.n[0], f.n[1], f.n[2], f.N, ex[l].p,
                                                                            view(l + 1), xghost.p, act);
                return;
            }
        }
        mg_prolong_kernel<DIM, T><<<L3.grid, L3.block, 0, stream>>>(f.n[0], f.n[1], f.n[2], f.N, ex[l].p, view(l + 1),
                                                                    bcdev.p + l + 1, per_mask, ex[l + 1].p, act);
    }

    // Direct bottom solve (vd_mg_direct.H): the V-cycle stops at the first
    // coarse level whose extents are all at most direct_max_n and solves it
    // exactly for the components whose operator separates there (uniform
    // coefficient, no second-cell ghost term); other components recurse as
    // before.  0: off.
    int direct_max_n = 128;
    DevBuf<T> dQ, dTmp;
    DevBuf<double> dLam;
    std::vector<char> d_sig;
    std::vector<T> d_qh;
    std::vector<double> d_lh;
    unsigned d_ready = 0u;       // components whose basis is on the device
    long long d_qs = 0, d_ls = 0;
    int d_qo[3] = {0, 0, 0}, d_lo[3] = {0, 0, 0};
    // the finest level itself qualifies too (single-level solves that ask
    // for it): the solve is then the direct solve and its residual check
    bool direct_finest = false;
    int direct_level() const
    {
        if (direct_max_n <= 0) return -1;
        for (int l = direct_finest ? 0 : 1; l < (int)lv.size(); ++l)
        {
            bool ok = true;
            for (int d = 0; d < DIM; ++d) ok = ok && lv[l].n[d] <= direct_max_n;
            if (ok) return l;
        }
        return -1;
    }
    bool direct_separable(int l, int c) const
    {
        if (rz) return false;   // the radial metric breaks the Kronecker form
        if (!(level_kind(l) == CK_CONST || ((uniform_mask >> c) & 1u))) return false;
        for (int d = 0; d < DIM; ++d)
            for (int s = 0; s < 2; ++s)
            {
                if ((per_mask >> d) & 1) continue;
                const SideBC b = side_model(l, c, d, s);
                if (b.kind == BK_CF || b.dep1 != 0.0) return false;
            }
        return true;
    }
    unsigned direct_mask(int l, unsigned act) const { return (l == direct_level()) ? (act & d_ready) : 0u; }
    // Eigenbases and eigenvalues of the direct level's separable components,
    // uploaded when the operator changed (before any V-cycle of a solve, so
    // never inside a captured graph).
    void direct_prepare(unsigned want)
    {
        d_ready = 0u;
        const int l = direct_level();
        if (l < 0) return;
        const auto& h = lv[l];
        d_qs = 0;
        d_ls = 1;
        for (int d = 0; d < DIM; ++d)
        {
            d_qo[d] = (int)d_qs;
            d_lo[d] = (int)(d_ls - 1);
            d_qs += (long long)h.n[d] * h.n[d];
            d_ls += h.n[d];
        }
        std::vector<char> sig;
        auto put = [&sig](const void* p, size_t n) { sig.insert(sig.end(), (const char*)p, (const char*)p + n); };
        unsigned sepm = 0u;
        for (int c = 0; c < nc; ++c)
            if (((want >> c) & 1u) && direct_separable(l, c)) sepm |= 1u << c;
        put(&l, sizeof l);
        put(&sepm, sizeof sepm);
        put(&alpha_a, sizeof alpha_a);
        put(&beta, sizeof beta);
        put(&per_mask, sizeof per_mask);
        for (int c = 0; c < nc; ++c)
        {
            if (!((sepm >> c) & 1u)) continue;
            const double b = ((uniform_mask >> c) & 1u) ? bcomp_h[c] : bconst;
            put(&b, sizeof b);
            for (int d = 0; d < DIM; ++d)
                for (int s = 0; s < 2; ++s)
                {
                    const SideBC m = side_model(l, c, d, s);
                    put(&m, sizeof m);
                }
        }
        if (!sepm) return;
        if (sig == d_sig)
        {
            d_ready = sepm;
            return;
        }
        if (dQ.n < (size_t)(d_qs * nc)) dQ.alloc((size_t)(d_qs * nc));
        if (dLam.n < (size_t)(d_ls * nc)) dLam.alloc((size_t)(d_ls * nc));
        if (dTmp.n < (size_t)(h.N * nc)) dTmp.alloc((size_t)(h.N * nc));
        // host copies kept as members: the asynchronous copies read them
        std::vector<T>& qh = d_qh;
        std::vector<double>& lh = d_lh;
        qh.assign((size_t)(d_qs * nc), T(0));
        lh.assign((size_t)(d_ls * nc), 0.0);
        for (int c = 0; c < nc; ++c)
        {
            if (!((sepm >> c) & 1u)) continue;
            const double b = ((uniform_mask >> c) & 1u) ? bcomp_h[c] : bconst;
            double top = std::fabs(alpha_a);
            for (int d = 0; d < DIM; ++d)
            {
                const bool per = (per_mask >> d) & 1;
                const SepBasis& B = sep_basis(h.n[d], per, side_model(l, c, d, 0).dep0, side_model(l, c, d, 1).dep0);
                const double sc = beta * b * (1.0 / (h.dx[d] * h.dx[d]));
                for (size_t q = 0; q < B.Q.size(); ++q) qh[(size_t)(c * d_qs + d_qo[d]) + q] = T(B.Q[q]);
                double mx = 0.0;
                for (int i = 0; i < h.n[d]; ++i)
                {
                    lh[(size_t)(c * d_ls + d_lo[d] + i)] = sc * B.mu[i];
                    mx = std::max(mx, std::fabs(sc * B.mu[i]));
                }
                top += mx;
            }
            // eigenvalue sums below this are a singular operator's null modes
            lh[(size_t)(c * d_ls + d_ls - 1)] = 1e-12 * top;
        }
        hip_check(hipMemcpyAsync(dQ.p, qh.data(), sizeof(T) * qh.size(), hipMemcpyHostToDevice, stream), "direct Q");
        hip_check(hipMemcpyAsync(dLam.p, lh.data(), sizeof(double) * lh.size(), hipMemcpyHostToDevice, stream),
                  "direct eigenvalues");
        d_sig = sig;
        d_ready = sepm;
    }
    // Levels with an extent above kSepTileN (up to direct_max_n) run the
    // products as rocBLAS GEMMs, their summation order being the library's:
    // such a level's solve agrees with the sequential sums of the tiled
    // kernels to rounding, not bit for bit.
    rocblas_handle blas = nullptr;
    bool direct_blas(int l) const
    {
        if (l < 0 || l >= (int)lv.size()) return false;
        for (int d = 0; d < DIM; ++d)
            if (lv[l].n[d] > kSepTileN) return true;
        return false;
    }
    void direct_solve_blas(int l, unsigned m, SepArgs<T>& A)
    {
        if (!blas && rocblas_create_handle(&blas) != rocblas_status_success)
            throw std::runtime_error("rocblas_create_handle failed");
        if (rocblas_set_stream(blas, stream) != rocblas_status_success)
            throw std::runtime_error("rocblas_set_stream failed");
        const auto& h = lv[l];
        KScope ks(kt, kbase + 4);
        for (int c = 0; c < nc; ++c)
        {
            if (!((m >> c) & 1u)) continue;
            const T* Qc = dQ.p + (long long)c * d_qs;
            const T* src = fx[l].p + (long long)c * h.N;
            T* bufs[2] = {dTmp.p + (long long)c * h.N, ex[l].p + (long long)c * h.N};
            int step = 0;
            for (int pass = 0; pass < 2; ++pass)
                for (int q = 0; q < DIM; ++q)
                {
                    const int d = pass == 0 ? q : DIM - 1 - q;
                    T* out = bufs[step & 1];
                    sep_gemm_axis<T>(blas, h.n, DIM, d, pass == 1, Qc + d_qo[d], src, out);
                    if (pass == 0 && q == DIM - 1)
                        mg_sep_scale_kernel<T><<<(unsigned)((h.N + 255) / 256), 256, 0, stream>>>(A, c, out);
                    src = out;
                    ++step;
                }
        }
    }

    // ex[l] = L^{-1} fx[l] for the components of m: forward products along
    // x, y (, z), the last one divided by the eigenvalue sums, then back.
    void direct_solve(int l, unsigned m)
    {
        const auto& h = lv[l];
        SepArgs<T> A{};
        for (int d = 0; d < 3; ++d) A.n[d] = h.n[d];
        A.dim = DIM;
        A.Q = dQ.p;
        A.qs = d_qs;
        A.ls = d_ls;
        for (int d = 0; d < 3; ++d)
        {
            A.qo[d] = d_qo[d];
            A.lo[d] = d_lo[d];
        }
        A.lam = dLam.p;
        A.aa = alpha_a;
        A.N = h.N;
        A.act = m;
        const dim3 grid((unsigned)((h.N + 255) / 256), (unsigned)nc, 1);
        if (direct_blas(l))
        {
            direct_solve_blas(l, m, A);
            return;
        }
        bool tiled = true;
        for (int d = 0; d < DIM; ++d) tiled = tiled && h.n[d] <= kSepTileN;
        // f -> tmp -> x -> tmp ... ending in x (2 DIM products)
        const T* src = fx[l].p;
        T* bufs[2] = {dTmp.p, ex[l].p};
        int step = 0;
        KScope ks(kt, kbase + 4);
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < DIM; ++q)
            {
                const int d = pass == 0 ? q : DIM - 1 - q;
                A.axis = d;
                A.inverse = pass;
                A.scale = (pass == 0 && q == DIM - 1) ? 1 : 0;
                A.in = src;
                A.out = bufs[step & 1];
                if (tiled)
                {
                    // (line tiles) x (output tiles) workgroups
                    const long long nk = (h.n[d] + kSepTK - 1) / kSepTK;
                    long long tiles;
                    if (d == 0) tiles = (h.N / h.n[0] + kSepTI - 1) / kSepTI;
                    else
                    {
                        const long long S = d == 1 ? (long long)h.n[0] : (long long)h.n[0] * h.n[1];
                        tiles = h.N / (S * h.n[d]) * ((S + kSepTI - 1) / kSepTI);
                    }
                    const dim3 g2((unsigned)(tiles * nk), (unsigned)nc, 1);
                    if (d == 0) mg_sep_tile_kernel<T, true><<<g2, 256, 0, stream>>>(A);
                    else mg_sep_tile_kernel<T, false><<<g2, 256, 0, stream>>>(A);
                } else
                    mg_sep_transform_kernel<T><<<grid, 256, 0, stream>>>(A);
                src = A.out;
                ++step;
            }
    }

    // V-cycle on level l solving L e_l = rhs (rhs = fx[l]; level 0 uses fx[0] = r).
    void vcycle(int l, unsigned act)
    {
        if (const unsigned dm = direct_mask(l, act))
        {
            direct_solve(l, dm);
            act &= ~dm;
            if (!act) return;
        }
        const int lf = fuse_level();
        if (l >= lf && l < (int)lv.size() - 1)
        {
            dispatch(l, act, [&](auto K, unsigned m) {
                if constexpr (decltype(K)::value == CK_CELLINV) throw std::runtime_error("coarse cycle on level 0");
                else coarse_cycle_k<decltype(K)::value>(l, m);
            });
            return;
        }
        if (l == (int)lv.size() - 1)
        {
            bottom(l, act);
            return;
        }
        // the components with constant coefficients here run the fused
        // passes, the others the per-colour kernels (components do not couple)
        const unsigned fm = fused_mask(l, act), rm = act & ~fm;
        if (fm)
        {
            fused_pass(l, fm, true);
            if (rm)
            {
                gsrb(l, ex[l].p, fx[l].p, rm, nu1, true);
                resid_restrict(l, ex[l].p, fx[l].p, rm);
            }
            if (l == 0) vcycle_coarse(1, act);
            else vcycle(l + 1, act);
            if (rm)
            {
                prolong(l, rm);
                gsrb(l, ex[l].p, fx[l].p, rm, nu2);
            }
            fused_pass(l, fm, false, rm);
            return;
        }
        if (nu1 > 0) gsrb(l, ex[l].p, fx[l].p, act, nu1, true);
        else zero(l, act);
        resid_restrict(l, ex[l].p, fx[l].p, act);
        if (l == 0) vcycle_coarse(1, act);
        else vcycle(l + 1, act);
        prolong(l, act);
        gsrb(l, ex[l].p, fx[l].p, act, nu2);
    }

    // Graph replay of the coarse part of the V-cycle (levels >= 1) of a
    // one-component solver.  Its launches and their arguments are fixed by
    // the solver's set-up, so a sequence seen twice with the same signature
    // (every coarse level's operator view and buffers) is captured once and
    // replayed: about 40 small launches per cycle cost the host more time
    // than the GPU takes to run them.  The coarse cycle's level table lives
    // in device memory and is re-uploaded, if it changed, before a replay.
    // Timed runs of the coarse class, several components, or a capture that
    // fails take the eager path, and so do solvers whose finest level has
    // more than graph_max_cells cells: there the coarse launches are bound
    // by the GPU, not the host, and a replay measured slightly slower
    // (air3d 256^3: 22.3 against 22.1 ms per step), while the 2D 512^2
    // argonrf2d forward ran 16% faster with it.
    bool coarse_graphs = true;
    long long graph_max_cells = 1LL << 22;
    struct CoarseGraph
    {
        std::vector<char> sig;
        hipGraphExec_t exec = nullptr;
    };
    CoarseGraph cg[2];
    std::vector<char> cg_pending;
    int cg_next = 0;
    std::vector<char> coarse_signature(int l, unsigned act) const
    {
        std::vector<char> s;
        auto put = [&s](const void* p, size_t n) { s.insert(s.end(), (const char*)p, (const char*)p + n); };
        put(&l, sizeof l);
        put(&act, sizeof act);
        put(&nu1, sizeof nu1);
        put(&nu2, sizeof nu2);
        put(&per_mask, sizeof per_mask);
        put(&uniform_mask, sizeof uniform_mask);
        put(&fused_on, sizeof fused_on);
        put(&fused_min_cells, sizeof fused_min_cells);
        const void* gp[5] = {xghost.p, cargs.p, dQ.p, dLam.p, dTmp.p};
        put(gp, sizeof gp);
        put(&d_ready, sizeof d_ready);
        const void* bp = bcdev.p;
        put(&bp, sizeof bp);
        for (int q = l; q < (int)lv.size(); ++q)
        {
            const MGView<T> v = view(q);
            put(v.n, sizeof v.n);
            put(&v.Nl, sizeof v.Nl);
            put(v.h2inv, sizeof v.h2inv);
            put(&v.alpha_a, sizeof v.alpha_a);
            put(&v.beta, sizeof v.beta);
            put(&v.bconst, sizeof v.bconst);
            const void* ptr[6] = {v.bcomp, v.iD, v.bf[0], v.bf[1], v.bf[2], nullptr};
            put(ptr, sizeof ptr);
            put(v.nf, sizeof v.nf);
            const int kk = level_kind(q);
            put(&kk, sizeof kk);
            const void* buf[3] = {ex[q].p, fx[q].p, q < (int)ex_alt.size() ? ex_alt[q].p : nullptr};
            put(buf, sizeof buf);
        }
        return s;
    }
    void vcycle_coarse(int l, unsigned act)
    {
        const bool timed = kt && kt->on && ((kt->mask >> (kbase + 4)) & 1ULL);
        bool eager = !coarse_graphs || nc != 1 || timed || l >= (int)lv.size() || lv[0].N > graph_max_cells;
        // a fused pass swaps the host's buffer pointers after its launch,
        // which a replay would not repeat
        for (int q = l; !eager && q < (int)lv.size() - 1; ++q)
            if (fused_mask(q, act)) eager = true;
        // library GEMMs stay out of captured graphs
        if (d_ready && direct_blas(direct_level())) eager = true;
        if (eager)
        {
            vcycle(l, act);
            return;
        }
        const int lf = fuse_level();
        const bool table = (lf < (int)lv.size() - 1);
        const std::vector<char> sig = coarse_signature(l, act);
        for (auto& g : cg)
            if (g.exec && g.sig == sig)
            {
                if (table) upload_coarse_table(lf < l ? l : lf);
                hip_check(hipGraphLaunch(g.exec, stream), "coarse graph launch");
                return;
            }
        if (sig != cg_pending)
        {
            // first sight: run eagerly (allocates and uploads what the
            // sequence needs); capture if the next cycle looks the same
            cg_pending = sig;
            vcycle(l, act);
            return;
        }
        CoarseGraph& g = cg[cg_next];
        cg_next ^= 1;
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.sig.clear();
        hipGraph_t graph = nullptr;
        bool ok = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok)
        {
            try
            {
                vcycle(l, act);
            } catch (...)
            {
                ok = false;
            }
            if (hipStreamEndCapture(stream, &graph) != hipSuccess) ok = false;
        }
        if (ok && graph && hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess)
        {
            g.exec = nullptr;
            ok = false;
        }
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        if (!ok || !g.exec)
        {
            // nothing of the capture ran: do the cycle eagerly from now on
            coarse_graphs = false;
            vcycle(l, act);
            return;
        }
        g.sig = sig;
        hip_check(hipGraphLaunch(g.exec, stream), "coarse graph launch");
    }
    void drop_coarse_graphs()
    {
        for (auto& g : cg)
        {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            g.sig.clear();
        }
        cg_pending.clear();
    }

    // The components of act that level l runs through the fused passes:
    // those with constant coefficients there (all of them, or the uniform
    // ones), if their cells are many enough.  fine: the level of a refined
    // solver (one level, no coarse level of its own), whose pre pass hands
    // out the residual and whose post pass has no prolongation.
    unsigned fused_mask(int l, unsigned act, bool fine = false) const
    {
        if constexpr (DIM != 3) return 0u;
        if (!fused_on || nu1 != 2 || nu2 != 2) return 0u;
        if (l >= (int)ex_alt.size() || ex_alt[l].n == 0) return 0u;
        if (!fine && l + 1 >= (int)lv.size()) return 0u;
        const unsigned m = level_kind(l) == CK_CONST ? act : (act & uniform_mask);
        if (!m) return 0u;
        // below ~4M cells (over the fused components) the per-colour
        // kernels measured faster: too few workgroups, and a chunk's 8
        // pipeline-fill steps are a large share of a short march
        if (lv[l].N * (long long)__builtin_popcount(m) < fused_min_cells) return 0u;
        for (int d = 0; d < 3; ++d)
        {
            if (lv[l].n[d] % 2 != 0 || lv[l].n[d] < 8) return 0u;
            if (!fine && lv[l + 1].n[d] * 2 != lv[l].n[d]) return 0u;
        }
        return m;
    }
    // z chunking of a fused pass: a workgroup (one per CU: its LDS) marches
    // zc planes plus 8 pipeline-fill steps, so a launch takes about
    // ceil(workgroups / CUs) * (zc + 8) plane steps; the chunk count with the
    // fewest (then the fewest chunks) is taken, chunks of at least 16 planes.
    static void fused_chunks(int n2, long long tiles, int& zc, int& nchunk)
    {
        static int cus = 0;
        if (!cus)
        {
            int dev = 0;
            hip_check(hipGetDevice(&dev), "hipGetDevice");
            hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
            if (cus < 1) cus = 1;
        }
        long long best = -1;
        for (int want = 1; want <= std::max(1, n2 / 16); ++want)
        {
            int z = (n2 + want - 1) / want;
            z += z & 1;
            const int nch = (n2 + z - 1) / z;
            const long long rounds = (tiles * nch + cus - 1) / cus;
            const long long cost = rounds * (z + 8);
            if (best < 0 || cost < best)
            {
                best = cost;
                zc = z;
                nchunk = nch;
            }
        }
    }

    // A refined solver's buffer for its post pass.
    void enable_fine_fused()
    {
        if (DIM == 3 && !ex_alt.empty() && ex_alt[0].n == 0) ex_alt[0].alloc(lv[0].N * nc);
    }

    // pre: x = 2 red-black sweeps from zero on fx[l] into ex[l], then the
    // residual restricted into fx[l+1] or (fine) the residual itself into
    // rout.  post: ex[l] + P ex[l+1] (fine: ex[l]; from_zero: 0), 2 sweeps,
    // into ex_alt[l], which then becomes ex[l]; the components in carry
    // (smoothed by the per-colour kernels) are copied across first.
    void fused_pass(int l, unsigned act, bool pre, unsigned carry = 0, bool fine = false, T* rout = nullptr,
                    bool from_zero = false)
    {
        const auto& h = lv[l];
        const int ntx = (h.n[0] + kFTX - 1) / kFTX, nty = (h.n[1] + kFTY - 1) / kFTY;
        const long long tiles = (long long)ntx * nty * __builtin_popcount(act);
        int zc = 0, nchunk = 0;
        fused_chunks(h.n[2], tiles, zc, nchunk);
        FusedArgs<T> A{};
        A.v = view(l);
        A.bcs = bcdev.p + l;
        A.pm = per_mask;
        A.f = fx[l].p;
        A.zc = zc;
        A.nchunk = nchunk;
        A.active = act;
        int nact = 0;
        for (int c = 0; c < nc; ++c)
            if ((act >> c) & 1u) A.comp[nact++] = c;
        if (!nact) throw std::runtime_error("fused pass: no active component");
        dim3 grid(ntx, nty, nchunk * nact);
        if (pre)
        {
            A.xout = ex[l].p;
            if (fine)
            {
                if (!rout) throw std::runtime_error("fused pass: no residual buffer");
                A.rout = rout;
            } else
            {
                A.fc = fx[l + 1].p;
                A.Nc = lv[l + 1].N;
                A.cn0 = lv[l + 1].n[0];
                A.cn1 = lv[l + 1].n[1];
            }
            KScope ks(kt, l == 0 ? kfused : kbase + 4);
            mg_fused_pre_kernel<T><<<grid, kFThreads, 0, stream>>>(A);
            return;
        }
        if (ex_alt[l].n != ex[l].n) throw std::runtime_error("fused pass: buffers not sized for this level");
        A.xin = from_zero ? nullptr : ex[l].p;
        A.xout = ex_alt[l].p;
        const bool add = (l == 0 && add_sol != nullptr);
        if (add)
        {
            A.sol = add_sol;
            A.sstride = add_stride;
            A.compmap = compmap.p;
        }
        {
            KScope ks(kt, l == 0 ? kfused + 1 : kbase + 4);
            if (!fine)
            {
                const auto& hc = lv[l + 1];
                const int gc[3] = {hc.n[0] + 2, hc.n[1] + 2, hc.n[2] + 2};
                if (xghost.n < (size_t)gc[0] * gc[1] * gc[2] * nc)
                    throw std::runtime_error("fused pass: buffers not sized for this level");
                A.vc = view(l + 1);
                A.xcg = xghost.p;
                int ge[3] = {gc[0], gc[1], gc[2] * nc};
                Launch3 G = cell_launch(ge);
                mg_coarse_ghost_kernel<T><<<G.grid, G.block, 0, stream>>>(A.vc, bcdev.p + l + 1, per_mask, ex[l + 1].p,
                                                                         xghost.p, act);
            }
            mg_fused_post_kernel<T><<<grid, kFThreads, 0, stream>>>(A);
        }
        if (add)
        {
            add_done |= act;   // ex[0] is not needed after the update
            return;
        }
        for (int c = 0; c < nc; ++c)
            if ((carry >> c) & 1u)
                hip_check(hipMemcpyAsync(ex_alt[l].p + (size_t)c * h.N, ex[l].p + (size_t)c * h.N, sizeof(T) * h.N,
                                         hipMemcpyDeviceToDevice, stream),
                          "fused pass carry");
        std::swap(ex[l].p, ex_alt[l].p);
        std::swap(ex[l].n, ex_alt[l].n);
    }

    // One V-cycle on the residual and the solution update x += e.
    void vcycle_update(double* sol, long long stride, unsigned act)
    {
        // a fused post pass on level 0 adds its result to the solution itself
        add_sol = sol;
        add_stride = stride;
        add_done = 0u;
        vcycle(0, act);
        add_sol = nullptr;
        if (act & ~add_done) add_to_state(sol, stride, act & ~add_done);
    }
    // One smoothing iteration on the residual and the solution update
    // x += e: e = nu2 red-black sweeps from zero on the finest level (the
    // fused post pass without prolongation where it applies), no coarse
    // correction.  For the components of smooth_mask().
    void smooth_update(double* sol, long long stride, unsigned act)
    {
        add_sol = sol;
        add_stride = stride;
        add_done = 0u;
        const unsigned fm = fused_mask(0, act, true);
        if (fm) fused_pass(0, fm, false, 0, true, nullptr, true);
        add_sol = nullptr;
        const unsigned rm = act & ~add_done;
        if (rm)
        {
            gsrb(0, ex[0].p, fx[0].p, rm, nu2, true);
            add_to_state(sol, stride, rm);
        }
    }
    // Smoothing iterations instead of V-cycles.  For an operator
    // alpha_a x - beta div(b grad x) with every face b <= bmax_c, the Jacobi
    // iteration matrix has infinity norm at most
    //     q_c = 2 beta bmax_c sum_d 1/dx_d^2 / alpha_a
    // (a boundary ghost only lowers a row's off-diagonal sum), and so does
    // Gauss-Seidel's: a red-black sweep contracts the error by at least
    // 1/q_c.  Where q_c <= kSmoothQ, a V-cycle's coarse levels add nothing a
    // sweep does not, and component c iterates smooth_update instead (the
    // same stopping test).  Empty: V-cycles for every component.
    static constexpr double kSmoothQ = 1e-2;
    std::vector<double> smooth_bmax;
    unsigned smooth_mask(unsigned act) const
    {
        if (smooth_bmax.empty() || !(alpha_a > 0.0)) return 0u;
        double hs = 0.0;
        for (int d = 0; d < DIM; ++d) hs += 1.0 / (lv[0].dx[d] * lv[0].dx[d]);
        unsigned m = 0u;
        for (int c = 0; c < nc && c < (int)smooth_bmax.size(); ++c)
        {
            const double q = 2.0 * beta * smooth_bmax[c] * hs / alpha_a;
            if (((act >> c) & 1u) && q <= kSmoothQ) m |= 1u << c;
        }
        return m;
    }
    double* add_sol = nullptr;   // solution the level-0 post pass adds into
    long long add_stride = 0;
    unsigned add_done = 0u;      // components the post pass added

    // Coarse face coefficients from the finer level (after the caller
    // filled iD0 or level-1.. sources).
    void coarsen_coeffs();

    // Read per-component norms (synchronises the stream).
    void read_norms(std::vector<double>& out)
    {
        std::vector<unsigned long long> h(kMaxComp);
        d2h_sync(h.data(), nrm.p, sizeof(unsigned long long) * kMaxComp, stream, "norms d2h");
        out.resize(nc);
        for (int c = 0; c < nc; ++c) out[c] = bits_to_double(h[c]);
    }
    void clear_norms()
    {
        zero_u64_kernel<<<1, 64, 0, stream>>>(nrm.p, kMaxComp);
    }

    // Solve L sol = f0 with per-component stopping tests.  sol: state
    // pointer; component c lives at sol + compmap[c]*sol_stride (the solution
    // is updated in place).  bnorm: |f|_inf of the unfolded rhs per comp.
    // Returns the iteration count (max over comps), -1 on non-convergence.
    int solve(double* sol, long long sol_stride, const std::vector<double>& bnorm, double reltol, double abstol,
              std::vector<int>* iters_out, std::vector<double>* rnorm_out, std::vector<double>* r0_out,
              unsigned mask = 0xffffffffu)
    {
        const long long N0 = lv[0].N;
        unsigned act = ((nc >= 32) ? 0xffffffffu : ((1u << nc) - 1u)) & mask;
        // residual needs the solution in solver layout: use the gather view
        // through compmap by solving on a contiguous copy when the state is
        // strided differently; here components are gathered by pointer math.
        // r0 = f0 - L sol
        std::vector<double> rn, target(nc), r0;
        require_finite(bnorm.data(), nc, "right-hand side norm", act);
        residual_state(sol, sol_stride, act, true);
        read_norms(r0);
        require_finite(r0.data(), nc, "residual norm", act);
        if (r0_out) *r0_out = r0;
        unsigned active = 0;
        for (int c = 0; c < nc; ++c)
        {
            target[c] = std::max(abstol, reltol * std::max(bnorm[c], r0[c]));
            if (((act >> c) & 1u) && !(r0[c] <= target[c])) active |= (1u << c);
        }
        if (iters_out) iters_out->assign(nc, 0);
        int used = 0;
        rn = r0;
        const unsigned sm_all = smooth_mask(active);
        direct_prepare(active & ~sm_all);
        for (int it = 1; it <= maxiter && active; ++it)
        {
            const unsigned sm = active & sm_all, vm = active & ~sm_all;
            if (vm) vcycle_update(sol, sol_stride, vm);
            if (sm) smooth_update(sol, sol_stride, sm);
            // after smoothing iterations only (contraction >= 1/kSmoothQ per
            // sweep) the test almost always ends the solve: the residual's
            // norms alone, the residual itself only for the components that
            // go on (the same values: the same kernel, which just does not
            // store)
            const bool norm_only = (vm == 0);
            residual_state(sol, sol_stride, active, true, !norm_only);
            read_norms(rn);
            require_finite(rn.data(), nc, "residual norm", active);
            for (int c = 0; c < nc; ++c)
                if ((active >> c) & 1u)
                {
                    if (iters_out) (*iters_out)[c] = it;
                    used = it;
                    if (rn[c] <= target[c]) active &= ~(1u << c);
                }
            if (norm_only && active && it < maxiter) residual_state(sol, sol_stride, active, false, true);
        }
        if (rnorm_out) *rnorm_out = rn;
        return active ? -1 : used;
    }

    // residual of the state-resident solution into fx[0] (store) or its
    // norms only (!store)
    void residual_state(const double* sol, long long stride, unsigned act, bool norm, bool store = true);
    // r = f - L x on the finest level for component-major x, f, r (N per comp)
    void residual_plain(const double* x, const double* f, double* r, unsigned act);
    void add_to_state(double* sol, long long stride, unsigned act)
    {
        long long N = lv[0].N;
        KScope ks(kt, K_OTHER);
        mg_add_kernel<T, double><<<dim3((unsigned)((N + 255) / 256), nc, 1), 256, 0, stream>>>(N, sol, ex[0].p, act, stride,
                                                                                          compmap.p);
    }
};

// Gather kernel: residual of a solution stored at sol + compmap[c]*stride,
// one workgroup maximum of |r| per (component, workgroup) into partials.
// A thread takes kz consecutive planes of its (i, j) column (the workgroup's
// reduction and index set-up paid once per kz cells).
template <int DIM, int KIND, class T>
__global__ void __launch_bounds__(256) mg_residual_state_kernel(MGView<double> v, const LevelBC* __restrict__ bcs,
                                                                int per_mask, const double* __restrict__ sol,
                                                                long long stride, const int* __restrict__ compmap,
                                                                const double* __restrict__ f, T* __restrict__ r,
                                                                unsigned active, double* __restrict__ partials, int kz,
                                                                int nkb)
{
    const dim3 bI_ = xcd_block();
    __shared__ double scratch[16];
    const int c = bI_.z / nkb;
    const int kb = bI_.z - c * nkb;
    if (!((active >> c) & 1u)) return;   // uniform over the workgroup
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const bool in = (i < v.n[0] && j < v.n[1]);
    const long long nb = (long long)gridDim.x * gridDim.y * nkb;
    const long long bid = (long long)bI_.x + (long long)gridDim.x * ((long long)bI_.y + (long long)gridDim.y * kb);
    double m = 0.0;
    if (in)
    {
        // view the state component as a one-component array at offset 0
        const double* xs = sol + (long long)compmap[c] * stride - (long long)c * v.Nl;
        const int k0 = kb * kz, k1 = min(k0 + kz, v.n[2]);
        for (int k = k0; k < k1; ++k)
        {
            const long long ic = (long long)i + (long long)v.n[0] * ((long long)j + (long long)v.n[1] * k);
            const long long o = (long long)c * v.Nl + ic;
            const double fo = f[o];
            double Lx, dg;
            stencil<DIM, KIND, double>(v, bcs->s[c], per_mask, xs, c, i, j, k, Lx, dg);
            const double rr = fo - Lx;
            if (r) r[o] = T(rr);
            m = vmax(m, fabs(rr));
        }
    }
    block_max_store(m, partials + (long long)c * nb + bid, scratch);
}

// planes per thread of mg_residual_state_kernel and its workgroup grid
inline int resid_kz(int n2) { return n2 >= 64 ? 8 : 1; }

// 3D form of mg_residual_state_kernel marching up its kz planes: the z
// neighbours (and, for cell coefficients, their 1/b) stay in registers from
// one plane to the next, so a cell costs five loads of x instead of seven.
// The values handed to stencil_core are the ones stencil_b loads, so the
// residual is the same bit for bit.
template <int KIND, class T>
__global__ void __launch_bounds__(256) mg_residual_march_kernel(MGView<double> v, const LevelBC* __restrict__ bcs,
                                                                int per_mask, const double* __restrict__ sol,
                                                                long long stride, const int* __restrict__ compmap,
                                                                const double* __restrict__ f, T* __restrict__ r,
                                                                unsigned active, double* __restrict__ partials, int kz,
                                                                int nkb)
{
    static_assert(KIND == CK_CONST || KIND == CK_CELLINV, "level-0 coefficient kinds");
    const dim3 bI_ = xcd_block();
    __shared__ double scratch[16];
    const int c = bI_.z / nkb;
    const int kb = bI_.z - c * nkb;
    if (!((active >> c) & 1u)) return;   // uniform over the workgroup
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const bool in = (i < v.n[0] && j < v.n[1]);
    const long long nb = (long long)gridDim.x * gridDim.y * nkb;
    const long long bid = (long long)bI_.x + (long long)gridDim.x * ((long long)bI_.y + (long long)gridDim.y * kb);
    double m = 0.0;
    if (in)
    {
        const SideBC(*bc)[2] = bcs->s[c];
        const double* __restrict__ xs = sol + (long long)compmap[c] * stride;
        const double* __restrict__ ib = nullptr;
        if constexpr (KIND == CK_CELLINV) ib = v.iD + (long long)c * v.Nl;
        const int n0 = v.n[0], n1 = v.n[1], n2 = v.n[2];
        const long long sz = (long long)n0 * n1;
        const long long col = (long long)i + (long long)n0 * j;
        // raw in-plane neighbours as stencil_b indexes them
        const bool px = per_mask & 1, py = (per_mask >> 1) & 1, pz = (per_mask >> 2) & 1;
        const int dxm = (i > 0) ? -1 : (px ? n0 - 1 : 0), dxp = (i < n0 - 1) ? 1 : (px ? -(n0 - 1) : 0);
        const int dym = (j > 0) ? -n0 : (py ? (n1 - 1) * n0 : 0), dyp = (j < n1 - 1) ? n0 : (py ? -(n1 - 1) * n0 : 0);
        auto kraw = [&](int k, int s) {   // plane of the raw z neighbour on side s
            if (s == 0) return k > 0 ? k - 1 : (pz ? n2 - 1 : k);
            return k < n2 - 1 ? k + 1 : (pz ? 0 : k);
        };
        const int k0 = kb * kz, k1 = min(k0 + kz, n2);
        double xzm = xs[col + sz * kraw(k0, 0)], xzc = xs[col + sz * k0], xzp = xs[col + sz * kraw(k0, 1)];
        double dzm = 0.0, dzc = 0.0, dzp = 0.0;
        if constexpr (KIND == CK_CELLINV)
        {
            dzm = ib[col + sz * kraw(k0, 0)];
            dzc = ib[col + sz * k0];
            dzp = ib[col + sz * kraw(k0, 1)];
        }
        const double bcst = v.bcomp ? v.bcomp[c] : v.bconst;
        // the in-plane neighbours and rhs of plane k are loaded one plane
        // ahead, so a thread's loads of the next plane are in flight while
        // it evaluates the current one
        const double* __restrict__ fc = f + (long long)c * v.Nl;
        struct Plane
        {
            double fo, xm0, xm1, xp0, xp1, im0, im1, ip0, ip1;
        };
        auto load = [&](int k) __attribute__((always_inline)) {
            Plane P;
            const long long ic = col + sz * k;
            P.fo = fc[ic];
            P.xm0 = xs[ic + dxm];
            P.xm1 = xs[ic + dym];
            P.xp0 = xs[ic + dxp];
            P.xp1 = xs[ic + dyp];
            P.im0 = P.im1 = P.ip0 = P.ip1 = 0.0;
            if constexpr (KIND == CK_CELLINV)
            {
                P.im0 = ib[ic + dxm];
                P.im1 = ib[ic + dym];
                P.ip0 = ib[ic + dxp];
                P.ip1 = ib[ic + dyp];
            }
            return P;
        };
        Plane cur = load(k0);
        for (int k = k0; k < k1; ++k)
        {
            const long long ic = col + sz * k;
            const Plane P = cur;
            double nzp = 0.0, ndzp = 0.0;
            if (k + 1 < k1)
            {
                cur = load(k + 1);
                // raw z neighbours of the next plane are planes k and k+2
                // (or their images / itself at a side)
                nzp = xs[col + sz * kraw(k + 1, 1)];
                if constexpr (KIND == CK_CELLINV) ndzp = ib[col + sz * kraw(k + 1, 1)];
            }
            double xmr[3] = {P.xm0, P.xm1, xzm};
            double xpr[3] = {P.xp0, P.xp1, xzp};
            bool lob[3] = {i == 0 && !px, j == 0 && !py, k == 0 && !pz};
            bool hib[3] = {i == n0 - 1 && !px, j == n1 - 1 && !py, k == n2 - 1 && !pz};
            double bm[3], bp[3];
            if constexpr (KIND == CK_CONST)
            {
#pragma unroll
                for (int d = 0; d < 3; ++d) bm[d] = bp[d] = bcst;
            } else
            {
                const double iDm[3] = {P.im0, P.im1, dzm};
                const double iDp[3] = {P.ip0, P.ip1, dzp};
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    bm[d] = 2.0 / (dzc + iDm[d]);
                    bp[d] = 2.0 / (dzc + iDp[d]);
                }
            }
            double Lx, dg;
            stencil_core<3, double>(v, bc, xzc, xmr, xpr, lob, hib, bm, bp, Lx, dg);
            const double rr = P.fo - Lx;
            if (r) r[(long long)c * v.Nl + ic] = T(rr);
            m = vmax(m, fabs(rr));
            if (k + 1 < k1)
            {
                xzm = xzc;
                xzc = xzp;
                xzp = nzp;
                if constexpr (KIND == CK_CELLINV)
                {
                    dzm = dzc;
                    dzc = dzp;
                    dzp = ndzp;
                }
            }
        }
    }
    block_max_store(m, partials + (long long)c * nb + bid, scratch);
}

template <int DIM, class T>
void MGSolver<DIM, T>::residual_state(const double* sol, long long stride, unsigned act, bool norm, bool store)
{
    T* const r = store ? fx[0].p : nullptr;
    auto v = view0d();
    const int kz = resid_kz(v.n[2]), nkb = (v.n[2] + kz - 1) / kz;
    int ce[3] = {v.n[0], v.n[1], nkb * nc};
    Launch3 L3 = cell_launch(ce);
    const long long nb = (long long)L3.grid.x * L3.grid.y * nkb;
    dispatch(0, act, [&](auto K, unsigned m) {
        constexpr int KK = decltype(K)::value;
        if constexpr (KK == CK_FACE) throw std::runtime_error("face coefficients on level 0");
        else
        {
            KScope ks(kt, kbase + 1);
            if constexpr (DIM == 3)
                mg_residual_march_kernel<KK, T><<<L3.grid, L3.block, 0, stream>>>(
                    v, bcdev.p, per_mask, sol, stride, compmap.p, f0.p, r, m, partials.p, kz, nkb);
            else
                mg_residual_state_kernel<DIM, KK, T><<<L3.grid, L3.block, 0, stream>>>(
                    v, bcdev.p, per_mask, sol, stride, compmap.p, f0.p, r, m, partials.p, kz, nkb);
        }
    });
    if (norm) finalize_max_kernel<<<nc, 1024, 0, stream>>>(partials.p, nb, nrm.p);
}

template <int DIM, class T>
void MGSolver<DIM, T>::residual_plain(const double* x, const double* f, double* r, unsigned act)
{
    auto v = view0d();
    const int kz = resid_kz(v.n[2]), nkb = (v.n[2] + kz - 1) / kz;
    int ce[3] = {v.n[0], v.n[1], nkb * nc};
    Launch3 L3 = cell_launch(ce);
    dispatch(0, act, [&](auto K, unsigned m) {
        constexpr int KK = decltype(K)::value;
        if constexpr (KK == CK_FACE) throw std::runtime_error("face coefficients on level 0");
        else
        {
            KScope ks(kt, kbase + 1);
            if constexpr (DIM == 3)
                mg_residual_march_kernel<KK, double><<<L3.grid, L3.block, 0, stream>>>(
                    v, bcdev.p, per_mask, x, v.Nl, idmap.p, f, r, m, partials.p, kz, nkb);
            else
                mg_residual_state_kernel<DIM, KK, double><<<L3.grid, L3.block, 0, stream>>>(
                    v, bcdev.p, per_mask, x, v.Nl, idmap.p, f, r, m, partials.p, kz, nkb);
        }
    });
}

// Level-1 faces from level-0 harmonic means of cell values; coarser levels
// average 2^(D-1) finer faces.
template <int DIM, class T>
__global__ void __launch_bounds__(256) mg_face_from_cellinv_kernel(int d, int nf0, int nf1, int nf2, long long Nf,
                                                                   const double* __restrict__ iD, int per, int nc0, int nc1,
                                                                   int nc2, long long nfc, T* __restrict__ bc, int nc)
{
    // one thread per coarse face of direction d
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int e0 = nc0 + (d == 0), e1 = nc1 + (d == 1), e2 = nc2 + (d == 2);
    if (i >= e0 || j >= e1 || k >= e2) return;
    const int nt = 1 << (DIM - 1);
    for (int c = 0; c < nc; ++c)
    {
        T s = T(0);
        for (int t2 = 0; t2 < ((DIM > 2 && d != 2) ? 2 : 1); ++t2)
            for (int t1 = 0; t1 < ((d != 1) ? 2 : 1); ++t1)
                for (int t0 = 0; t0 < ((d != 0) ? 2 : 1); ++t0)
                {
                    int F[3] = {2 * i + t0, 2 * j + t1, (DIM > 2) ? 2 * k + t2 : 0};
                    int nF[3] = {nf0, nf1, nf2};
                    int R[3] = {F[0], F[1], F[2]};
                    int Lc[3] = {F[0], F[1], F[2]};
                    Lc[d] -= 1;
                    // neighbour cells of the fine face, FillPatch-mapped
                    R[d] = ghost_map(R[d], nF[d], per);
                    Lc[d] = ghost_map(Lc[d], nF[d], per);
                    long long oR = (long long)R[0] + (long long)nf0 * ((long long)R[1] + (long long)nf1 * R[2]);
                    long long oL = (long long)Lc[0] + (long long)nf0 * ((long long)Lc[1] + (long long)nf1 * Lc[2]);
                    s += T(2.0 / (iD[(long long)c * Nf + oL] + iD[(long long)c * Nf + oR]));
                }
        bc[(long long)c * nfc + (long long)i + (long long)e0 * ((long long)j + (long long)e1 * k)] = s / T(nt);
    }
}

template <int DIM, class T>
__global__ void __launch_bounds__(256) mg_face_coarsen_kernel(int d, int nf0, int nf1, int nf2, long long nff,
                                                              const T* __restrict__ bf, int nc0, int nc1, int nc2,
                                                              long long nfc, T* __restrict__ bc, int nc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int e0 = nc0 + (d == 0), e1 = nc1 + (d == 1), e2 = nc2 + (d == 2);
    if (i >= e0 || j >= e1 || k >= e2) return;
    const int f0 = nf0 + (d == 0), f1 = nf1 + (d == 1);
    const int nt = 1 << (DIM - 1);
    for (int c = 0; c < nc; ++c)
    {
        T s = T(0);
        for (int t2 = 0; t2 < ((DIM > 2 && d != 2) ? 2 : 1); ++t2)
            for (int t1 = 0; t1 < ((d != 1) ? 2 : 1); ++t1)
                for (int t0 = 0; t0 < ((d != 0) ? 2 : 1); ++t0)
                {
                    int F[3] = {2 * i + t0, 2 * j + t1, (DIM > 2) ? 2 * k + t2 : 0};
                    s += bf[(long long)c * nff + (long long)F[0] + (long long)f0 * ((long long)F[1] + (long long)f1 * F[2])];
                }
        bc[(long long)c * nfc + (long long)i + (long long)e0 * ((long long)j + (long long)e1 * k)] = s / T(nt);
        (void)nf2;
    }
}

template <class T>
__global__ void __launch_bounds__(256) convert_kernel(long long n, const double* __restrict__ a, T* __restrict__ b)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = T(a[q]);
}

template <int DIM, class T>
void MGSolver<DIM, T>::coarsen_coeffs()
{
    if (kind == CK_CONST) return;
    // every component on the constant-coefficient kernels: the coarse face
    // arrays would not be read
    const unsigned all = (nc >= 32) ? 0xffffffffu : ((1u << nc) - 1u);
    if ((uniform_mask & all) == all) return;
    if constexpr (!std::is_same<T, double>::value)
    {
        const long long n = lv[0].N * nc;
        KScope ks(kt, kbase + 4);
        convert_kernel<T><<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(n, iD0.p, iD0t.p);
    }
    for (size_t l = 1; l < lv.size(); ++l)
    {
        const auto& f = lv[l - 1];
        const auto& c = lv[l];
        for (int d = 0; d < DIM; ++d)
        {
            int e[3] = {c.n[0] + (d == 0), c.n[1] + (d == 1), c.n[2] + (d == 2)};
            Launch3 L3 = cell_launch(e);
            KScope ks(kt, kbase + 4);
            if (l == 1)
                mg_face_from_cellinv_kernel<DIM, T><<<L3.grid, L3.block, 0, stream>>>(
                    d, f.n[0], f.n[1], f.n[2], f.N, iD0.p, (per_mask >> d) & 1, c.n[0], c.n[1], c.n[2], c.nf[d],
                    bface[d][l].p, nc);
            else
                mg_face_coarsen_kernel<DIM, T><<<L3.grid, L3.block, 0, stream>>>(
                    d, f.n[0], f.n[1], f.n[2], f.nf[d], bface[d][l - 1].p, c.n[0], c.n[1], c.n[2], c.nf[d],
                    bface[d][l].p, nc);
        }
    }
}

} // namespace vd
#endif
