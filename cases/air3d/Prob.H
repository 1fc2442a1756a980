#ifndef _PROB_H_
#define _PROB_H_

#include <AMReX_Box.H>
#include <AMReX_FArrayBox.H>
#include <AMReX_Geometry.H>
#include <AMReX_ParmParse.H>
#include <Chemistry.H>
#include <ProbParm.H>
#include <UnivConstants.H>
#include <VarDefines.H>

using namespace amrex;

AMREX_INLINE
void amrex_probinit(ProbParm& h_prob_parm, ProbParm& d_prob_parm)
{
    amrex::ParmParse pp("prob");
    pp.query("peak", h_prob_parm.peak);
    pp.query("sigma", h_prob_parm.sigma);
    pp.query("xc", h_prob_parm.xc);
    pp.query("yc", h_prob_parm.yc);
    pp.query("zc", h_prob_parm.zc);
    pp.query("background", h_prob_parm.background);
    pp.query("ngas", h_prob_parm.ngas);
    pp.query("te0", h_prob_parm.te0);
    amrex::Gpu::htod_memcpy(&d_prob_parm, &h_prob_parm, sizeof(ProbParm));
}

// Quasi-neutral Gaussian seed (electrons; N2+ and O2+ in the 80/20 gas
// ratio) over a uniform pre-ionisation, in N2/O2 background gas.
AMREX_GPU_DEVICE
AMREX_INLINE
void initdomaindata(
    Box const& bx,
    Array4<Real> const& phi,
    GeometryData const& geomdata,
    ProbParm* prob_parm)
{
    const auto lo = lbound(bx);
    const auto hi = ubound(bx);
    const Real* prob_lo = geomdata.ProbLo();
    const Real* dx = geomdata.CellSize();
    for (int k = lo.z; k <= hi.z; ++k)
    for (int j = lo.y; j <= hi.y; ++j)
    for (int i = lo.x; i <= hi.x; ++i)
    {
        for (int c = 0; c < NVAR; c++) phi(i, j, k, c) = 0.0;
        const Real x = prob_lo[0] + (i + 0.5) * dx[0] - prob_parm->xc;
        const Real y = prob_lo[1] + (j + 0.5) * dx[1] - prob_parm->yc;
        const Real z = prob_lo[2] + (k + 0.5) * dx[2] - prob_parm->zc;
        const Real s = prob_parm->sigma;
        const Real ne = prob_parm->peak * std::exp(-(x * x + y * y + z * z) / (2.0 * s * s)) +
                        prob_parm->background;
        phi(i, j, k, E_ID) = ne;
        phi(i, j, k, N2p_ID) = 0.8 * ne;
        phi(i, j, k, O2p_ID) = 0.2 * ne;
        phi(i, j, k, N2_ID) = 0.8 * prob_parm->ngas;
        phi(i, j, k, O2_ID) = 0.2 * prob_parm->ngas;
        phi(i, j, k, ETEMP_ID) = prob_parm->te0;
        phi(i, j, k, EEN_ID) = 1.5 * K_B * ne * prob_parm->te0;
        phi(i, j, k, CMASK_ID) = 1.0;
    }
}
#endif
