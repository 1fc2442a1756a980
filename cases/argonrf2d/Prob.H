// Initial state of the argonrf2d case: a broad Gaussian plasma on a uniform
// plasma background (so the plasma reaches both electrodes), independent
// per-cell noise on electrons and ions (net space charge), Ar* zero,
// uniform background Ar.
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
#include <vd_rng.H>

using namespace amrex;

AMREX_INLINE
void amrex_probinit(ProbParm& h_prob_parm, ProbParm& d_prob_parm)
{
    amrex::ParmParse pp("prob");
    pp.query("peak", h_prob_parm.peak);
    pp.query("background", h_prob_parm.background);
    pp.query("sigma", h_prob_parm.sigma);
    pp.query("noise", h_prob_parm.noise);
    pp.query("ngas", h_prob_parm.ngas);
    pp.query("te0", h_prob_parm.te0);
    pp.query("seed", h_prob_parm.seed);
    amrex::Gpu::htod_memcpy(&d_prob_parm, &h_prob_parm, sizeof(ProbParm));
}

// Gaussian seed plasma at the domain centre with per-cell N(0,1) noise drawn
// from a counter-based generator keyed on (seed, species, global cell index).
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
    const Real* prob_hi = geomdata.ProbHi();
    const Real* dx = geomdata.CellSize();
    const int nx = geomdata.Domain().length(0);
    const int ny = geomdata.Domain().length(1);

    for (int k = lo.z; k <= hi.z; ++k)
    for (int j = lo.y; j <= hi.y; ++j)
    for (int i = lo.x; i <= hi.x; ++i)
    {
        for (int c = 0; c < NVAR; c++) phi(i, j, k, c) = 0.0;
        const Real x = prob_lo[0] + (i + 0.5) * dx[0];
        const Real y = prob_lo[1] + (j + 0.5) * dx[1];
        const Real xc = 0.5 * (prob_lo[0] + prob_hi[0]);
        const Real yc = 0.5 * (prob_lo[1] + prob_hi[1]);
        const Real r2 = (x - xc) * (x - xc) + (y - yc) * (y - yc);
        const Real s = prob_parm->sigma;
        const Real g = prob_parm->peak * std::exp(-r2 / (2.0 * s * s)) + prob_parm->background;
        const unsigned long long idx =
            (unsigned long long)i + (unsigned long long)nx * ((unsigned long long)j + (unsigned long long)ny * k);
        const Real xe = vd_rng::normal(prob_parm->seed, E_ID, idx);
        const Real xi = vd_rng::normal(prob_parm->seed, ARp_ID, idx);
        phi(i, j, k, E_ID) = g * (1.0 + prob_parm->noise * xe);
        phi(i, j, k, ARp_ID) = g * (1.0 + prob_parm->noise * xi);
        phi(i, j, k, ARm_ID) = 0.0;
        phi(i, j, k, AR_ID) = prob_parm->ngas;
        phi(i, j, k, ETEMP_ID) = prob_parm->te0;
        phi(i, j, k, EEN_ID) = 1.5 * K_B * phi(i, j, k, E_ID) * prob_parm->te0;
        phi(i, j, k, CMASK_ID) = 1.0;
    }
}
#endif
