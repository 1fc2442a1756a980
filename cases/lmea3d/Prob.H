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
    pp.query("n0", h_prob_parm.n0);
    pp.query("noise", h_prob_parm.noise);
    pp.query("ngas", h_prob_parm.ngas);
    pp.query("te0", h_prob_parm.te0);
    pp.query("nexc", h_prob_parm.nexc);
    amrex::Gpu::htod_memcpy(&d_prob_parm, &h_prob_parm, sizeof(ProbParm));
}

// Uniform quasi-neutral plasma with seeded lognormal noise: n_e = n_Ar+ =
// n0 exp(s xi - s^2/2), xi ~ N(0,1) from a counter-based generator keyed on
// (seed, global cell index of the level's domain), so every level and the
// oracle see the same field for the same cell.
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
    const int nx = geomdata.Domain().length(0);
    const int ny = geomdata.Domain().length(1);

    for (int k = lo.z; k <= hi.z; ++k)
    for (int j = lo.y; j <= hi.y; ++j)
    for (int i = lo.x; i <= hi.x; ++i)
    {
        for (int c = 0; c < NVAR; c++) phi(i, j, k, c) = 0.0;
        const unsigned long long idx =
            (unsigned long long)i + (unsigned long long)nx * ((unsigned long long)j + (unsigned long long)ny * k);
        const Real s = prob_parm->noise;
        const Real xi = vd_rng::normal(prob_parm->seed, 0, idx);
        const Real n = prob_parm->n0 * std::exp(s * xi - 0.5 * s * s);
        phi(i, j, k, E_ID) = n;
        phi(i, j, k, ARp_ID) = n;
        phi(i, j, k, AR2p_ID) = 0.0;
        phi(i, j, k, ARm_ID) = prob_parm->nexc;
        phi(i, j, k, ARr_ID) = prob_parm->nexc;
        phi(i, j, k, AR_ID) = prob_parm->ngas;
        phi(i, j, k, ETEMP_ID) = prob_parm->te0;
        phi(i, j, k, EEN_ID) = 1.5 * K_B * n * prob_parm->te0;
        phi(i, j, k, CMASK_ID) = 1.0;
    }
}
#endif
