// User transport hooks of the air3d case.  The inputs file leaves
// user_defined_potential/species/vel at 0, so the solver's default hooks
// apply; these definitions complete the plugin interface.
#ifndef _TRANSPORT_H_
#define _TRANSPORT_H_

#include <AMReX_Geometry.H>
#include <AMReX_FArrayBox.H>
#include <AMReX_REAL.H>
#include <Chemistry.H>
#include <ProbParm.H>
#include <VarDefines.H>

using namespace amrex;
namespace user_transport {

// Current collectors: the boundary cells of the surface a location names
// (surface code 2*dir + side), e.g. the z-lo electrode for code 4.
AMREX_GPU_HOST_DEVICE AMREX_INLINE amrex::Real current_collector_value(
    IntVect cellid,
    int locid,
    int surfaceloc,
    GpuArray<Real, AMREX_SPACEDIM> problo,
    GpuArray<Real, AMREX_SPACEDIM> probhi,
    GpuArray<int, AMREX_SPACEDIM> domlo,
    GpuArray<int, AMREX_SPACEDIM> domhi,
    GpuArray<Real, AMREX_SPACEDIM> dx)
{
    const int dir = surfaceloc / 2;
    const int cell = cellid[dir];
    const bool on = (surfaceloc % 2 == 0) ? (cell == domlo[dir]) : (cell == domhi[dir]);
    return on ? 1.0 : 0.0;
}

AMREX_GPU_DEVICE AMREX_INLINE void potential_bc(
    int i, int j, int k, int dir, int sgn,
    Array4<Real> const& phi, Array4<Real> const& bcarr,
    Array4<Real> const& robin_a, Array4<Real> const& robin_b, Array4<Real> const& robin_f,
    GpuArray<Real, AMREX_SPACEDIM> prob_lo, GpuArray<Real, AMREX_SPACEDIM> prob_hi,
    GpuArray<Real, AMREX_SPACEDIM> dx, const Real time, ProbParm const& prob_parm,
    Real Tg, Real Pg, Real app_voltage,
    GpuArray<Real, MAX_CURRENT_LOCS> int_currents,
    GpuArray<Real, MAX_CURRENT_LOCS> int_current_areas,
    GpuArray<int, MAX_CURRENT_LOCS> int_current_surfaces)
{
    IntVect ghost{AMREX_D_DECL(i, j, k)};
    if (sgn == -1) ghost[dir] -= 1;
    robin_a(ghost) = 1.0;
    robin_b(ghost) = 0.0;
    robin_f(ghost) = app_voltage;
    bcarr(ghost) = app_voltage;
}

AMREX_GPU_DEVICE AMREX_INLINE void species_bc(
    int i, int j, int k, int dir, int sgn, int spec_id, int bcspec_id,
    Array4<Real> const& phi, Array4<Real> const& bcarr,
    Array4<Real> const& robin_a, Array4<Real> const& robin_b, Array4<Real> const& robin_f,
    GpuArray<Real, AMREX_SPACEDIM> prob_lo, GpuArray<Real, AMREX_SPACEDIM> prob_hi,
    GpuArray<Real, AMREX_SPACEDIM> dx, const Real time, ProbParm const& prob_parm,
    Real Tg, Real Pg,
    GpuArray<Real, MAX_CURRENT_LOCS> int_currents,
    GpuArray<Real, MAX_CURRENT_LOCS> int_current_areas,
    GpuArray<int, MAX_CURRENT_LOCS> int_current_surfaces)
{
    IntVect ghost{AMREX_D_DECL(i, j, k)};
    if (sgn == -1) ghost[dir] -= 1;
    robin_a(ghost, bcspec_id) = 0.0;
    robin_b(ghost, bcspec_id) = 1.0;
    robin_f(ghost, bcspec_id) = 0.0;
    bcarr(ghost, bcspec_id) = 0.0;
}

AMREX_GPU_DEVICE AMREX_INLINE amrex::Real compute_vel(
    IntVect cellid, int dir, int specid, Array4<Real> const& phi,
    GpuArray<Real, AMREX_SPACEDIM> dx, const Real time, ProbParm const& prob_parm,
    Real Tg, Real Pg)
{
    amrex::Real emag = 0.0;
    for (int d = 0; d < AMREX_SPACEDIM; ++d)
        emag += amrex::Math::powi<2>(phi(cellid, EFX_ID + d));
    emag = std::sqrt(emag);
    amrex::Real ndens = 0.0;
    for (int sp = 0; sp < NUM_SPECIES; sp++) ndens += phi(cellid, sp);
    return specMob(specid, phi(cellid, ETEMP_ID), ndens, emag, Tg) *
           phi(cellid, EFX_ID + dir);
}

// Photoionization hooks (Source/PhotoIonizationSolve.cpp): this case does not
// run photoionization; coefficient and rhs stay at the zero the solver sets,
// and the sides are homogeneous Neumann in Robin form.
AMREX_GPU_DEVICE AMREX_INLINE void get_photoion_acoeff(
    int, int, int, int, amrex::Array4<amrex::Real> const&, amrex::Array4<amrex::Real> const&,
    amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>, amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>,
    amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>, const amrex::Real, ProbParm const&, amrex::Real, amrex::Real)
{
}

AMREX_GPU_DEVICE AMREX_INLINE void get_photoion_rhs(
    int, int, int, int, amrex::Array4<amrex::Real> const&, amrex::Array4<amrex::Real> const&,
    amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>, amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>,
    amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>, const amrex::Real, ProbParm const&, amrex::Real, amrex::Real)
{
}

AMREX_GPU_DEVICE AMREX_INLINE void photoionization_bc(
    int i, int j, int k, int dir, int sgn, amrex::Array4<amrex::Real> const&, amrex::Array4<amrex::Real> const& bc_arr,
    amrex::Array4<amrex::Real> const& robin_a, amrex::Array4<amrex::Real> const& robin_b,
    amrex::Array4<amrex::Real> const& robin_f, amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>,
    amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>, amrex::GpuArray<amrex::Real, AMREX_SPACEDIM>, const amrex::Real,
    ProbParm const&, amrex::Real, amrex::Real)
{
    amrex::IntVect ghost{AMREX_D_DECL(i, j, k)};
    if (sgn == -1) ghost[dir] -= 1;
    robin_a(ghost) = 0.0;
    robin_b(ghost) = 1.0;
    robin_f(ghost) = 0.0;
    bc_arr(ghost) = 0.0;
}

} // namespace user_transport
#endif
