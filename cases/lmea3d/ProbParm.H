#ifndef CNS_PROB_PARM_H_
#define CNS_PROB_PARM_H_
#include <AMReX_REAL.H>
#include <AMReX_GpuMemory.H>

struct ProbParm
{
    amrex::Real n0 = 1.0e15;       // uniform plasma density, m^-3
    amrex::Real noise = 0.01;      // lognormal noise, sigma of ln(n)
    amrex::Real ngas = 3.22e22;    // Ar density (1 Torr, 300 K), m^-3
    amrex::Real te0 = 11604.5;     // initial electron temperature (1 eV), K
    amrex::Real nexc = 1.0e12;     // initial Ar2+, Ar*, Ar^r densities, m^-3
    unsigned long long seed = 20261203ULL;
};
#endif
