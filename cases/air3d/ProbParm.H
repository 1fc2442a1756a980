#ifndef CNS_PROB_PARM_H_
#define CNS_PROB_PARM_H_
#include <AMReX_REAL.H>
#include <AMReX_GpuMemory.H>

struct ProbParm
{
    amrex::Real peak = 1.0e18;    // seed plasma peak density, m^-3
    amrex::Real sigma = 0.5e-3;   // m
    amrex::Real xc = 0.01, yc = 0.01, zc = 0.002;  // m
    amrex::Real background = 1.0e12;              // uniform pre-ionisation, m^-3
    amrex::Real ngas = 2.5e25;    // total gas density, m^-3 (80% N2, 20% O2)
    amrex::Real te0 = 23209.0;    // K
};
#endif
