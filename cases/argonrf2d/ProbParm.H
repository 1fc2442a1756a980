#ifndef CNS_PROB_PARM_H_
#define CNS_PROB_PARM_H_
#include <AMReX_REAL.H>
#include <AMReX_GpuMemory.H>

struct ProbParm
{
    amrex::Real peak = 1.0e16;     // m^-3
    amrex::Real background = 1.0e15; // uniform plasma density added to the seed, m^-3
    amrex::Real sigma = 1.0e-3;    // m
    amrex::Real noise = 0.01;      // relative amplitude of N(0,1) noise
    amrex::Real ngas = 2.4e25;     // background Ar density, m^-3
    amrex::Real te0 = 11604.5;     // initial electron temperature, K
    long long seed = 20261203;
};
#endif
