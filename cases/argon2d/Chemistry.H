// Argon LFA discharge chemistry (2D, 256x256 test problem).
//
// Species: electrons, Ar+, Ar* (metastable), Ar (background gas).
// Reactions (rate constants in m^3/s; E/N in Td):
//   R0  E + AR  -> 2E + ARp      k = 1e-16 exp(-200/(E/N))   ionization
//   R1  E + AR  -> E + ARm       k = 5e-16 exp(-100/(E/N))   excitation
//   R2  E + ARp -> AR            k = 1e-13                   recombination
//   R3  ARm + AR -> 2 AR         k = 2e-16                   quenching
// Transport: mu_e = -0.03 (charge-signed), D_e = 0.1, mu_i = 1e-4,
// D_i = 1e-5 (m^2/V/s, m^2/s); Ar* diffuses with D_i.
// The plugin interface (species macros, specMob, specDiff, CKWC, CKCHRG,
// etempLFA, mw/imw) is the vidyut3d case-plugin interface.
#ifndef MECHANISM_H
#define MECHANISM_H

#include <AMReX_Gpu.H>
#include <AMReX_REAL.H>
#include <UnivConstants.H>
#include <VarDefines.H>

#define E_ID 0
#define ARp_ID 1
#define ARm_ID 2
#define AR_ID 3

#define NUM_ELEMENTS 2
#define NUM_SPECIES 4
#define FIRST_ION 1
#define NUM_IONS 1
#define FIRST_NEUTRAL 2
#define NUM_NEUTRALS 1
#define NUM_REACTIONS 4

void CKSYMS_STR(amrex::Vector<std::string>& kname);
void CKSYME_STR(amrex::Vector<std::string>& ename);

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real mw(const int n)
{
    // g/mol
    return (n == E_ID) ? 5.485799e-4 : ((n == ARp_ID) ? 39.947451 : 39.948);
}
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real imw(const int n)
{
    return 1.0 / mw(n);
}
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void get_mw(amrex::Real* w)
{
    for (int n = 0; n < NUM_SPECIES; ++n) w[n] = mw(n);
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real specMob(
    const int specid,
    const amrex::Real Te,
    const amrex::Real ndens,
    const amrex::Real emag,
    const amrex::Real T)
{
    amrex::Real mu = 0.0;
    if (specid == E_ID) mu = -0.03;
    if (specid == ARp_ID) mu = 1.0e-4;
    if (specid == NUM_SPECIES) mu = fivebythree * (-0.03);
    return mu;
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real specDiff(
    const int specid,
    const amrex::Real Te,
    const amrex::Real ndens,
    const amrex::Real emag,
    const amrex::Real T)
{
    amrex::Real d = 0.0;
    if (specid == E_ID) d = 0.1;
    if (specid == ARp_ID) d = 1.0e-5;
    if (specid == ARm_ID) d = 1.0e-5;
    if (specid == NUM_SPECIES) d = fivebythree * 0.1;
    return d;
}

// Molar production rates wdot [mol/m^3/s] from concentrations C [mol/m^3].
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKWC(
    const amrex::Real T,
    amrex::Real C[],
    amrex::Real wdot[],
    const amrex::Real Te,
    const amrex::Real EN,
    amrex::Real* ener_exch)
{
    for (int n = 0; n < NUM_SPECIES; ++n) wdot[n] = 0.0;
    *ener_exch = 0.0;
    const amrex::Real ENs = amrex::max(EN, 1.0e-30);
    // m^3/s -> m^3/mol/s
    const amrex::Real k0 = 1.0e-16 * std::exp(-200.0 / ENs) * N_A;
    const amrex::Real k1 = 5.0e-16 * std::exp(-100.0 / ENs) * N_A;
    const amrex::Real k2 = 1.0e-13 * N_A;
    const amrex::Real k3 = 2.0e-16 * N_A;

    const amrex::Real q0 = k0 * C[E_ID] * C[AR_ID];
    const amrex::Real q1 = k1 * C[E_ID] * C[AR_ID];
    const amrex::Real q2 = k2 * C[E_ID] * C[ARp_ID];
    const amrex::Real q3 = k3 * C[ARm_ID] * C[AR_ID];

    wdot[E_ID] += q0 - q2;
    wdot[ARp_ID] += q0 - q2;
    wdot[AR_ID] += -q0 - q1 + q2 + q3;
    wdot[ARm_ID] += q1 - q3;
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real
etempLFA(const amrex::Real Tg, amrex::Real C[], const amrex::Real EN)
{
    return 11604.5; // 1 eV
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKCHRG(int kcharge[])
{
    kcharge[E_ID] = -1;
    kcharge[ARp_ID] = 1;
    kcharge[ARm_ID] = 0;
    kcharge[AR_ID] = 0;
}
#endif
