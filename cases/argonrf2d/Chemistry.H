// Argon RF discharge chemistry (2D, 512x512), table-driven, with run-time
// reaction prefactors and electron transport coefficients.
//
// Species: electrons, Ar+, Ar* (metastable), Ar (background gas).
// Reactions (two-reactant rates in m^3/s, one-reactant in 1/s; E/N in Td),
// k_r = p_r A_r exp(-B_r/(E/N)), p_r = case_param(r) (default 1):
//   R0  E + AR    -> 2E + ARp        A 1e-16    B 200   ionization
//   R1  E + AR    -> E + ARm         A 5e-16    B 100   excitation
//   R2  E + ARp   -> AR              A 1e-13            recombination
//   R3  ARm + AR  -> 2 AR            A 2e-16            quenching
//   R4  E + ARm   -> 2E + ARp        A 2e-14    B 50    stepwise ionization
//   R5  E + ARm   -> E + AR          A 2e-13            superelastic
//   R6  ARm + ARm -> E + ARp + AR    A 6.2e-16          metastable pooling
//   R7  E + AR    -> 2E + ARp        A 3e-17    B 400   ionization (high-energy channel)
//   R8  E + ARp   -> ARm             A 5e-14            recombination to Ar*
//   R9  ARm + AR  -> E + ARp + AR    A 1e-20            Penning-type ionization
//   R10 ARm       -> AR              A 1e5              radiative decay
//   R11 E + AR    -> E + ARm         A 2e-16    B 300   excitation (second channel)
// R0-R3 are the argon2d mechanism.  Transport: mu_e = case_param(12)
// (default -0.03, charge-signed), D_e = case_param(13) (default 0.1),
// mu_i = 1e-4, D_i = 1e-5 for Ar+ and Ar*.
// The plugin interface (species macros, specMob, specDiff, CKWC, CKCHRG,
// etempLFA, mw/imw) is the vidyut3d case-plugin interface.
#ifndef MECHANISM_H
#define MECHANISM_H

#include <AMReX_Gpu.H>
#include <AMReX_REAL.H>
#include <UnivConstants.H>

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
#define NUM_REACTIONS 12

// run-time parameters: 12 reaction prefactors, mu_e, D_e
#define VD_NUM_CASE_PARAMS 14
#define VD_CASE_PARAM_DEFAULTS 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -0.03, 0.1
// which case parameters are the electron mobility and diffusivity (their
// derivatives enter the adjoint through the drift and diffusion terms)
#define VD_PARAM_MU_E 12
#define VD_PARAM_D_E 13
#include <vd_case_params.H>

#define VD_NUM_RXN 12
constexpr int vd_rxn_reac[VD_NUM_RXN][2] = {{E_ID, AR_ID},  {E_ID, AR_ID},   {E_ID, ARp_ID}, {ARm_ID, AR_ID},
                                            {E_ID, ARm_ID}, {E_ID, ARm_ID},  {ARm_ID, ARm_ID}, {E_ID, AR_ID},
                                            {E_ID, ARp_ID}, {ARm_ID, AR_ID}, {ARm_ID, -1},    {E_ID, AR_ID}};
//                                        E  ARp ARm AR
constexpr int vd_rxn_nu[VD_NUM_RXN][4] = {{1, 1, 0, -1},  {0, 0, 1, -1}, {-1, -1, 0, 1}, {0, 0, -1, 1},
                                          {1, 1, -1, 0},  {0, 0, -1, 1}, {1, 1, -2, 1},  {1, 1, 0, -1},
                                          {-1, -1, 1, 0}, {1, 1, -1, 0}, {0, 0, -1, 1},  {0, 0, 1, -1}};
constexpr double vd_rxn_A[VD_NUM_RXN] = {1e-16, 5e-16, 1e-13, 2e-16,   2e-14, 2e-13,
                                         6.2e-16, 3e-17, 5e-14, 1e-20, 1e5,   2e-16};
constexpr double vd_rxn_B[VD_NUM_RXN] = {200.0, 100.0, 0.0, 0.0, 50.0, 0.0, 0.0, 400.0, 0.0, 0.0, 0.0, 300.0};
#define VD_RXN_PREFACTOR(r) case_param(r)
#include <vd_rxn_table.H>

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
    if (specid == E_ID) mu = case_param(12);
    if (specid == ARp_ID) mu = 1.0e-4;
    if (specid == NUM_SPECIES) mu = fivebythree * case_param(12);
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
    if (specid == E_ID) d = case_param(13);
    if (specid == ARp_ID) d = 1.0e-5;
    if (specid == ARm_ID) d = 1.0e-5;
    if (specid == NUM_SPECIES) d = fivebythree * case_param(13);
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
    vd_table_ckwc(C, wdot, EN);
    *ener_exch = 0.0;
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
