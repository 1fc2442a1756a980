// Generated by gen_chemistry.py -- 3D argon LMEA mechanism (do not edit).
// Electron-impact k = A exp(-C/eps) [m^3/s, eV], seed 20261203; constant rates otherwise.
#ifndef MECHANISM_H
#define MECHANISM_H

#include <AMReX_Gpu.H>
#include <AMReX_REAL.H>
#include <UnivConstants.H>
#include <VarDefines.H>

#define E_ID 0
#define ARp_ID 1
#define AR2p_ID 2
#define ARm_ID 3
#define ARr_ID 4
#define AR_ID 5

#define NUM_ELEMENTS 2
#define NUM_SPECIES 6
#define FIRST_ION 1
#define NUM_IONS 2
#define FIRST_NEUTRAL 3
#define NUM_NEUTRALS 2
#define NUM_REACTIONS 11

void CKSYMS_STR(amrex::Vector<std::string>& kname);
void CKSYME_STR(amrex::Vector<std::string>& ename);

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real mw(const int n)
{
    const amrex::Real w[NUM_SPECIES] = {0.0005485799, 39.947451, 79.895451, 39.948, 39.948, 39.948};
    return w[n];
}
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real imw(const int n) { return 1.0 / mw(n); }
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void get_mw(amrex::Real* w)
{
    for (int n = 0; n < NUM_SPECIES; ++n) w[n] = mw(n);
}

// Charge-signed mobilities [m^2/V/s] and diffusivities [m^2/s]: electrons
// -0.03 / 0.1, ions 1e-4 / 1e-5, excited neutrals 0 / 1e-5; NUM_SPECIES
// selects the electron-energy coefficients (5/3 of the electron ones).
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real specMob(
    const int specid, const amrex::Real Te, const amrex::Real ndens,
    const amrex::Real emag, const amrex::Real T)
{
    if (specid == E_ID) return -0.03;
    if (specid == ARp_ID || specid == AR2p_ID) return 1.0e-4;
    if (specid == NUM_SPECIES) return fivebythree * (-0.03);
    return 0.0;
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real specDiff(
    const int specid, const amrex::Real Te, const amrex::Real ndens,
    const amrex::Real emag, const amrex::Real T)
{
#ifdef LMEA_TE_DIFFUSION
    // cases/lmea3d_te: proportional to T_e (1 eV = 11604.5 K)
    if (specid == E_ID) return 0.1 * (Te / 11604.5);
    if (specid == NUM_SPECIES) return fivebythree * 0.1 * (Te / 11604.5);
#else
    if (specid == E_ID) return 0.1;
    if (specid == NUM_SPECIES) return fivebythree * 0.1;
#endif
    return 1.0e-5;
}

// electron-impact table: A [m^3/s], C [eV], threshold [eV]
constexpr int lm_ne = 5;
constexpr double lm_A[lm_ne] = {3.261354034789929e-14, 6.3246830307730816e-15, 3.6223866197644208e-16, 1.2174924055823978e-15, 3.992606119074599e-15};
constexpr double lm_C[lm_ne] = {5.452196941409241, 5.3338585136703109, 18.883026184237714, 12.73487380611377, 15.088071212386595};
constexpr double lm_eexci[lm_ne] = {15.8, 11.55, 11.62, 4.21, 4.14};
constexpr int lm_cn = 6;
constexpr double lm_K[lm_cn] = {2.5e-19, 1e-13, 6.2e-16, 2.1e-21, 100000.0, 1e-13};

// Molar production rates wdot [mol/m^3/s] from concentrations C [mol/m^3],
// and the electron energy exchange [W/m^3] (inelastic losses and elastic
// exchange, the form of the reference's comp_ener_exch).
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKWC(
    const amrex::Real T, amrex::Real C[], amrex::Real wdot[], const amrex::Real Te,
    const amrex::Real EN, amrex::Real* ener_exch)
{
    const amrex::Real eps = 1.5 * K_B * Te / ECHARGE;   // mean energy, eV
    const amrex::Real ieps = 1.0 / (eps > 1.0e-3 ? eps : 1.0e-3);
    const amrex::Real sefact = (std::tanh((Te - T) / 0.5) + 1.0) / 2.0;
    amrex::Real q[NUM_REACTIONS];
    q[0] = lm_A[0] * N_A * std::exp(-lm_C[0] * ieps) * C[E_ID] * C[AR_ID];
    q[1] = lm_A[1] * N_A * std::exp(-lm_C[1] * ieps) * C[E_ID] * C[AR_ID];
    q[2] = lm_A[2] * N_A * std::exp(-lm_C[2] * ieps) * C[E_ID] * C[AR_ID];
    q[3] = lm_A[3] * N_A * std::exp(-lm_C[3] * ieps) * C[E_ID] * C[ARm_ID];
    q[4] = lm_A[4] * N_A * std::exp(-lm_C[4] * ieps) * C[E_ID] * C[ARr_ID];
    q[5] = lm_K[0] * N_A * C[ARp_ID] * C[AR_ID];
    q[6] = lm_K[1] * N_A * C[E_ID] * C[AR2p_ID];
    q[7] = lm_K[2] * N_A * C[ARm_ID] * C[ARm_ID];
    q[8] = lm_K[3] * N_A * C[ARm_ID] * C[AR_ID];
    q[9] = lm_K[4] * C[ARr_ID];
    q[10] = lm_K[5] * N_A * C[E_ID] * C[ARp_ID];
    for (int s = 0; s < NUM_SPECIES; ++s) wdot[s] = 0.0;
    wdot[E_ID] += 1.0 * q[0];
    wdot[AR_ID] += -1.0 * q[0];
    wdot[ARp_ID] += 1.0 * q[0];
    wdot[AR_ID] += -1.0 * q[1];
    wdot[ARm_ID] += 1.0 * q[1];
    wdot[AR_ID] += -1.0 * q[2];
    wdot[ARr_ID] += 1.0 * q[2];
    wdot[E_ID] += 1.0 * q[3];
    wdot[ARm_ID] += -1.0 * q[3];
    wdot[ARp_ID] += 1.0 * q[3];
    wdot[E_ID] += 1.0 * q[4];
    wdot[ARr_ID] += -1.0 * q[4];
    wdot[ARp_ID] += 1.0 * q[4];
    wdot[ARp_ID] += -1.0 * q[5];
    wdot[AR_ID] += -1.0 * q[5];
    wdot[AR2p_ID] += 1.0 * q[5];
    wdot[E_ID] += -1.0 * q[6];
    wdot[AR2p_ID] += -1.0 * q[6];
    wdot[ARm_ID] += 1.0 * q[6];
    wdot[AR_ID] += 1.0 * q[6];
    wdot[ARm_ID] += -2.0 * q[7];
    wdot[E_ID] += 1.0 * q[7];
    wdot[ARp_ID] += 1.0 * q[7];
    wdot[AR_ID] += 1.0 * q[7];
    wdot[ARm_ID] += -1.0 * q[8];
    wdot[AR_ID] += 1.0 * q[8];
    wdot[ARr_ID] += -1.0 * q[9];
    wdot[AR_ID] += 1.0 * q[9];
    wdot[E_ID] += -1.0 * q[10];
    wdot[ARp_ID] += -1.0 * q[10];
    wdot[AR_ID] += 1.0 * q[10];
    amrex::Real ex = 0.0;
    for (int r = 0; r < lm_ne; ++r) ex -= (lm_eexci[r] * ECHARGE) * N_A * q[r] * sefact;
    const amrex::Real qel = 1e-13 * N_A * C[E_ID] * C[AR_ID];
    ex -= 3.0 * K_B * (mw(E_ID) / mw(AR_ID)) * (Te - T) * qel * N_A;
    *ener_exch = ex;
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
    kcharge[AR2p_ID] = 1;
    kcharge[ARm_ID] = 0;
    kcharge[ARr_ID] = 0;
    kcharge[AR_ID] = 0;
}
#endif
