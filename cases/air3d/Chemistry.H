// Generated by gen_chemistry.py -- 3D air streamer mechanism (do not edit).
// 20 reactions k = A exp(-B/(E/N)) [m^3/s, Td], seed 20261203.
#ifndef MECHANISM_H
#define MECHANISM_H

#include <AMReX_Gpu.H>
#include <AMReX_REAL.H>
#include <UnivConstants.H>
#include <VarDefines.H>

#define E_ID 0
#define N2p_ID 1
#define O2p_ID 2
#define N4p_ID 3
#define O4p_ID 4
#define O2m_ID 5
#define Om_ID 6
#define N2s_ID 7
#define N2_ID 8
#define O2_ID 9

#define NUM_ELEMENTS 3
#define NUM_SPECIES 10
#define FIRST_ION 1
#define NUM_IONS 6
#define FIRST_NEUTRAL 7
#define NUM_NEUTRALS 1
#define NUM_REACTIONS 20

void CKSYMS_STR(amrex::Vector<std::string>& kname);
void CKSYME_STR(amrex::Vector<std::string>& ename);

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real mw(const int n)
{
    const amrex::Real w[NUM_SPECIES] = {0.00054858, 28.0134, 31.9988, 56.0268, 63.9976, 31.9988, 15.9994, 28.0134, 28.0134, 31.9988};
    return w[n];
}
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real imw(const int n) { return 1.0 / mw(n); }
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void get_mw(amrex::Real* w)
{
    for (int n = 0; n < NUM_SPECIES; ++n) w[n] = mw(n);
}

// Charge-signed mobilities [m^2/V/s] and diffusivities [m^2/s]: electrons
// -0.04 / 0.15, positive ions 2.0e-4 / 5.0e-6, negative ions -2.2e-4 / 5.5e-6,
// N2* 0 / 5.0e-6; NUM_SPECIES selects the electron-energy coefficients.
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real specMob(
    const int specid, const amrex::Real Te, const amrex::Real ndens,
    const amrex::Real emag, const amrex::Real T)
{
    if (specid == E_ID) return -0.04;
    if (specid == N2p_ID || specid == O2p_ID || specid == N4p_ID || specid == O4p_ID) return 2.0e-4;
    if (specid == O2m_ID || specid == Om_ID) return -2.2e-4;
    if (specid == NUM_SPECIES) return fivebythree * (-0.04);
    return 0.0;
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real specDiff(
    const int specid, const amrex::Real Te, const amrex::Real ndens,
    const amrex::Real emag, const amrex::Real T)
{
    if (specid == E_ID) return 0.15;
    if (specid == N2p_ID || specid == O2p_ID || specid == N4p_ID || specid == O4p_ID) return 5.0e-6;
    if (specid == O2m_ID || specid == Om_ID) return 5.5e-6;
    if (specid == N2s_ID) return 5.0e-6;
    if (specid == NUM_SPECIES) return fivebythree * 0.15;
    return 0.0;
}

// Molar production rates wdot [mol/m^3/s] from concentrations C [mol/m^3].
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKWC(
    const amrex::Real T, amrex::Real C[], amrex::Real wdot[],
    const amrex::Real Te, const amrex::Real EN, amrex::Real* ener_exch)
{
    for (int n = 0; n < NUM_SPECIES; ++n) wdot[n] = 0.0;
    *ener_exch = 0.0;
    const amrex::Real iEN = 1.0 / amrex::max(EN, 1.0e-30);
    { // R0: E + N2 -> E + E + N2p
        const amrex::Real k = 2.244901351169259e-15 * std::exp(-58.34706143438205 * iEN) * N_A;
        const amrex::Real q = k * C[E_ID] * C[N2_ID];
        wdot[E_ID] += q;
        wdot[N2_ID] -= q;
        wdot[N2p_ID] += q;
    }
    { // R1: E + O2 -> E + E + O2p
        const amrex::Real k = 2.5199099435512475e-16 * std::exp(-140.06640343410064 * iEN) * N_A;
        const amrex::Real q = k * C[E_ID] * C[O2_ID];
        wdot[E_ID] += q;
        wdot[O2_ID] -= q;
        wdot[O2p_ID] += q;
    }
    { // R2: E + N2 -> E + N2s
        const amrex::Real k = 5.563222135107969e-18 * std::exp(-248.25846047038115 * iEN) * N_A;
        const amrex::Real q = k * C[E_ID] * C[N2_ID];
        wdot[N2_ID] -= q;
        wdot[N2s_ID] += q;
    }
    { // R3: E + O2 -> O2m
        const amrex::Real k = 2.8008403075187547e-17 * std::exp(-188.26640873560072 * iEN) * N_A;
        const amrex::Real q = k * C[E_ID] * C[O2_ID];
        wdot[E_ID] -= q;
        wdot[O2_ID] -= q;
        wdot[O2m_ID] += q;
    }
    { // R4: E + O2 -> Om
        const amrex::Real k = 1.36461023419777e-16 * std::exp(-289.36424739755944 * iEN) * N_A;
        const amrex::Real q = k * C[E_ID] * C[O2_ID];
        wdot[E_ID] -= q;
        wdot[O2_ID] -= q;
        wdot[Om_ID] += q;
    }
    { // R5: N2p + N2 -> N4p
        const amrex::Real k = 1.3200362297954753e-18 * std::exp(-75.27187898360509 * iEN) * N_A;
        const amrex::Real q = k * C[N2p_ID] * C[N2_ID];
        wdot[N2p_ID] -= q;
        wdot[N2_ID] -= q;
        wdot[N4p_ID] += q;
    }
    { // R6: O2p + O2 -> O4p
        const amrex::Real k = 1.2275210179822035e-18 * std::exp(-263.8078588883028 * iEN) * N_A;
        const amrex::Real q = k * C[O2p_ID] * C[O2_ID];
        wdot[O2p_ID] -= q;
        wdot[O2_ID] -= q;
        wdot[O4p_ID] += q;
    }
    { // R7: N4p + O2 -> O2p + N2 + N2
        const amrex::Real k = 5.036633106094121e-15 * std::exp(-70.24738286177936 * iEN) * N_A;
        const amrex::Real q = k * C[N4p_ID] * C[O2_ID];
        wdot[N4p_ID] -= q;
        wdot[O2_ID] -= q;
        wdot[O2p_ID] += q;
        wdot[N2_ID] += 2.0 * q;
    }
    { // R8: N2p + O2 -> O2p + N2
        const amrex::Real k = 1.1551357234893301e-16 * std::exp(-101.54516783175623 * iEN) * N_A;
        const amrex::Real q = k * C[N2p_ID] * C[O2_ID];
        wdot[N2p_ID] -= q;
        wdot[O2_ID] -= q;
        wdot[O2p_ID] += q;
        wdot[N2_ID] += q;
    }
    { // R9: N4p + E -> N2 + N2
        const amrex::Real k = 4.899506332309811e-16 * std::exp(-272.8226520992474 * iEN) * N_A;
        const amrex::Real q = k * C[N4p_ID] * C[E_ID];
        wdot[N4p_ID] -= q;
        wdot[E_ID] -= q;
        wdot[N2_ID] += 2.0 * q;
    }
    { // R10: O4p + E -> O2 + O2
        const amrex::Real k = 1.1863399244901919e-16 * std::exp(-285.2344884050292 * iEN) * N_A;
        const amrex::Real q = k * C[O4p_ID] * C[E_ID];
        wdot[O4p_ID] -= q;
        wdot[E_ID] -= q;
        wdot[O2_ID] += 2.0 * q;
    }
    { // R11: N2p + E -> N2
        const amrex::Real k = 1.2500827712455404e-17 * std::exp(-115.42772267712526 * iEN) * N_A;
        const amrex::Real q = k * C[N2p_ID] * C[E_ID];
        wdot[N2p_ID] -= q;
        wdot[E_ID] -= q;
        wdot[N2_ID] += q;
    }
    { // R12: O2p + E -> O2
        const amrex::Real k = 4.6104873968569935e-17 * std::exp(-263.9590049629252 * iEN) * N_A;
        const amrex::Real q = k * C[O2p_ID] * C[E_ID];
        wdot[O2p_ID] -= q;
        wdot[E_ID] -= q;
        wdot[O2_ID] += q;
    }
    { // R13: O2m + N4p -> O2 + N2 + N2
        const amrex::Real k = 9.266845229473798e-18 * std::exp(-256.4505787626766 * iEN) * N_A;
        const amrex::Real q = k * C[O2m_ID] * C[N4p_ID];
        wdot[O2m_ID] -= q;
        wdot[N4p_ID] -= q;
        wdot[O2_ID] += q;
        wdot[N2_ID] += 2.0 * q;
    }
    { // R14: O2m + O4p -> O2 + O2 + O2
        const amrex::Real k = 9.833797460462614e-16 * std::exp(-183.82895097065878 * iEN) * N_A;
        const amrex::Real q = k * C[O2m_ID] * C[O4p_ID];
        wdot[O2m_ID] -= q;
        wdot[O4p_ID] -= q;
        wdot[O2_ID] += 3.0 * q;
    }
    { // R15: Om + O2p -> O2
        const amrex::Real k = 9.875220672012254e-17 * std::exp(-178.06649862747207 * iEN) * N_A;
        const amrex::Real q = k * C[Om_ID] * C[O2p_ID];
        wdot[Om_ID] -= q;
        wdot[O2p_ID] -= q;
        wdot[O2_ID] += q;
    }
    { // R16: O2m + N2s -> E + O2 + N2
        const amrex::Real k = 1.3465953394963778e-18 * std::exp(-162.9771819212857 * iEN) * N_A;
        const amrex::Real q = k * C[O2m_ID] * C[N2s_ID];
        wdot[O2m_ID] -= q;
        wdot[N2s_ID] -= q;
        wdot[E_ID] += q;
        wdot[O2_ID] += q;
        wdot[N2_ID] += q;
    }
    { // R17: Om + N2 -> E + N2
        const amrex::Real k = 4.714342256028545e-16 * std::exp(-113.9296011607804 * iEN) * N_A;
        const amrex::Real q = k * C[Om_ID] * C[N2_ID];
        wdot[Om_ID] -= q;
        wdot[E_ID] += q;
    }
    { // R18: N2s + O2 -> N2 + O2
        const amrex::Real k = 2.2089926705383792e-18 * std::exp(-290.5172865366043 * iEN) * N_A;
        const amrex::Real q = k * C[N2s_ID] * C[O2_ID];
        wdot[N2s_ID] -= q;
        wdot[N2_ID] += q;
    }
    { // R19: N2s + N2 -> N2 + N2
        const amrex::Real k = 5.671059868510319e-15 * std::exp(-134.77054557428136 * iEN) * N_A;
        const amrex::Real q = k * C[N2s_ID] * C[N2_ID];
        wdot[N2s_ID] -= q;
        wdot[N2_ID] += q;
    }
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real
etempLFA(const amrex::Real Tg, amrex::Real C[], const amrex::Real EN)
{
    return 23209.0; // 2 eV
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKCHRG(int kcharge[])
{
    kcharge[0] = -1;
    kcharge[1] = 1;
    kcharge[2] = 1;
    kcharge[3] = 1;
    kcharge[4] = 1;
    kcharge[5] = -1;
    kcharge[6] = -1;
    kcharge[7] = 0;
    kcharge[8] = 0;
    kcharge[9] = 0;
}
#endif
