// Table-driven mass-action chemistry for case plugins.
//
// A case that describes its reactions as a table (rather than hand-written
// CKWC code) defines, before including this header:
//   VD_NUM_RXN                         number of reactions R
//   vd_rxn_reac[R][2]                  reactant species ids (-1: none)
//   vd_rxn_nu[R][NUM_SPECIES]          net stoichiometric change
//   vd_rxn_A[R]                        m^3/s (two reactants) or 1/s (one)
//   vd_rxn_B[R]                        Td; k = p_r A exp(-B/(E/N)), B = 0: constant
//   VD_RXN_PREFACTOR(r)                run-time multiplier p_r of reaction r
// and implements CKWC with vd_table_ckwc.  The same table drives the
// hand-written reverse-mode derivative (vd_table_ckwc_vjp) used by the
// adjoint: no derivative code is written per case.
//
// Units follow CKWC (Source/ScalarSolve.cpp:181-209 calls it with molar
// concentrations): C [mol/m^3], wdot [mol/m^3/s]; a two-reactant rate in
// m^3/s becomes m^3/mol/s through N_A.  E/N is floored at 1e-30 Td.
#ifndef VD_RXN_TABLE_H_
#define VD_RXN_TABLE_H_

#include <AMReX_REAL.H>
#include <AMReX_Gpu.H>
#include <UnivConstants.H>
#include <cmath>

// q_r without the prefactor (q0) for every reaction
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void vd_table_rates0(const amrex::Real C[], amrex::Real EN,
                                                             amrex::Real q0[])
{
    const amrex::Real ENs = (EN > 1.0e-30) ? EN : 1.0e-30;
#pragma unroll
    for (int r = 0; r < VD_NUM_RXN; ++r)
    {
        const int a = vd_rxn_reac[r][0], b = vd_rxn_reac[r][1];
        amrex::Real k = vd_rxn_A[r];
        if (b >= 0) k *= N_A;
        if (vd_rxn_B[r] != 0.0) k *= std::exp(-vd_rxn_B[r] / ENs);
        amrex::Real q = k;
        if (a >= 0) q *= C[a];
        if (b >= 0) q *= C[b];
        q0[r] = q;
    }
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void vd_table_ckwc(const amrex::Real C[], amrex::Real wdot[],
                                                           amrex::Real EN)
{
    amrex::Real q0[VD_NUM_RXN];
    vd_table_rates0(C, EN, q0);
    for (int s = 0; s < NUM_SPECIES; ++s) wdot[s] = 0.0;
#pragma unroll
    for (int r = 0; r < VD_NUM_RXN; ++r)
    {
        const amrex::Real q = VD_RXN_PREFACTOR(r) * q0[r];
#pragma unroll
        for (int s = 0; s < NUM_SPECIES; ++s)
            if (vd_rxn_nu[r][s] != 0) wdot[s] += amrex::Real(vd_rxn_nu[r][s]) * q;
    }
}

// Reverse mode of rxn = N_A wdot(n / N_A, EN): given rbar[s] = dJ/d rxn[s],
// adds dJ/dn[s] to nbar, returns dJ/dEN, and adds dJ/dp_r to pbar[r].
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real vd_table_ckwc_vjp(const amrex::Real n[], amrex::Real EN,
                                                                      const amrex::Real rbar[], amrex::Real nbar[],
                                                                      amrex::Real pbar[])
{
    amrex::Real C[NUM_SPECIES];
    for (int s = 0; s < NUM_SPECIES; ++s) C[s] = n[s] / N_A;
    amrex::Real q0[VD_NUM_RXN];
    vd_table_rates0(C, EN, q0);
    const bool floored = !(EN > 1.0e-30);
    amrex::Real enbar = 0.0;
#pragma unroll
    for (int r = 0; r < VD_NUM_RXN; ++r)
    {
        // dJ/dq_r = N_A sum_s nu_rs rbar_s
        amrex::Real w = 0.0;
#pragma unroll
        for (int s = 0; s < NUM_SPECIES; ++s)
            if (vd_rxn_nu[r][s] != 0) w += amrex::Real(vd_rxn_nu[r][s]) * rbar[s];
        w *= N_A;
        const amrex::Real p = VD_RXN_PREFACTOR(r);
        pbar[r] += w * q0[r];
        const amrex::Real wq = w * p;   // dJ/dq0 scaled by p: dJ/d(q) with q = p q0
        const int a = vd_rxn_reac[r][0], b = vd_rxn_reac[r][1];
        if (vd_rxn_B[r] != 0.0 && !floored) enbar += wq * q0[r] * vd_rxn_B[r] / (EN * EN);
        // q0 = k C_a C_b: dq0/dn_a = q0/C_a/N_A written without the division
        amrex::Real k = vd_rxn_A[r];
        if (b >= 0) k *= N_A;
        if (vd_rxn_B[r] != 0.0) k *= std::exp(-vd_rxn_B[r] / (floored ? 1.0e-30 : EN));
        if (a >= 0 && b >= 0)
        {
            nbar[a] += wq * k * C[b] / N_A;
            nbar[b] += wq * k * C[a] / N_A;
        } else if (a >= 0)
        {
            nbar[a] += wq * k / N_A;
        }
    }
    return enbar;
}

#endif
