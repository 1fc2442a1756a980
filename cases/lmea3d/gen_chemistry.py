"""Writes cases/lmea3d/Chemistry.H: a 3D argon mechanism for the local
mean energy approximation (electron energy equation).  The five
electron-impact rates are k = A exp(-C/eps) with eps = 1.5 k_B Te / e the mean
electron energy in eV, A log-uniform in [1e-16, 1e-13] m^3/s and C uniform in
[5, 20] eV, drawn from seed 20261203 (numpy PCG64); the heavy-particle rates
are constants.  Each inelastic electron-impact reaction removes its
threshold energy from the electrons (15.8 eV per ionization), scaled by the
superelastic factor (tanh((Te - Tg)/0.5) + 1)/2, and elastic collisions with
Ar exchange 3 k_B (m_e/m_Ar)(Te - Tg) per collision.  Rerun to regenerate."""
import os

import numpy as np

SEED = 20261203
SPECIES = ["E", "ARp", "AR2p", "ARm", "ARr", "AR"]
CHARGE = [-1, 1, 1, 0, 0, 0]
MW = [5.485799e-4, 39.947451, 79.895451, 39.948, 39.948, 39.948]
# (reactants, products, threshold eV or None); electron-impact first
EIMPACT = [
    (("E", "AR"), ("E", "E", "ARp"), 15.8),    # ionization
    (("E", "AR"), ("E", "ARm"), 11.55),        # metastable excitation
    (("E", "AR"), ("E", "ARr"), 11.62),        # resonant excitation
    (("E", "ARm"), ("E", "E", "ARp"), 4.21),   # stepwise ionization
    (("E", "ARr"), ("E", "E", "ARp"), 4.14),   # stepwise ionization
]
CONSTANT = [
    (("ARp", "AR"), ("AR2p",), 2.5e-19),       # molecular ion conversion (effective two-body)
    (("E", "AR2p"), ("ARm", "AR"), 1.0e-13),   # dissociative recombination
    (("ARm", "ARm"), ("E", "ARp", "AR"), 6.2e-16),  # metastable pooling
    (("ARm", "AR"), ("AR", "AR"), 2.1e-21),    # two-body quenching
    (("ARr",), ("AR",), 1.0e5),                # radiative decay (trapped), 1/s
    (("E", "ARp"), ("AR",), 1.0e-13),          # recombination
]
K_ELASTIC = 1.0e-13                            # m^3/s, E + AR elastic


def draw():
    rng = np.random.default_rng(SEED)
    A = 10.0 ** rng.uniform(-16.0, -13.0, size=len(EIMPACT))
    C = rng.uniform(5.0, 20.0, size=len(EIMPACT))
    return A, C


def main():
    A, C = draw()
    ix = {s: i for i, s in enumerate(SPECIES)}
    allr = [(r, p) for r, p, _ in EIMPACT] + [(r, p) for r, p, _ in CONSTANT]
    for r, p in allr:
        assert sum(CHARGE[ix[s]] for s in r) == sum(CHARGE[ix[s]] for s in p)
    L = []
    w = L.append
    w("// Generated by gen_chemistry.py -- 3D argon LMEA mechanism (do not edit).")
    w("// Electron-impact k = A exp(-C/eps) [m^3/s, eV], seed %d; constant rates otherwise." % SEED)
    w("#ifndef MECHANISM_H\n#define MECHANISM_H\n")
    w("#include <AMReX_Gpu.H>\n#include <AMReX_REAL.H>\n#include <UnivConstants.H>\n#include <VarDefines.H>\n")
    for i, s in enumerate(SPECIES):
        w(f"#define {s}_ID {i}")
    w("\n#define NUM_ELEMENTS 2\n#define NUM_SPECIES %d" % len(SPECIES))
    w("#define FIRST_ION 1\n#define NUM_IONS 2\n#define FIRST_NEUTRAL 3\n#define NUM_NEUTRALS 2")
    w("#define NUM_REACTIONS %d\n" % len(allr))
    w("void CKSYMS_STR(amrex::Vector<std::string>& kname);")
    w("void CKSYME_STR(amrex::Vector<std::string>& ename);\n")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real mw(const int n)\n{")
    w("    const amrex::Real w[NUM_SPECIES] = {%s};" % ", ".join(repr(x) for x in MW))
    w("    return w[n];\n}")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real imw(const int n) { return 1.0 / mw(n); }")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void get_mw(amrex::Real* w)\n{")
    w("    for (int n = 0; n < NUM_SPECIES; ++n) w[n] = mw(n);\n}\n")
    w("""// Charge-signed mobilities [m^2/V/s] and diffusivities [m^2/s]: electrons
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
    if (specid == E_ID) return 0.1;
    if (specid == NUM_SPECIES) return fivebythree * 0.1;
    return 1.0e-5;
}
""")
    w("// electron-impact table: A [m^3/s], C [eV], threshold [eV]")
    w("constexpr int lm_ne = %d;" % len(EIMPACT))
    w("constexpr double lm_A[lm_ne] = {%s};" % ", ".join(f"{a:.17g}" for a in A))
    w("constexpr double lm_C[lm_ne] = {%s};" % ", ".join(f"{c:.17g}" for c in C))
    w("constexpr double lm_eexci[lm_ne] = {%s};" % ", ".join(f"{e!r}" for _, _, e in EIMPACT))
    w("constexpr int lm_cn = %d;" % len(CONSTANT))
    w("constexpr double lm_K[lm_cn] = {%s};\n" % ", ".join(f"{k!r}" for _, _, k in CONSTANT))
    w("""// Molar production rates wdot [mol/m^3/s] from concentrations C [mol/m^3],
// and the electron energy exchange [W/m^3] (inelastic losses and elastic
// exchange, the form of the reference's comp_ener_exch).
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKWC(
    const amrex::Real T, amrex::Real C[], amrex::Real wdot[], const amrex::Real Te,
    const amrex::Real EN, amrex::Real* ener_exch)
{
    const amrex::Real eps = 1.5 * K_B * Te / ECHARGE;   // mean energy, eV
    const amrex::Real ieps = 1.0 / (eps > 1.0e-3 ? eps : 1.0e-3);
    const amrex::Real sefact = (std::tanh((Te - T) / 0.5) + 1.0) / 2.0;
    amrex::Real q[NUM_REACTIONS];""")
    for r, (re, pr, _) in enumerate(EIMPACT):
        terms = " * ".join(f"C[{s}_ID]" for s in re)
        w(f"    q[{r}] = lm_A[{r}] * N_A * std::exp(-lm_C[{r}] * ieps) * {terms};")
    for r, (re, pr, _) in enumerate(CONSTANT):
        k = f"lm_K[{r}]" + (" * N_A" if len(re) == 2 else "")
        terms = " * ".join(f"C[{s}_ID]" for s in re)
        w(f"    q[{len(EIMPACT) + r}] = {k} * {terms};")
    w("    for (int s = 0; s < NUM_SPECIES; ++s) wdot[s] = 0.0;")
    for r, (re, pr) in enumerate(allr):
        nu = {}
        for s in re:
            nu[s] = nu.get(s, 0) - 1
        for s in pr:
            nu[s] = nu.get(s, 0) + 1
        for s, v in nu.items():
            if v:
                w(f"    wdot[{s}_ID] += {float(v)!r} * q[{r}];")
    w("""    amrex::Real ex = 0.0;
    for (int r = 0; r < lm_ne; ++r) ex -= (lm_eexci[r] * ECHARGE) * N_A * q[r] * sefact;
    const amrex::Real qel = %r * N_A * C[E_ID] * C[AR_ID];
    ex -= 3.0 * K_B * (mw(E_ID) / mw(AR_ID)) * (Te - T) * qel * N_A;
    *ener_exch = ex;
}
""" % K_ELASTIC)
    w("""AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real
etempLFA(const amrex::Real Tg, amrex::Real C[], const amrex::Real EN)
{
    return 11604.5; // 1 eV
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKCHRG(int kcharge[])
{""")
    for i, s in enumerate(SPECIES):
        w(f"    kcharge[{s}_ID] = {CHARGE[i]};")
    w("}\n#endif")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "Chemistry.H")
    with open(out, "w") as f:
        f.write("\n".join(L) + "\n")
    print("wrote", out)
    for r in range(len(EIMPACT)):
        print(f"  R{r}: A={A[r]:.3e} C={C[r]:.2f} eV")


if __name__ == "__main__":
    main()
