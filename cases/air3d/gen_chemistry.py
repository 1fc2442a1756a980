"""Writes cases/air3d/Chemistry.H: the 3D air streamer mechanism with 20
Arrhenius-in-E/N reactions k = A exp(-B/(E/N)), A log-uniform in
[1e-18, 1e-14] m^3/s and B uniform in [50, 300] Td, drawn from seed 20261203
(numpy PCG64).  Like the chemistry generator the reference uses, this emits a
plain compile-time header; rerun it to regenerate."""
import os

import numpy as np

SEED = 20261203
SPECIES = ["E", "N2p", "O2p", "N4p", "O4p", "O2m", "Om", "N2s", "N2", "O2"]
CHARGE = [-1, 1, 1, 1, 1, -1, -1, 0, 0, 0]
MW = [5.4858e-4, 28.0134, 31.9988, 56.0268, 63.9976, 31.9988, 15.9994, 28.0134, 28.0134, 31.9988]
# (reactants, products); all bimolecular, charge-conserving
REACTIONS = [
    (("E", "N2"), ("E", "E", "N2p")),
    (("E", "O2"), ("E", "E", "O2p")),
    (("E", "N2"), ("E", "N2s")),
    (("E", "O2"), ("O2m",)),
    (("E", "O2"), ("Om",)),
    (("N2p", "N2"), ("N4p",)),
    (("O2p", "O2"), ("O4p",)),
    (("N4p", "O2"), ("O2p", "N2", "N2")),
    (("N2p", "O2"), ("O2p", "N2")),
    (("N4p", "E"), ("N2", "N2")),
    (("O4p", "E"), ("O2", "O2")),
    (("N2p", "E"), ("N2",)),
    (("O2p", "E"), ("O2",)),
    (("O2m", "N4p"), ("O2", "N2", "N2")),
    (("O2m", "O4p"), ("O2", "O2", "O2")),
    (("Om", "O2p"), ("O2",)),
    (("O2m", "N2s"), ("E", "O2", "N2")),
    (("Om", "N2"), ("E", "N2")),
    (("N2s", "O2"), ("N2", "O2")),
    (("N2s", "N2"), ("N2", "N2")),
]


def draw():
    rng = np.random.default_rng(SEED)
    A = 10.0 ** rng.uniform(-18.0, -14.0, size=len(REACTIONS))
    B = rng.uniform(50.0, 300.0, size=len(REACTIONS))
    return A, B


def main():
    A, B = draw()
    ix = {s: i for i, s in enumerate(SPECIES)}
    for r, p in REACTIONS:
        assert sum(CHARGE[ix[s]] for s in r) == sum(CHARGE[ix[s]] for s in p)
    L = []
    w = L.append
    w("// Generated by gen_chemistry.py -- 3D air streamer mechanism (do not edit).")
    w("// 20 reactions k = A exp(-B/(E/N)) [m^3/s, Td], seed %d." % SEED)
    w("#ifndef MECHANISM_H\n#define MECHANISM_H\n")
    w("#include <AMReX_Gpu.H>\n#include <AMReX_REAL.H>\n#include <UnivConstants.H>\n#include <VarDefines.H>\n")
    for i, s in enumerate(SPECIES):
        w(f"#define {s}_ID {i}")
    w("\n#define NUM_ELEMENTS 3\n#define NUM_SPECIES %d" % len(SPECIES))
    w("#define FIRST_ION 1\n#define NUM_IONS 6\n#define FIRST_NEUTRAL 7\n#define NUM_NEUTRALS 1")
    w("#define NUM_REACTIONS %d\n" % len(REACTIONS))
    w("void CKSYMS_STR(amrex::Vector<std::string>& kname);")
    w("void CKSYME_STR(amrex::Vector<std::string>& ename);\n")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real mw(const int n)\n{")
    w("    const amrex::Real w[NUM_SPECIES] = {%s};" % ", ".join(repr(x) for x in MW))
    w("    return w[n];\n}")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real imw(const int n) { return 1.0 / mw(n); }")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void get_mw(amrex::Real* w)\n{")
    w("    for (int n = 0; n < NUM_SPECIES; ++n) w[n] = mw(n);\n}\n")
    w("""// Charge-signed mobilities [m^2/V/s] and diffusivities [m^2/s]: electrons
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
""")
    w("// Molar production rates wdot [mol/m^3/s] from concentrations C [mol/m^3].")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKWC(\n    const amrex::Real T, amrex::Real C[], amrex::Real wdot[],")
    w("    const amrex::Real Te, const amrex::Real EN, amrex::Real* ener_exch)\n{")
    w("    for (int n = 0; n < NUM_SPECIES; ++n) wdot[n] = 0.0;")
    w("    *ener_exch = 0.0;")
    w("    const amrex::Real iEN = 1.0 / amrex::max(EN, 1.0e-30);")
    for r, ((ra, rb), prods) in enumerate(REACTIONS):
        w(f"    {{ // R{r}: {' + '.join((ra, rb))} -> {' + '.join(prods)}")
        w(f"        const amrex::Real k = {float(A[r])!r} * std::exp({-float(B[r])!r} * iEN) * N_A;")
        w(f"        const amrex::Real q = k * C[{ra}_ID] * C[{rb}_ID];")
        net = {}
        for s in (ra, rb):
            net[s] = net.get(s, 0) - 1
        for s in prods:
            net[s] = net.get(s, 0) + 1
        for s, v in net.items():
            if v == 0:
                continue
            op = "+=" if v > 0 else "-="
            mag = abs(v)
            w(f"        wdot[{s}_ID] {op} {'' if mag == 1 else f'{mag}.0 * '}q;")
        w("    }")
    w("}\n")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real\netempLFA(const amrex::Real Tg, amrex::Real C[], const amrex::Real EN)\n{")
    w("    return 23209.0; // 2 eV\n}\n")
    w("AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void CKCHRG(int kcharge[])\n{")
    for i, c in enumerate(CHARGE):
        w(f"    kcharge[{i}] = {c};")
    w("}\n#endif")
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "Chemistry.H"), "w") as f:
        f.write("\n".join(L) + "\n")
    with open(os.path.join(here, "Chemistry.cpp"), "w") as f:
        f.write('#include "Chemistry.H"\n\nvoid CKSYMS_STR(amrex::Vector<std::string>& kname)\n{\n    kname = {%s};\n}\n'
                % ", ".join(f'"{s}"' for s in SPECIES))
        f.write('void CKSYME_STR(amrex::Vector<std::string>& ename) { ename = {"E", "N", "O"}; }\n')


if __name__ == "__main__":
    main()
