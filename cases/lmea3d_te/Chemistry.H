// The LMEA argon case (cases/lmea3d) with the electron diffusivity (and the
// electron energy's, 5/3 of it) proportional to the electron temperature,
// D_e = 0.1 m^2/s x T_e / 1 eV: the initial T_e is uniform, so D_e is
// uniform in the first step and varies in space from the second on, which
// exercises the harmonic-mean (cell coefficient) species operator after a
// step whose coefficients were uniform.
#define LMEA_TE_DIFFUSION 1
#include "../lmea3d/Chemistry.H"
