// The LMEA argon case (cases/lmea3d) with an electron-temperature dependent
// electron diffusivity (see Chemistry.H): everything else is lmea3d's.
#include "../lmea3d/Chemistry.cpp"
