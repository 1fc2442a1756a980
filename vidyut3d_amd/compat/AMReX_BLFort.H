// Forwarding header: the subset of this AMReX header that case plugins use.
#include "AMReX_compat.H"
