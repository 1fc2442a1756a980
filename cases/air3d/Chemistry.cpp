#include "Chemistry.H"

void CKSYMS_STR(amrex::Vector<std::string>& kname)
{
    kname = {"E", "N2p", "O2p", "N4p", "O4p", "O2m", "Om", "N2s", "N2", "O2"};
}
void CKSYME_STR(amrex::Vector<std::string>& ename) { ename = {"E", "N", "O"}; }
