#include "Chemistry.H"

void CKSYMS_STR(amrex::Vector<std::string>& kname) { kname = {"E", "ARp", "AR2p", "ARm", "ARr", "AR"}; }
void CKSYME_STR(amrex::Vector<std::string>& ename) { ename = {"E", "AR"}; }
