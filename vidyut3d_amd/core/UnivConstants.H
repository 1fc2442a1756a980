// Physical constants and unit helpers used by vidyut3d case plugins.
// The values are the ones the reference solver uses
// (Source/UnivConstants.H:4-20); case headers include this file by name, so
// the names and values stay as the reference defines them.
#ifndef VD_UNIVCONSTANTS_H_
#define VD_UNIVCONSTANTS_H_

#define N_A 6.02214076e23
#define K_B 1.380649e-23
#define PI 3.14159265358979323846264338327950288
#define P_NTP 101325.0
#define T_NTP 293.15
#define N_NTP (P_NTP / K_B / T_NTP)
#define FARADCONST 96485.332
#define GASCONST 8.31446261815324
#define ECHARGE 1.60217662e-19
#define EPS0 8.854187817e-12
#define eV (ECHARGE / K_B)
#define FBYR (FARADCONST / GASCONST)
#define ME 9.10938188e-31
#define M_AMU 1.66054e-27
#define TDUNIT 1e-21
#define fivebythree (5.0 / 3.0)
#define twothird (2.0 / 3.0)

#endif
