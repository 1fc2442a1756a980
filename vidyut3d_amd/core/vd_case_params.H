// Runtime case parameters.  The reference's case constants (rate
// coefficients, mobilities) are compile-time; a case that wants them
// settable at run time -- for parameter studies and for the gradients of the
// adjoint -- defines, before including this header,
//   VD_NUM_CASE_PARAMS      number of parameters
//   VD_CASE_PARAM_DEFAULTS  comma-separated default values
// and reads them with case_param(i) in its plugin functions (specMob,
// specDiff, CKWC, ...).  Device code reads a constant-memory copy, host code
// a host copy.  Both are process-wide, so each solver context keeps its own
// parameter vector and copies it in whenever it becomes the active context.
#ifndef VD_CASE_PARAMS_H_
#define VD_CASE_PARAMS_H_

#include <AMReX_REAL.H>
#include <AMReX_Gpu.H>

#if defined(VD_NUM_CASE_PARAMS)
#if defined(__HIPCC__)
static __constant__ amrex::Real vd_case_param_dev[VD_NUM_CASE_PARAMS] = {VD_CASE_PARAM_DEFAULTS};
#endif
static amrex::Real vd_case_param_host[VD_NUM_CASE_PARAMS] = {VD_CASE_PARAM_DEFAULTS};
static const amrex::Real vd_case_param_defaults[VD_NUM_CASE_PARAMS] = {VD_CASE_PARAM_DEFAULTS};

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real case_param(int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return vd_case_param_dev[i];
#else
    return vd_case_param_host[i];
#endif
}
#endif
#endif
