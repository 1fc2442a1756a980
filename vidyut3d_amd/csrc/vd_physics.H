// Device physics of the vidyut3d step on gfx950: default transport hooks,
// drift-flux reconstruction (upwind, WAF, WENO-JS/M/Z), flux divergence plus
// reaction source, charge density, E-field, time scales and LFA.
//
// Included after the case plugin headers (Chemistry.H, ProbParm.H, Prob.H,
// UserFunctions.H, UserSources.H), so NUM_SPECIES, specMob, specDiff, CKWC,
// CKCHRG and the user_transport/user_sources hooks resolve to the case's.
#ifndef VD_PHYSICS_H_
#define VD_PHYSICS_H_

#include "vd_grid.H"
#include "vd_reduce.H"
#include "vd_amr.H"
#include <type_traits>

namespace vd {

using amrex::Real;
constexpr int SD = AMREX_SPACEDIM;

// Everything a kernel needs to build the reference's hook arguments.
struct PhysParams
{
    Grid g;
    Real Tg, Pg;
    Real time, dt;
    Real timescale;
    int user_vel, user_species, user_potential, user_photo;
    int do_transport, do_reactions;
    int eidx;
};

// View of a level's state for the plugin hooks, indexed by global level
// indices (the box starts at g.off).
__device__ inline amrex::Array4<Real> state_view(Real* s, const Grid& g)
{
    amrex::Array4<Real> a(s, amrex::Dim3{g.off[0], g.off[1], g.off[2]},
                          amrex::Dim3{g.off[0] + g.n[0], g.off[1] + g.n[1], g.off[2] + g.n[2]}, NVAR);
    a.ghost = amrex::VD_GHOST_ACTIVE | (g.periodic[0] ? 1 : 0) | (g.periodic[1] ? 2 : 0) | (g.periodic[2] ? 4 : 0);
    return a;
}

template <class T>
__device__ inline amrex::GpuArray<Real, SD> garr(const T* a)
{
    amrex::GpuArray<Real, SD> r;
    for (int d = 0; d < SD; ++d) r[d] = a[d];
    return r;
}

__host__ __device__ inline int species_charge(int sp)
{
    int ch[NUM_SPECIES];
    CKCHRG(ch);
    return ch[sp];
}

// Applied electrode voltage (DC, sine, triangle), Source/HelperFuncs.H:125-161.
__host__ __device__ inline Real applied_voltage(Real t, int end, int vprof, Real v1, Real v2, Real vfreq, Real vdur,
                                                Real vcen)
{
    const Real amp = (end == -1) ? v1 : v2;
    if (vprof == 1) return sin(2.0 * PI * vfreq * t) * amp;
    if (vprof == 2)
    {
        const Real half = vdur / 2.0;
        const Real off = (t <= vcen) ? (vcen - t) : (t - vcen);
        return (off < half) ? (1.0 - off / half) * amp : 0.0;
    }
    return amp;
}

// ---------------------------------------------------------------------------
// Cell quantities every drift velocity needs: sum of densities, |E|, Te.
struct CellQ
{
    Real ndens, emag, te;
};

__device__ inline CellQ cell_q(const Real* __restrict__ s, long long N, long long o)
{
    CellQ q;
    Real nd = 0.0;
#pragma unroll
    for (int sp = 0; sp < NUM_SPECIES; ++sp) nd += s[(long long)sp * N + o];
    q.ndens = nd;
    const Real ex = s[(long long)EFX_ID * N + o], ey = s[(long long)EFY_ID * N + o], ez = s[(long long)EFZ_ID * N + o];
    q.emag = sqrt(ex * ex + ey * ey + ez * ez);
    q.te = s[(long long)ETEMP_ID * N + o];
    return q;
}

// compute_axisym_correction (Source/ScalarSolve.cpp:331-378): in RZ the
// explicit source of species sp loses the radial drift term mu n E_r / r
// (the Cartesian divergence lacks it), evaluated as the reference writes it.
__device__ __forceinline__ Real axisym_drift(const Grid& g, const Real* __restrict__ s, long long N, long long o,
                                             int i, int sp, const CellQ& q, Real Tg)
{
    const Real r = fabs(g.problo[0] + (i + g.off[0] + 0.5) * g.dx[0]);
    const Real mu = specMob(sp, q.te, q.ndens, q.emag, Tg);
    return mu * s[(long long)sp * N + o] * s[(long long)EFX_ID * N + o] / r;
}

// ---------------------------------------------------------------------------
// Reconstructions (Source/compute_explicit_flux.H:16-214).

// x / y; in FP32 (the mixed form only) the hardware reciprocal (1 ulp)
// times x: every divisor here is at least eps = 1e-12 in scaled units or a
// sum of positive weights, far from the denormal range, and the FP32 faces
// carry 6e-8 of rounding anyway
template <class T>
__device__ __forceinline__ T wdiv(T x, T y)
{
    if constexpr (std::is_same<T, float>::value) return x * __builtin_amdgcn_rcpf(y);
    else return x / y;
}

// num / den in FP64 for a den in the normal range and a quotient far from
// overflow: the hardware reciprocal (relative error d) refined by one
// Newton-Raphson step (e = 1 - den r, r += r e: error d^2), then the
// quotient corrected once by its residual (q += r (num - den q)), which
// leaves d^2 (d^2 + 2^-53) beside the final rounding: within an ulp of the
// rounded quotient for any d below 2^-14, in 6 instructions instead of the
// 11 of the IEEE sequence with its scaling and special-case fix-up
__device__ __forceinline__ double ddiv_nr(double num, double den)
{
    double r = __builtin_amdgcn_rcp(den);
    const double e = fma(-den, r, 1.0);
    r = fma(r, e, r);
    const double q = num * r;
    return fma(r, fma(-den, q, num), q);
}

template <class T>
__device__ __forceinline__ T weno_gk(T w, T g)
{
    return wdiv(w * (g + g * g - T(3) * g * w + w * w), g * g + w * (T(1) - T(2) * g));
}

// Fifth-order WENO value at i+1/2 from a..e = phi(i-2..i+2); scheme 1 JS,
// 2 M, 3 Z (Jiang-Shu 1996; Henrick et al. 2005; Borges et al. 2008).
// weno5_combine: the nonlinear weights from the smoothness indicators
// b1..b3 and the value from the three candidate polynomials p1..p3.
template <class T, int SCHEME>
__device__ __forceinline__ T weno5_combine(T b1, T b2, T b3, T p1, T p2, T p3);

template <class T, int SCHEME>
__device__ __forceinline__ T weno5(T a, T b, T c, T d, T e)
{
    const T p1 = (T(1) / T(3)) * a - (T(7) / T(6)) * b + (T(11) / T(6)) * c;
    const T p2 = -(T(1) / T(6)) * b + (T(5) / T(6)) * c + (T(1) / T(3)) * d;
    const T p3 = (T(1) / T(3)) * c + (T(5) / T(6)) * d - (T(1) / T(6)) * e;
    const T s1a = a - T(2) * b + c, s1b = a - T(4) * b + T(3) * c;
    const T s2a = b - T(2) * c + d, s2b = b - d;
    const T s3a = c - T(2) * d + e, s3b = T(3) * c - T(4) * d + e;
    const T b1 = (T(13) / T(12)) * (s1a * s1a) + (T(1) / T(4)) * (s1b * s1b);
    const T b2 = (T(13) / T(12)) * (s2a * s2a) + (T(1) / T(4)) * (s2b * s2b);
    const T b3 = (T(13) / T(12)) * (s3a * s3a) + (T(1) / T(4)) * (s3b * s3b);
    return weno5_combine<T, SCHEME>(b1, b2, b3, p1, p2, p3);
}

// The reconstruction from q_k = beta_k + eps (every scheme needs the
// indicators only through q_k and tau = |beta_1 - beta_3|, here taken as
// |q_1 - q_3|).
template <class T, int SCHEME>
__device__ __forceinline__ T weno5_combine_q(T q1, T q2, T q3, T p1, T p2, T p3)
{
    const T g1 = T(0.1), g2 = T(0.6), g3 = T(0.3);
    T a1, a2, a3;
    if constexpr (std::is_same<T, double>::value && SCHEME != 2)
    {
        // The normalised weights a_k / sum a_j do not change when every a_k
        // is multiplied by the same factor:
        //   JS: a_k = g_k / q_k^2            -> g_k prod_{j!=k} q_j^2
        //   Z:  a_k = g_k (1 + (tau/q_k)^2)  -> g_k (q_k^2 + tau^2) prod_{j!=k} q_j^2
        // after scaling the q by 2^-e (exact; e the frexp exponent of the
        // largest q, so no product overflows).  The q are positive, so the
        // largest has the largest high word, whose exponent field gives e.
        // One division per call instead of four.  If the sum underflows (q
        // ratios beyond 1e139, far outside any density or flux range), the
        // per-weight form runs.
        const unsigned h = max(max((unsigned)__double2hiint(q1), (unsigned)__double2hiint(q2)),
                               (unsigned)__double2hiint(q3));
        const int ex = (int)((h >> 20) & 0x7ffu) - 1022;
        const T Q1 = ldexp(q1, -ex), Q2 = ldexp(q2, -ex), Q3 = ldexp(q3, -ex);
        const T P1 = Q1 * Q1, P2 = Q2 * Q2, P3 = Q3 * Q3;
        if constexpr (SCHEME == 3)
        {
            const T tau = fabs(Q1 - Q3);
            const T t2 = tau * tau;
            a1 = g1 * (P1 + t2) * (P2 * P3);
            a2 = g2 * (P2 + t2) * (P1 * P3);
            a3 = g3 * (P3 + t2) * (P1 * P2);
        } else
        {
            a1 = g1 * (P2 * P3);
            a2 = g2 * (P1 * P3);
            a3 = g3 * (P1 * P2);
        }
        const T s = a1 + a2 + a3;
        // s in [1e-280, 3]: the sum of positive weights scaled by 2^-2e
        if (s >= T(1e-280)) return ddiv_nr(a1 * p1 + a2 * p2 + a3 * p3, s);
    }
    if constexpr (SCHEME == 3)
    {
        const T tau = fabs(q1 - q3);
        const T r1 = wdiv(tau, q1), r2 = wdiv(tau, q2), r3 = wdiv(tau, q3);
        a1 = g1 * (T(1) + r1 * r1);
        a2 = g2 * (T(1) + r2 * r2);
        a3 = g3 * (T(1) + r3 * r3);
    } else
    {
        a1 = wdiv(g1, q1 * q1);
        a2 = wdiv(g2, q2 * q2);
        a3 = wdiv(g3, q3 * q3);
        if constexpr (SCHEME == 2)
        {
            const T s = a1 + a2 + a3;
            const T w1 = wdiv(a1, s), w2 = wdiv(a2, s), w3 = wdiv(a3, s);
            a1 = weno_gk(w1, g1);
            a2 = weno_gk(w2, g2);
            a3 = weno_gk(w3, g3);
        }
    }
    // sum_k (a_k/s) p_k with one division: (sum_k a_k p_k)/s
    const T s = a1 + a2 + a3;
    return wdiv(a1 * p1 + a2 * p2 + a3 * p3, s);
}

template <class T, int SCHEME>
__device__ __forceinline__ T weno5_combine(T b1, T b2, T b3, T p1, T p2, T p3)
{
    // 1e-20 is the reference's epsilon (Source/compute_explicit_flux.H:41).
    // The FP32 path works on values scaled to [-1, 1], where that epsilon is
    // effectively zero; 1e-12 keeps (tau/(beta+eps))^2 and 1/(beta+eps)^2
    // inside the FP32 range when a sub-stencil is flat.
    const T eps = std::is_same<T, double>::value ? T(1e-20) : T(1e-12);
    return weno5_combine_q<T, SCHEME>(b1 + eps, b2 + eps, b3 + eps, p1, p2, p3);
}

// Left- (a..e = w[0..4]) and right-biased (w[5..1]) WENO5 values of the
// face between cells w[2] and w[3] in one pass: the two reconstructions
// share the terms S = 13/12 (x - 2y + z)^2 of the triples centred on w[2]
// and w[3], and two of their three candidate values (the left p2 is the
// right p3 and the other way round), so 4 of 6 S terms and 4 of 6
// candidates are evaluated.  Same formulas as weno5 (summation order aside).
template <class T, int SCHEME>
__device__ __forceinline__ void weno5_pair(const T* w, T& ph, T& mh)
{
    const T sa0 = w[0] - T(2) * w[1] + w[2], sa1 = w[1] - T(2) * w[2] + w[3];
    const T sa2 = w[2] - T(2) * w[3] + w[4], sa3 = w[3] - T(2) * w[4] + w[5];
    const T c13 = T(13) / T(12), q = T(1) / T(4);
    const T S0 = c13 * (sa0 * sa0), S1 = c13 * (sa1 * sa1), S2 = c13 * (sa2 * sa2), S3 = c13 * (sa3 * sa3);
    const T l1 = w[0] - T(4) * w[1] + T(3) * w[2], l2 = w[1] - w[3], l3 = T(3) * w[2] - T(4) * w[3] + w[4];
    const T r1 = w[5] - T(4) * w[4] + T(3) * w[3], r2 = w[4] - w[2], r3 = T(3) * w[3] - T(4) * w[2] + w[1];
    const T E = (T(1) / T(3)) * w[0] - (T(7) / T(6)) * w[1] + (T(11) / T(6)) * w[2];
    const T R = -(T(1) / T(6)) * w[1] + (T(5) / T(6)) * w[2] + (T(1) / T(3)) * w[3];
    const T L = (T(1) / T(3)) * w[2] + (T(5) / T(6)) * w[3] - (T(1) / T(6)) * w[4];
    const T Er = (T(1) / T(3)) * w[5] - (T(7) / T(6)) * w[4] + (T(11) / T(6)) * w[3];
    ph = weno5_combine<T, SCHEME>(S0 + q * (l1 * l1), S1 + q * (l2 * l2), S2 + q * (l3 * l3), E, R, L);
    mh = weno5_combine<T, SCHEME>(S3 + q * (r1 * r1), S2 + q * (r2 * r2), S1 + q * (r3 * r3), Er, L, R);
}

// WENO5 in a reduced precision T with exact power-of-two scaling: the five
// values are scaled by 2^-e (e from the largest magnitude) so the squared
// smoothness indicators stay inside T's range, and the result is scaled
// back; scaling by a power of two is exact.
template <class T, int SCHEME>
__device__ __forceinline__ T weno5_scaled(T a, T b, T c, T d, T e)
{
    const T m = fmax(fmax(fmax(fabs(a), fabs(b)), fmax(fabs(c), fabs(d))), fabs(e));
    if (m == T(0)) return T(0);
    int ex;
    frexp(m, &ex);
    const T sa = ldexp(a, -ex), sb = ldexp(b, -ex), sc = ldexp(c, -ex), sd = ldexp(d, -ex), se = ldexp(e, -ex);
    return ldexp(weno5<T, SCHEME>(sa, sb, sc, sd, se), ex);
}

// MINBEE slope limiter for the WAF flux (Toro, Riemann Solvers, ch. 14).
__device__ inline Real waf_limiter(Real c, Real dl, Real dm, Real dr)
{
    const Real heps = 1.0e-30;
    const Real r = (c >= 0.0) ? dl / (dm + heps) : dr / (dm + heps);
    if (r <= 0.0) return 1.0;
    if (r <= 1.0) return 1.0 - (1.0 - fabs(c)) * r;
    return fabs(c);
}

// ---------------------------------------------------------------------------
// Face flux kernel for direction DIR and reconstruction order HYP (1, 2, 5).
// One thread per face; loops the transported species.  Writes
// F[c*nf + face].  Faces on IHNEU/AXIS domain boundaries get zero flux
// (Source/compute_explicit_flux.H:401-414).
struct FluxArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state;
    int nt;                  // transported species
    int tsp[vd::kMaxComp];   // species ids
    int bclo[vd::kMaxComp][3], bchi[vd::kMaxComp][3];
    int weno_scheme;
    Real dxmin;
    void* F;                 // nt * nf faces, in the flux precision TF
    long long nf;
    // fine levels: the state read through a ghosted copy (components
    // [0, EFZ_ID] on the box grown by sng; null on level 0, where ghosts
    // are index-mapped)
    const Real* sg;
    int sng;
};

// TF = double: the reference arithmetic.  TF = float (mixed form): the
// reconstruction and face flux in FP32 (WENO with exact power-of-two
// scaling), faces stored in FP32.
template <int DIR, int HYP, int SCHEME, class TF>
__global__ void __launch_bounds__(256) face_flux_kernel(FluxArgs A)
{
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int e0 = g.n[0] + (DIR == 0), e1 = g.n[1] + (DIR == 1), e2 = g.n[2] + (DIR == 2);
    // one thread per face, the faces of a k-plane numbered x fastest (a
    // row of e0 = n+1 faces does not leave a mostly idle wavefront per row)
    const unsigned t = bI_.x * blockDim.x + threadIdx.x;
    const int k = bI_.z;
    if (t >= (unsigned)e0 * (unsigned)e1 || k >= e2) return;
    const int j = (int)(t / (unsigned)e0);
    const int i = (int)(t - (unsigned)j * (unsigned)e0);
    const long long fo = (long long)i + (long long)e0 * ((long long)j + (long long)e1 * k);
    const int I[3] = {i, j, k};
    const int nd = g.n[DIR];
    const int per = g.periodic[DIR];
    constexpr int NS = (HYP == 1) ? 2 : (HYP == 2 ? 4 : 6);
    constexpr int OFF = (HYP == 1) ? 1 : (HYP == 2 ? 2 : 3);
    // stencil cells: face index + s - OFF, s = 0..NS-1; (OFF-1) is the left cell
    long long co[NS];
    int civ[NS];
    // source array: the level state (ghosts index-mapped) or its ghosted copy
    const Real* src = A.state;
    long long N = g.N;
    long long st, base;
    if (A.sg)
    {
        const int q = A.sng;
        const long long gx = g.n[0] + 2 * q, gy = (SD > 1) ? g.n[1] + 2 * q : 1, gz = (SD > 2) ? g.n[2] + 2 * q : 1;
        src = A.sg;
        N = gx * gy * gz;
        st = (DIR == 0) ? 1 : (DIR == 1 ? gx : gx * gy);
        base = (long long)(DIR == 0 ? 0 : i + q) +
               gx * ((long long)(DIR == 1 ? 0 : (SD > 1 ? j + q : 0)) + gy * (DIR == 2 ? 0 : (SD > 2 ? k + q : 0)));
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            civ[s] = I[DIR] + s - OFF;
            co[s] = base + (long long)(civ[s] + q) * st;
        }
    } else
    {
        st = (DIR == 0) ? 1 : (DIR == 1 ? (long long)g.n[0] : (long long)g.n[0] * g.n[1]);
        base = (long long)(DIR == 0 ? 0 : i) + (long long)g.n[0] * ((long long)(DIR == 1 ? 0 : j) +
                                                                    (long long)g.n[1] * (DIR == 2 ? 0 : k));
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            civ[s] = I[DIR] + s - OFF;
            co[s] = base + (long long)ghost_map(civ[s], nd, per) * st;
        }
    }
    CellQ q[NS];
    Real E[NS];
    if (!A.P.user_vel)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            q[s] = cell_q(src, N, co[s]);
            E[s] = src[(long long)(EFX_ID + DIR) * N + co[s]];
        }
    }
    // hook view: the ghosted copy (global indices, components up to EFZ_ID)
    // or the level state with index-mapped ghosts
    amrex::Array4<Real> sv = state_view(A.state, g);
    if (A.sg)
    {
        const int q = A.sng;
        sv = amrex::Array4<Real>(const_cast<Real*>(A.sg), amrex::Dim3{g.off[0] - q, SD > 1 ? g.off[1] - q : 0, SD > 2 ? g.off[2] - q : 0},
                                 amrex::Dim3{g.off[0] + g.n[0] + q, SD > 1 ? g.off[1] + g.n[1] + q : 1,
                                             SD > 2 ? g.off[2] + g.n[2] + q : 1},
                                 EFZ_ID + 1);
    }
    const Real Tg = A.P.Tg;
    for (int c = 0; c < A.nt; ++c)
    {
        const int sp = A.tsp[c];
        Real v[NS], u[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            u[s] = src[(long long)sp * N + co[s]];
            if (A.P.user_vel)
            {
                amrex::IntVect cv;
                for (int d = 0; d < SD; ++d) cv[d] = I[d] + g.off[d];
                cv[DIR] = civ[s] + g.off[DIR];
                v[s] = user_transport::compute_vel(cv, DIR, sp, sv, garr(g.dx), A.P.time, *A.pp, Tg, A.P.Pg);
            } else
            {
                v[s] = specMob(sp, q[s].te, q[s].ndens, q[s].emag, Tg) * E[s];
            }
        }
        TF fx;
        const TF vL = TF(v[OFF - 1]), vR = TF(v[OFF]);
        const TF pL = TF(u[OFF - 1]), pR = TF(u[OFF]);
        if constexpr (HYP == 1)
        {
            const TF fL = pL * vL, fR = pR * vR;
            const TF sg = copysign(TF(1), TF(0.5) * (vL + vR));
            fx = fL * TF(0.5) * (TF(1) + sg) + fR * TF(0.5) * (TF(1) - sg);
        } else if constexpr (HYP == 2)
        {
            const TF dl = pL - TF(u[0]), dm = pR - pL, dr = TF(u[3]) - pR;
            const TF vm = TF(0.5) * (vL + vR);
            const TF sg = copysign(TF(1), vm);
            const TF cc = vm * TF(A.P.dt) / TF(A.dxmin);
            const TF fL = pL * vL, fR = pR * vR;
            const TF lim = TF(waf_limiter(double(cc), double(dl), double(dm), double(dr)));
            fx = TF(0.5) * (fL + fR) - TF(0.5) * sg * lim * (fR - fL);
        } else
        {
            TF fv[6], uv[6];
#pragma unroll
            for (int q = 0; q < 6; ++q)
            {
                uv[q] = TF(u[q]);
                fv[q] = TF(u[q] * v[q]);
            }
            const TF lam = fabs(TF(0.5) * (vL + vR));
            TF fph, fmh, uph, umh;
            if constexpr (std::is_same<TF, double>::value)
            {
                weno5_pair<TF, SCHEME>(fv, fph, fmh);
                weno5_pair<TF, SCHEME>(uv, uph, umh);
            } else
            {
                fph = weno5_scaled<TF, SCHEME>(fv[0], fv[1], fv[2], fv[3], fv[4]);
                fmh = weno5_scaled<TF, SCHEME>(fv[5], fv[4], fv[3], fv[2], fv[1]);
                uph = weno5_scaled<TF, SCHEME>(uv[0], uv[1], uv[2], uv[3], uv[4]);
                umh = weno5_scaled<TF, SCHEME>(uv[5], uv[4], uv[3], uv[2], uv[1]);
            }
            fx = TF(0.5) * (fph + lam * uph) + TF(0.5) * (fmh - lam * umh);
        }
        if (I[DIR] == 0 && (A.bclo[c][DIR] == IHNEUBC || A.bclo[c][DIR] == AXISBC)) fx = TF(0);
        if (I[DIR] == nd && (A.bchi[c][DIR] == IHNEUBC || A.bchi[c][DIR] == AXISBC)) fx = TF(0);
        static_cast<TF*>(A.F)[(long long)c * A.nf + fo] = fx;
    }
}

// ---------------------------------------------------------------------------
// WENO5 face fluxes along DIR = 1 (y) or 2 (z) by marching: one thread per
// (transverse column, chunk of faces, species), lanes along x (coalesced),
// walking the faces of its chunk in order.  Each cell enters the window
// once: its density, drift velocity and f = n v are computed once (the
// per-face kernel recomputes them for each of the six faces that read the
// cell), and the smoothness indicators are rebuilt from per-triple pieces:
// for the triple (x, y, z) of cells m-1, m, m+1
//     S = 13/12 (x - 2y + z)^2,   P = 1/4 (x - 4y + 3z)^2,
//     Q = 1/4 (x - z)^2,          R = 1/4 (3x - 4y + z)^2,
// the face between cells k-1 and k takes, left-biased, b = (S+P)_{k-2},
// (S+Q)_{k-1}, (S+R)_k and, right-biased, (S+R)_{k+1}, (S+Q)_k, (S+P)_{k-1}
// (the indicators of Source/compute_explicit_flux.H:27-142, grouped by
// triple), so one new triple per face replaces six indicator evaluations.
// Same reconstruction and Lax-Friedrichs combination as face_flux_kernel
// (HYP = 5, FP64, default velocities, level 0 with index-mapped ghosts).
struct Triple
{
    double sp, sq, sr;   // S+P, S+Q, S+R, each + eps
};
// the pieces carry the epsilon 1e-20 of q = beta + eps, added once per
// triple instead of once per reconstruction (each piece enters two)
__device__ __forceinline__ Triple weno_triple(double x, double y, double z)
{
    const double sa = x - 2.0 * y + z;
    const double S = (13.0 / 12.0) * (sa * sa) + 1e-20;
    const double a = x - 4.0 * y + 3.0 * z, q = x - z, r = 3.0 * x - 4.0 * y + z;
    Triple t;
    t.sp = S + 0.25 * (a * a);
    t.sq = S + 0.25 * (q * q);
    t.sr = S + 0.25 * (r * r);
    return t;
}

// FP32 form (the mixed precision): a triple is built from its three cells
// scaled by 2^-e, e the exponent of their largest magnitude, and keeps e.
// A face brings its three indicators to the exponent E of its whole
// five-cell stencil (the largest of the three) and its candidate values to
// 2^-E, combines, and scales the result back.  Scaling by a power of two is
// exact, so this is weno5_scaled's per-stencil scaling (squared indicators
// inside the FP32 range, eps = 1e-12 in scaled units) without rescaling
// every cell for each of the six faces that read it.
struct TripleF
{
    float sp, sq, sr;
    int e;
};
__device__ __forceinline__ TripleF weno_triple_f(float x, float y, float z)
{
    const float m = fmaxf(fmaxf(fabsf(x), fabsf(y)), fabsf(z));
    int e = 0;
    if (m != 0.0f) (void)frexpf(m, &e);
    x = ldexpf(x, -e);
    y = ldexpf(y, -e);
    z = ldexpf(z, -e);
    const float sa = x - 2.0f * y + z;
    const float S = (13.0f / 12.0f) * (sa * sa);
    const float a = x - 4.0f * y + 3.0f * z, q = x - z, r = 3.0f * x - 4.0f * y + z;
    TripleF t;
    t.sp = S + 0.25f * (a * a);
    t.sq = S + 0.25f * (q * q);
    t.sr = S + 0.25f * (r * r);
    t.e = e;
    return t;
}
template <int SCHEME>
__device__ __forceinline__ float weno_march_f(float b1, int e1, float b2, int e2, float b3, int e3, float p1,
                                              float p2, float p3)
{
    const int E = max(max(e1, e2), e3);
    b1 = ldexpf(b1, 2 * (e1 - E));
    b2 = ldexpf(b2, 2 * (e2 - E));
    b3 = ldexpf(b3, 2 * (e3 - E));
    const float r = weno5_combine<float, SCHEME>(b1, b2, b3, ldexpf(p1, -E), ldexpf(p2, -E), ldexpf(p3, -E));
    return ldexpf(r, E);
}

// The marching kernels' shared pieces.  A triple of cells in the precision
// TF (Triple, or TripleF with its exponent in the mixed form).
template <class TF>
using MarchTriple = typename std::conditional<std::is_same<TF, double>::value, Triple, TripleF>::type;

template <class TF>
__device__ __forceinline__ MarchTriple<TF> march_triple(TF x, TF y, TF z)
{
    if constexpr (std::is_same<TF, double>::value) return weno_triple(x, y, z);
    else return weno_triple_f(x, y, z);
}

// one reconstruction from the triples' pieces b1..b3 (in FP64 already q =
// beta + eps) and the candidate values p1..p3
template <int SCHEME, class TF>
__device__ __forceinline__ TF march_recon(const MarchTriple<TF>& t1, TF b1, const MarchTriple<TF>& t2, TF b2,
                                          const MarchTriple<TF>& t3, TF b3, TF p1, TF p2, TF p3)
{
    if constexpr (std::is_same<TF, double>::value) return weno5_combine_q<double, SCHEME>(b1, b2, b3, p1, p2, p3);
    else return weno_march_f<SCHEME>(b1, t1.e, b2, t2.e, b3, t3.e, p1, p2, p3);
}

// candidate values of face k from the window w[0..5] = cells k-3..k+2:
// left-biased p1 = E (cells k-3..k-1), p2 = R (k-2..k), p3 = L (k-1..k+1);
// right-biased p1 = E' (k..k+2), p2 = L, p3 = R -- the two reconstructions
// share their second and third candidates
template <class TF>
__device__ __forceinline__ void march_cands(const TF* w, TF& E, TF& R, TF& L, TF& Er)
{
    E = (TF(1) / TF(3)) * w[0] - (TF(7) / TF(6)) * w[1] + (TF(11) / TF(6)) * w[2];
    R = -(TF(1) / TF(6)) * w[1] + (TF(5) / TF(6)) * w[2] + (TF(1) / TF(3)) * w[3];
    L = (TF(1) / TF(3)) * w[2] + (TF(5) / TF(6)) * w[3] - (TF(1) / TF(6)) * w[4];
    Er = (TF(1) / TF(3)) * w[5] - (TF(7) / TF(6)) * w[4] + (TF(11) / TF(6)) * w[3];
}

// The face flux of the window's face (between cells 2 and 3 of the window
// u, fv, vv; triples tf, tu centred on cells 1..4): both reconstructions of
// n v and of n, Lax-Friedrichs combination with the local speed
// |(v_L + v_R)/2| (Source/compute_explicit_flux.H:307-394).
template <int SCHEME, class TF>
__device__ __forceinline__ TF march_face(const TF* u, const TF* fv, const TF* vv, const MarchTriple<TF>* tf,
                                         const MarchTriple<TF>* tu)
{
    TF fE, fR, fL, fEr, uE, uR, uL, uEr;
    march_cands(fv, fE, fR, fL, fEr);
    march_cands(u, uE, uR, uL, uEr);
    const TF fph = march_recon<SCHEME, TF>(tf[0], tf[0].sp, tf[1], tf[1].sq, tf[2], tf[2].sr, fE, fR, fL);
    const TF fmh = march_recon<SCHEME, TF>(tf[3], tf[3].sr, tf[2], tf[2].sq, tf[1], tf[1].sp, fEr, fL, fR);
    const TF uph = march_recon<SCHEME, TF>(tu[0], tu[0].sp, tu[1], tu[1].sq, tu[2], tu[2].sr, uE, uR, uL);
    const TF umh = march_recon<SCHEME, TF>(tu[3], tu[3].sr, tu[2], tu[2].sq, tu[1], tu[1].sp, uEr, uL, uR);
    const TF lam = fabs(TF(0.5) * (vv[2] + vv[3]));
    return TF(0.5) * (fph + lam * uph) + TF(0.5) * (fmh - lam * umh);
}

template <int DIR, int SCHEME, class TF>
__global__ void __launch_bounds__(256) face_flux_march_kernel(FluxArgs A, int chunk)
{
    static_assert(DIR == 1 || DIR == 2, "marching along y or z");
    using Tr = MarchTriple<TF>;
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    // species fastest in the block order: the waves of the nt species of a
    // column block run close together (xcd_block keeps consecutive blocks
    // on one XCD), so E_d, read by each of them, comes from L2 after the
    // first
    const int c = (int)(bI_.x % (unsigned)A.nt);
    const int i = (int)(bI_.x / (unsigned)A.nt) * blockDim.x + threadIdx.x;
    if (i >= g.n[0]) return;
    const int tr = bI_.y;                      // DIR = 2: j; DIR = 1: k
    const int ch = bI_.z;
    const int nd = g.n[DIR], per = g.periodic[DIR];
    const int f0 = ch * chunk, f1 = min(f0 + chunk, nd + 1);
    const long long N = g.N;
    const long long st = (DIR == 1) ? (long long)g.n[0] : (long long)g.n[0] * g.n[1];
    const long long base = (DIR == 1) ? (long long)i + (long long)g.n[0] * g.n[1] * tr
                                      : (long long)i + (long long)g.n[0] * tr;
    // face (i, f, k) or (i, j, f) in the face array of direction DIR
    const long long fbase = (DIR == 1) ? (long long)i + (long long)g.n[0] * (long long)(g.n[1] + 1) * tr
                                       : (long long)i + (long long)g.n[0] * tr;
    const long long fst = (long long)g.n[0] * ((DIR == 1) ? 1 : g.n[1]);
    const int sp = A.tsp[c];
    const Real* s = A.state;
    const Real Tg = A.P.Tg;
    // window of the six stencil cells k-3..k+2 of face k (slot 0 oldest):
    // density, f = n v and drift velocity, evaluated in FP64 and rounded to
    // TF as the per-face kernel does
    TF u[6], fv[6], vv[6];
    Tr tf[4], tu[4];                           // triples centred k-2..k+1
    // the density and E_DIR of a cell, and from them (and the cell's
    // mobility inputs) its window entries
    auto raw = [&](int m, double& nn, double& ed) __attribute__((always_inline)) {
        const long long o = base + (long long)ghost_map(m, nd, per) * st;
        nn = s[(long long)sp * N + o];
        ed = s[(long long)(EFX_ID + DIR) * N + o];
    };
    auto cell = [&](int m, double nn, double ed, TF& un, TF& fn, TF& vn) __attribute__((always_inline)) {
        const long long o = base + (long long)ghost_map(m, nd, per) * st;
        const CellQ q = cell_q(s, N, o);
        const double vel = specMob(sp, q.te, q.ndens, q.emag, Tg) * ed;
        un = TF(nn);
        fn = TF(nn * vel);
        vn = TF(vel);
    };
#pragma unroll
    for (int w = 0; w < 6; ++w)
    {
        double nn, ed;
        raw(f0 - 3 + w, nn, ed);
        cell(f0 - 3 + w, nn, ed, u[w], fv[w], vv[w]);
    }
    // the cell entering at the next advance, loaded one face ahead so that
    // its latency hides behind a face's arithmetic (clamped to the last
    // cell the chunk reads: the final prefetch is not used)
    double pn, pe;
    raw(min(f0 + 3, f1 + 1), pn, pe);
#pragma unroll
    for (int w = 0; w < 4; ++w)
    {
        tf[w] = march_triple<TF>(fv[w], fv[w + 1], fv[w + 2]);
        tu[w] = march_triple<TF>(u[w], u[w + 1], u[w + 2]);
    }
    TF* F = static_cast<TF*>(A.F) + (long long)c * A.nf + fbase;
    // unrolled by the window length: the window rotates by register renaming
    // instead of moves (about 70 moves per face otherwise)
#pragma unroll 6
    for (int k = f0; k < f1; ++k)
    {
        TF fx = march_face<SCHEME, TF>(u, fv, vv, tf, tu);
        if (k == 0 && (A.bclo[c][DIR] == IHNEUBC || A.bclo[c][DIR] == AXISBC)) fx = TF(0);
        if (k == nd && (A.bchi[c][DIR] == IHNEUBC || A.bchi[c][DIR] == AXISBC)) fx = TF(0);
        F[(long long)k * fst] = fx;
        if (k + 1 == f1) break;
        // advance: cell k+3 enters, triple centred k+2 enters
#pragma unroll
        for (int w = 0; w < 5; ++w)
        {
            u[w] = u[w + 1];
            fv[w] = fv[w + 1];
            vv[w] = vv[w + 1];
        }
#pragma unroll
        for (int w = 0; w < 3; ++w)
        {
            tf[w] = tf[w + 1];
            tu[w] = tu[w + 1];
        }
        cell(k + 3, pn, pe, u[5], fv[5], vv[5]);
        raw(min(k + 4, f1 + 1), pn, pe);
        tf[3] = march_triple<TF>(fv[3], fv[4], fv[5]);
        tu[3] = march_triple<TF>(u[3], u[4], u[5]);
    }
}

// WENO5 x faces by marching, the rows staged through LDS.  One wavefront
// per (64 rows (j, k), species); each lane walks the n0 + 1 faces of its
// row with the window and triples of face_flux_march_kernel.  Cells enter in
// blocks of XB = 8 per row: the wavefront reads a block's densities and E_x
// for its 64 rows with consecutive lanes on consecutive cells of a row
// (contiguous row segments), stores them transposed in LDS (cell-major,
// rows padded to 65 entries, so the lane-per-row reads do not conflict), and
// each lane takes its own row's cells from there.  The next block is read
// into registers while the current one is marched.  The faces of a block go
// back through LDS and are written as contiguous row segments (stored
// straight from the lane-per-row layout they cost a cache line per lane and
// face, which made the kernel slower than the per-face one).  The case's
// other per-cell inputs (density sum, |E|, T_e) are read from the state
// directly, and only when its mobility uses them.
template <int SCHEME, class TF>
__global__ void __launch_bounds__(64) face_flux_march_x_kernel(FluxArgs A)
{
    using Tr = MarchTriple<TF>;
    constexpr int XB = 8;
    constexpr int PAD = 65;
    constexpr int NSLOT = XB + 6;              // [0, XB): the block; [XB, XB + 6): the first window
    __shared__ double su[NSLOT * PAD], se[NSLOT * PAD];
    __shared__ TF so[XB * PAD];
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int lane = threadIdx.x;
    const int nrow = g.n[1] * g.n[2];
    // species fastest in the block order (E_x shared through L2, as in
    // face_flux_march_kernel)
    const int c = (int)(bI_.x % (unsigned)A.nt);
    const int r0 = (int)(bI_.x / (unsigned)A.nt) * 64;
    const int sp = A.tsp[c];
    const int nd = g.n[0], per = g.periodic[0];
    const int nf = nd + 1;
    const long long N = g.N;
    const Real* s = A.state;
    const Real Tg = A.P.Tg;
    // rows past the last (the final wavefront) repeat the last row and
    // store nothing
    const long long rowc = (long long)min(r0 + lane, nrow - 1) * nd;
    // the row sequence: entry q is cell q - 3 of the row (the first window
    // is entries 0..5, the cell entering after face k is entry k + 6)
    auto cell_at = [&](int r, int q) __attribute__((always_inline)) {
        return (long long)min(r0 + r, nrow - 1) * nd + ghost_map(q - 3, nd, per);
    };
    // entries 0..5 into slots XB..XB+5
#pragma unroll
    for (int t = 0; t < 6; ++t)
    {
        const int e = lane + 64 * t;
        const int r = e / 6, cc = e - r * 6;
        const long long o = cell_at(r, cc);
        su[(XB + cc) * PAD + r] = s[(long long)sp * N + o];
        se[(XB + cc) * PAD + r] = s[(long long)EFX_ID * N + o];
    }
    // a block of entries q0 .. q0 + XB - 1: read into registers, stored to
    // slots 0 .. XB - 1
    double pu[XB], pe[XB];
    auto fetch = [&](int q0) __attribute__((always_inline)) {
#pragma unroll
        for (int t = 0; t < XB; ++t)
        {
            const int e = lane + 64 * t;
            const long long o = cell_at(e / XB, q0 + e % XB);
            pu[t] = s[(long long)sp * N + o];
            pe[t] = s[(long long)EFX_ID * N + o];
        }
    };
    auto put = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int t = 0; t < XB; ++t)
        {
            const int e = lane + 64 * t;
            su[(e % XB) * PAD + e / XB] = pu[t];
            se[(e % XB) * PAD + e / XB] = pe[t];
        }
    };
    auto cell = [&](int slot, int m, TF& un, TF& fn, TF& vn) __attribute__((always_inline)) {
        const long long o = rowc + ghost_map(m, nd, per);
        const CellQ q = cell_q(s, N, o);
        const double nn = su[slot * PAD + lane];
        const double vel = specMob(sp, q.te, q.ndens, q.emag, Tg) * se[slot * PAD + lane];
        un = TF(nn);
        fn = TF(nn * vel);
        vn = TF(vel);
    };
    fetch(6);
    put();
    __syncthreads();
    TF u[6], fv[6], vv[6];
    Tr tf[4], tu[4];
#pragma unroll
    for (int w = 0; w < 6; ++w) cell(XB + w, w - 3, u[w], fv[w], vv[w]);
#pragma unroll
    for (int w = 0; w < 4; ++w)
    {
        tf[w] = march_triple<TF>(fv[w], fv[w + 1], fv[w + 2]);
        tu[w] = march_triple<TF>(u[w], u[w + 1], u[w + 2]);
    }
    TF* F = static_cast<TF*>(A.F) + (long long)c * A.nf;
    const bool zlo = A.bclo[c][0] == IHNEUBC || A.bclo[c][0] == AXISBC;
    const bool zhi = A.bchi[c][0] == IHNEUBC || A.bchi[c][0] == AXISBC;
    for (int k0 = 0; k0 < nf; k0 += XB)
    {
        // faces k0 .. k0 + XB - 1 (uniform over the wavefront); the cell
        // entering after face k is in slot k - k0
        const bool more = k0 + XB < nf;
        if (more) fetch(k0 + XB + 6);
#pragma unroll
        for (int t = 0; t < XB; ++t)
        {
            const int k = k0 + t;
            if (k < nf)
            {
                TF fx = march_face<SCHEME, TF>(u, fv, vv, tf, tu);
                if (k == 0 && zlo) fx = TF(0);
                if (k == nd && zhi) fx = TF(0);
                so[t * PAD + lane] = fx;
#pragma unroll
                for (int w = 0; w < 5; ++w)
                {
                    u[w] = u[w + 1];
                    fv[w] = fv[w + 1];
                    vv[w] = vv[w + 1];
                }
#pragma unroll
                for (int w = 0; w < 3; ++w)
                {
                    tf[w] = tf[w + 1];
                    tu[w] = tu[w + 1];
                }
                cell(t, k + 3, u[5], fv[5], vv[5]);
                tf[3] = march_triple<TF>(fv[3], fv[4], fv[5]);
                tu[3] = march_triple<TF>(u[3], u[4], u[5]);
            }
        }
        __syncthreads();
        // the block's faces as row segments: consecutive lanes on
        // consecutive faces of a row
#pragma unroll
        for (int t = 0; t < XB; ++t)
        {
            const int e = lane + 64 * t;
            const int r = e / XB, cc = e % XB;
            if (r0 + r < nrow && k0 + cc < nf) F[(long long)(r0 + r) * nf + k0 + cc] = so[cc * PAD + r];
        }
        if (more) put();
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Per-cell explicit source and implicit-solve inputs (one pass):
//   expl_c = sum_d (F_d(lo) - F_d(hi))/dx_d + rxn_c       (compute_dsdt,
//            Source/ScalarSolve.cpp:17-79, rxn from CKWC :181-209)
//   rhs_c  = ((1/dt) n_old_c + expl_c) * timescale/scale_c  (:626-628, :826-844)
//   iD_c   = 1/specDiff_c                                   (:658-676)
// with per-component max|rhs_c| for the MLMG stopping test.
struct CellArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state;          // Sborder (phi_new: new E, old densities)
    const Real* state_old;// Sborder_old densities
    int nt;
    int tsp[vd::kMaxComp];
    Real rscale[vd::kMaxComp];   // timescale / specscale
    const void* F[3];     // faces in the flux precision TF
    long long nf[3];
    Real* rhs;            // nt * N
    Real* iD;             // nt * N
    Real* expl_out;       // optional (NUM_SPECIES+1) * N, explicit source for tests
    Real* rxn_out;        // optional (NUM_SPECIES+1) * N reaction source
    double* partials;     // nt * nblocks * kCellWaves wave maxima of |rhs|
    double* partials_id;  // 2 * nt * nblocks * kCellWaves: maxima of 1/D and of -1/D
    int defer_c;          // component whose rhs waits for the Joule heating (-1: none)
    Real* defer_out;      // its explicit source (N)
    int write_eih;        // store the energy's reaction source as EIH
    // three-term photoionization (Source/Evolve.cpp:756-799): the terms'
    // solutions (null: off), added one after another to the reaction
    // source of the electrons and of photo_sp, their sum kept as
    // PHOTO_ION_SRC_ID
    const Real* photo[3];
    int photo_sp;
    // components whose 1/D is not stored (predicted spatially uniform: the
    // solver then takes the constant-coefficient kernels and never reads
    // it; idiff_fill_kernel stores it if the prediction fails)
    unsigned skip_iD;
};

// Reaction source of one cell (Source/ScalarSolve.cpp:156-216): C = n/N_A,
// the case's CKWC, wdot*N_A and the energy exchange in wdot[NUM_SPECIES],
// then the case's add_user_react_sources on a one-cell view of the values.
__device__ __forceinline__ void cell_reactions(const PhysParams& P, const ProbParm* pp, Real* state, int i, int j,
                                               int k, long long o, Real* wdot)
{
    const Grid& g = P.g;
    const long long N = g.N;
    Real C[NUM_SPECIES], W[NUM_SPECIES];
    for (int sp = 0; sp < NUM_SPECIES; ++sp) C[sp] = state[(long long)sp * N + o] / N_A;
    Real qx = 0.0;
    CKWC(P.Tg, C, W, state[(long long)ETEMP_ID * N + o], state[(long long)REF_ID * N + o], &qx);
    for (int sp = 0; sp < NUM_SPECIES; ++sp) wdot[sp] = W[sp] * N_A;
    wdot[NUM_SPECIES] = qx;
    const int gi = i + g.off[0], gj = j + g.off[1], gk = k + g.off[2];
    amrex::Array4<Real> rv(wdot, amrex::Dim3{gi, gj, gk}, amrex::Dim3{gi + 1, gj + 1, gk + 1}, NUM_SPECIES + 1);
    user_sources::add_user_react_sources(gi, gj, gk, state_view(state, g), rv, garr(g.problo), garr(g.probhi),
                                         garr(g.dx), P.time, *pp, P.Tg, P.Pg);
}

template <class TF>
__global__ void __launch_bounds__(256) cell_source_kernel(CellArgs A)
{
    const dim3 bI_ = xcd_block();
    const TF* F0 = static_cast<const TF*>(A.F[0]);
    const TF* F1 = static_cast<const TF*>(A.F[1]);
    const TF* F2 = static_cast<const TF*>(A.F[2]);
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    const long long o = in ? g.idx(i, j, k) : 0;
    Real wdot[NUM_SPECIES + 1];
    CellQ q{0, 0, 0};
    if (in)
    {
        for (int sp = 0; sp <= NUM_SPECIES; ++sp) wdot[sp] = 0.0;
        if (A.P.do_reactions)
        {
            cell_reactions(A.P, A.pp, A.state, i, j, k, o, wdot);
        }
        if (A.photo[0])
        {
            Real tot = 0.0;
#pragma unroll
            for (int t = 0; t < 3; ++t)
            {
                const Real sv = A.photo[t][o];
                tot += sv;
                wdot[E_ID] += sv;
                wdot[A.photo_sp] += sv;
            }
            A.state[(long long)PHOTO_ION_SRC_ID * N + o] = tot;
        }
        if (A.rxn_out)
            for (int sp = 0; sp <= NUM_SPECIES; ++sp) A.rxn_out[(long long)sp * N + o] = wdot[sp];
        // inelastic heating diagnostic, Source/ElecEnergySource.cpp:287-288
        if (A.write_eih) A.state[(long long)EIH_ID * N + o] = wdot[NUM_SPECIES];
        q = cell_q(A.state, N, o);
    }
    const long long sx = 1;
    (void)sx;
    for (int c = 0; c < A.nt; ++c)
    {
        double m = 0.0, idv = 0.0;
        if (in)
        {
            const int sp = A.tsp[c];
            Real v;
            {
                const long long e0 = g.n[0] + 1;
                const long long f = (long long)i + e0 * ((long long)j + (long long)g.n[1] * k);
                v = (Real(F0[(long long)c * A.nf[0] + f]) - Real(F0[(long long)c * A.nf[0] + f + 1])) / g.dx[0] +
                    wdot[sp];
            }
            if (SD > 1)
            {
                const long long e0 = g.n[0], e1 = g.n[1] + 1;
                const long long f = (long long)i + e0 * ((long long)j + e1 * k);
                v += (Real(F1[(long long)c * A.nf[1] + f]) - Real(F1[(long long)c * A.nf[1] + f + e0])) / g.dx[1];
            }
            if (SD > 2)
            {
                const long long e0 = g.n[0], e1 = g.n[1];
                const long long f = (long long)i + e0 * ((long long)j + e1 * k);
                v += (Real(F2[(long long)c * A.nf[2] + f]) - Real(F2[(long long)c * A.nf[2] + f + e0 * e1])) / g.dx[2];
            }
            if (g.coord == 1) v -= axisym_drift(g, A.state, N, o, i, sp, q, A.P.Tg);
            if (A.expl_out) A.expl_out[(long long)sp * N + o] = v;
            if (c == A.defer_c)
                A.defer_out[o] = v;
            else
            {
                const Real r = ((1.0 / A.P.dt) * A.state_old[(long long)sp * N + o] + 1.0 * v) * A.rscale[c];
                A.rhs[(long long)c * N + o] = r;
                m = fabs(r);
            }
            idv = 1.0 / specDiff(sp, q.te, q.ndens, q.emag, A.P.Tg);
            if (!((A.skip_iD >> c) & 1u)) A.iD[(long long)c * N + o] = idv;
        }
        // one partial per wave (kCellWaves per workgroup of 256)
        const long long nw = (long long)gridDim.x * gridDim.y * gridDim.z * kCellWaves;
        wave_max_store(m, A.partials + (long long)c * nw);
        // uniformity test of the coefficient: max(1/D) and max(-1/D)
        wave_max_store(in ? idv : -1e308, A.partials_id + (long long)c * nw);
        wave_max_store(in ? -idv : -1e308, A.partials_id + (long long)(A.nt + c) * nw);
    }
}

// 1/D of the components of mask, as cell_source_kernel forms it from the
// same state (for those it did not store).
__global__ void __launch_bounds__(256) idiff_fill_kernel(CellArgs A, unsigned mask)
{
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    const long long N = g.N;
    const long long o = g.idx(i, j, k);
    const CellQ q = cell_q(A.state, N, o);
    for (int c = 0; c < A.nt; ++c)
        if ((mask >> c) & 1u) A.iD[(long long)c * N + o] = 1.0 / specDiff(A.tsp[c], q.te, q.ndens, q.emag, A.P.Tg);
}

// ---------------------------------------------------------------------------
// update_rxnsrc_at_all_levels on a caller's state (Source/ScalarSolve.cpp:
// 156-216): the NUM_SPECIES + 1 reaction components of every cell, zero
// without reactions.
struct RxnArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state;
    Real* rxn;            // (NUM_SPECIES + 1) * N
};

__global__ void __launch_bounds__(256) reaction_source_kernel(RxnArgs A)
{
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    const long long N = g.N;
    const long long o = g.idx(i, j, k);
    Real wdot[NUM_SPECIES + 1];
    for (int sp = 0; sp <= NUM_SPECIES; ++sp) wdot[sp] = 0.0;
    if (A.P.do_reactions) cell_reactions(A.P, A.pp, A.state, i, j, k, o, wdot);
    for (int sp = 0; sp <= NUM_SPECIES; ++sp) A.rxn[(long long)sp * N + o] = wdot[sp];
}

// compute_dsdt (Source/ScalarSolve.cpp:17-79) for species [start, start +
// num): expl_c = sum_d (F_d(i) - F_d(i+1))/dx_d + rxn_c, in the reference's
// order (x difference plus rxn, then y, then z); the other NUM_SPECIES + 1
// components of expl are zero.  rxn == nullptr evaluates the reactions in
// this pass.
struct DsdtArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state;
    const Real* rxn;      // (NUM_SPECIES + 1) * N, or nullptr
    int start, num;
    const Real* F[3];     // num * nf[d] faces per direction
    long long nf[3];
    Real* expl;           // (NUM_SPECIES + 1) * N
};

__global__ void __launch_bounds__(256) dsdt_kernel(DsdtArgs A)
{
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    const long long N = g.N;
    const long long o = g.idx(i, j, k);
    Real wdot[NUM_SPECIES + 1];
    for (int sp = 0; sp <= NUM_SPECIES; ++sp) wdot[sp] = 0.0;
    if (A.rxn)
        for (int sp = A.start; sp < A.start + A.num; ++sp) wdot[sp] = A.rxn[(long long)sp * N + o];
    else if (A.P.do_reactions)
        cell_reactions(A.P, A.pp, A.state, i, j, k, o, wdot);
    for (int sp = 0; sp <= NUM_SPECIES; ++sp)
    {
        Real v = 0.0;
        if (sp >= A.start && sp < A.start + A.num)
        {
            const long long c = sp - A.start;
            {
                const long long e0 = g.n[0] + 1;
                const long long f = (long long)i + e0 * ((long long)j + (long long)g.n[1] * k);
                v = (A.F[0][c * A.nf[0] + f] - A.F[0][c * A.nf[0] + f + 1]) / g.dx[0] + wdot[sp];
            }
            if (SD > 1)
            {
                const long long e0 = g.n[0], e1 = g.n[1] + 1;
                const long long f = (long long)i + e0 * ((long long)j + e1 * k);
                v += (A.F[1][c * A.nf[1] + f] - A.F[1][c * A.nf[1] + f + e0]) / g.dx[1];
            }
            if (SD > 2)
            {
                const long long e0 = g.n[0], e1 = g.n[1];
                const long long f = (long long)i + e0 * ((long long)j + e1 * k);
                v += (A.F[2][c * A.nf[2] + f] - A.F[2][c * A.nf[2] + f + e0 * e1]) / g.dx[2];
            }
            if (g.coord == 1) v -= axisym_drift(g, A.state, N, o, i, sp, cell_q(A.state, N, o), A.P.Tg);
        }
        A.expl[(long long)sp * N + o] = v;
    }
}

// ---------------------------------------------------------------------------
// Boundary data of one domain side (direction d, side s) for nc components:
// runs the reference's boundary hooks on every boundary face
// (Source/PotentialSolve.cpp:350-434, Source/ScalarSolve.cpp:693-803) into
// slab arrays, then folds the inhomogeneous ghost constant into the rhs of
// the boundary cell:  rhs -= -beta*b_face*gconst/h^2.
struct SideArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state;
    int d, side;          // direction, 0 = lo, 1 = hi
    int n1, n2;           // tangential extents
    int nc;
    int tsp[vd::kMaxComp];
    int kind[vd::kMaxComp];   // BcKind per component
    int potential;        // 1: potential_bc, 2: photoionization_bc, 0: species_bc
    Real app_voltage;
    Real vscale;          // potential: 1/potscale; species: 1/specscale per comp in cscale
    Real cscale[vd::kMaxComp];
    int maxorder;
    Real beta;
    Real bconst;          // potential: face b constant
    const Real* iD;       // species: level-0 inverse coefficients (nc*N)
    unsigned skip_iD;     // components whose iD the cell pass did not store
    Real* slab_g;         // nc * n1*n2 : bc value as written by the hook
    Real* slab_a;
    Real* slab_b;
    Real* slab_f;
    Real* gconst;         // nc * n1*n2 : constant part of the inhomogeneous ghost
    Real* rhs;            // nc * N
    unsigned nofold;      // components whose rhs is folded later (the energy)
    // integrated conduction currents, areas and surfaces handed to the
    // hooks (Source/PotentialSolve.cpp:42-52, Source/ScalarSolve.cpp:404-414)
    amrex::GpuArray<Real, MAX_CURRENT_LOCS> ic, ia;
    amrex::GpuArray<int, MAX_CURRENT_LOCS> is;
};

// The lo and hi sides of one direction run in one launch (blockIdx.z = side
// slot): they never share a boundary cell, so their rhs folds cannot
// collide, while sides of different directions share edge cells and stay in
// separate launches (their folds keep the reference's order).
struct SidePair
{
    SideArgs s[2];
};

__global__ void __launch_bounds__(256) side_bc_kernel(const SidePair SP)
{
    const SideArgs& A = SP.s[blockIdx.z];
    const Grid& g = A.P.g;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (a >= A.n1 || b >= A.n2) return;
    const int d = A.d;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int F[3] = {0, 0, 0};
    F[t1] = a;
    F[t2] = b;
    F[d] = A.side ? g.n[d] : 0;
    int G[3] = {F[0], F[1], F[2]};
    G[d] = A.side ? g.n[d] : -1;
    int C[3] = {F[0], F[1], F[2]};
    C[d] = A.side ? g.n[d] - 1 : 0;
    const long long so = (long long)a + (long long)A.n1 * b;
    const long long slab = (long long)A.n1 * A.n2;
    // slab views: the ghost layer of the grown box restricted to this side
    amrex::Dim3 lo{0, 0, 0}, hi{g.n[0], g.n[1], g.n[2]};
    if (d == 0) { lo.x = G[0]; hi.x = G[0] + 1; }
    if (d == 1) { lo.y = G[1]; hi.y = G[1] + 1; }
    if (d == 2) { lo.z = G[2]; hi.z = G[2] + 1; }
    if (SD < 3) { lo.z = 0; hi.z = 1; }
    amrex::Array4<Real> vg(A.slab_g, lo, hi, A.nc), va(A.slab_a, lo, hi, A.nc), vb(A.slab_b, lo, hi, A.nc),
        vf(A.slab_f, lo, hi, A.nc);
    const amrex::Array4<Real> sv = state_view(A.state, g);
    const int sgn = A.side ? 1 : -1;
    const amrex::GpuArray<Real, MAX_CURRENT_LOCS> ic = A.ic, ia = A.ia;
    const amrex::GpuArray<int, MAX_CURRENT_LOCS> is = A.is;
    for (int c = 0; c < A.nc; ++c)
    {
        // defaults the reference sets before the hooks run (robin 0/1/0)
        vg.p[(long long)c * slab + so] = 0.0;
        va.p[(long long)c * slab + so] = 0.0;
        vb.p[(long long)c * slab + so] = 1.0;
        vf.p[(long long)c * slab + so] = 0.0;
    }
    if (A.potential == 2)
    {
        // photoionization (Source/PhotoIonizationSolve.cpp:213-359): the bc
        // array is PHOTO_ION_SRC_ID copied without ghosts, so 0 at the ghost;
        // robin 0/1/0 before the hook
        if (A.P.user_photo)
            user_transport::photoionization_bc(F[0], F[1], F[2], d, sgn, sv, vg, va, vb, vf, garr(g.problo),
                                               garr(g.probhi), garr(g.dx), A.P.time, *A.pp, A.P.Tg, A.P.Pg);
        else
        {
            // plasmachem_transport::photoionization_bc, Source/PlasmaChem.H:202-255
            va(G[0], G[1], G[2]) = 1.0;
            vb(G[0], G[1], G[2]) = 0.0;
            vf(G[0], G[1], G[2]) = 0.0;
            vg(G[0], G[1], G[2]) = 0.0;
        }
    } else if (A.potential == 1)
    {
        // the bc value defaults to the state's potential at the ghost
        // (Sborder[POT] copied into the bc array, PotentialSolve.cpp:219)
        vg.p[so] = sv(G[0], G[1], G[2], POT_ID);
        if (A.P.user_potential)
            user_transport::potential_bc(F[0], F[1], F[2], d, sgn, sv, vg, va, vb, vf, garr(g.problo),
                                         garr(g.probhi), garr(g.dx), A.P.time, *A.pp, A.P.Tg, A.P.Pg,
                                         A.app_voltage, ic, ia, is);
        else
        {
            // plasmachem_transport::potential_bc, Source/PlasmaChem.H:16-73
            va(G[0], G[1], G[2]) = 1.0;
            vb(G[0], G[1], G[2]) = 0.0;
            vf(G[0], G[1], G[2]) = A.app_voltage;
            vg(G[0], G[1], G[2]) = A.app_voltage;
        }
    } else
    {
        // specdata ghost defaults to the Sborder species value
        for (int c = 0; c < A.nc; ++c) vg.p[(long long)c * slab + so] = sv(G[0], G[1], G[2], A.tsp[c]);
        for (int c = 0; c < A.nc; ++c)
        {
            const int sp = A.tsp[c];
            if (A.P.user_species)
                user_transport::species_bc(F[0], F[1], F[2], d, sgn, sp, c, sv, vg, va, vb, vf, garr(g.problo),
                                           garr(g.probhi), garr(g.dx), A.P.time, *A.pp, A.P.Tg, A.P.Pg, ic, ia, is);
            else
            {
                // plasmachem_transport::species_bc, Source/PlasmaChem.H:75-200
                Real nd = 0.0;
                for (int s = 0; s < NUM_SPECIES; s++) nd += sv(C[0], C[1], C[2], s);
                const Real ex = sv(C[0], C[1], C[2], EFX_ID), ey = sv(C[0], C[1], C[2], EFY_ID),
                           ez = sv(C[0], C[1], C[2], EFZ_ID);
                const Real emag = sqrt(ex * ex + ey * ey + ez * ez);
                const Real Te = sv(C[0], C[1], C[2], ETEMP_ID);
                Real val;
                if (sp == E_ID)
                {
                    const Real ne = sv(C[0], C[1], C[2], E_ID);
                    const Real Dc = specDiff(sp, Te, nd, emag, A.P.Tg);
                    const Real cbar = sqrt(8.0 * K_B * Te / PI / ME);
                    const Real flx = ne * cbar / 4.0 * Real(sgn);
                    val = -flx / Dc;
                } else if (sp == EEN_ID)
                {
                    const Real ne = sv(C[0], C[1], C[2], E_ID);
                    const Real Dc = specDiff(sp, Te, nd, emag, A.P.Tg);
                    const Real cbar = sqrt(8.0 * K_B * Te / PI / ME);
                    const Real flx = (2.0 * K_B * Te) * ne * cbar / 4.0 * Real(sgn);
                    val = -flx / Dc;
                } else if (species_charge(sp) != 0)
                {
                    const Real mu = specMob(sp, Te, nd, emag, A.P.Tg);
                    const Real Dc = specDiff(sp, Te, nd, emag, A.P.Tg);
                    const Real fl = mu * sv(C[0], C[1], C[2], sp) * sv(C[0], C[1], C[2], EFX_ID + d);
                    val = (fl * Real(sgn) >= 0.0) ? (-fl / Dc) : 0.0;
                } else
                {
                    const Real Dc = specDiff(sp, Te, nd, emag, A.P.Tg);
                    const Real cbar = sqrt(8.0 * K_B * A.P.Tg / PI / (mw(sp) * M_AMU));
                    // the state read at the face index (a ghost on the hi side)
                    const Real flx = sv(F[0], F[1], F[2], sp) * cbar / 4.0 * Real(sgn);
                    val = -flx / Dc;
                }
                vg.p[(long long)c * slab + so] = val;
            }
        }
    }
    // scale and fold
    const Real h = g.dx[d];
    const long long cell = g.idx(C[0], C[1], C[2]);
    for (int c = 0; c < A.nc; ++c)
    {
        const Real sc = A.potential ? A.vscale : A.cscale[c];
        const long long q = (long long)c * slab + so;
        const Real gv = vg.p[q] * sc;
        const Real fv = vf.p[q] * sc;
        vg.p[q] = gv;
        vf.p[q] = fv;
        const int kd = A.kind[c];
        Real gc = 0.0;
        if (kd == BK_DIR)
        {
            const int m = min(g.n[d] + 1, A.maxorder);
            gc = (m >= 3) ? (8.0 / 3.0) * gv : 2.0 * gv;
        } else if (kd == BK_INHNEU)
        {
            gc = A.side ? gv * h : -gv * h;
        } else if (kd == BK_ROBIN)
        {
            gc = fv / (0.5 * va.p[q] + vb.p[q] / h);
        }
        A.gconst[q] = gc;
        if ((kd == BK_DIR || kd == BK_INHNEU || kd == BK_ROBIN) && !((A.nofold >> c) & 1u))
        {
            Real idv = 0.0;
            if (!A.potential)
            {
                if ((A.skip_iD >> c) & 1u)
                {
                    // not stored by the cell pass: its value, from the same state
                    const CellQ q = cell_q(A.state, g.N, cell);
                    idv = 1.0 / specDiff(A.tsp[c], q.te, q.ndens, q.emag, A.P.Tg);
                } else
                    idv = A.iD[(long long)c * g.N + cell];
            }
            Real bface = A.potential ? A.bconst : 1.0 / idv;
            if (d == 0 && g.coord == 1)
            {
                double rm, rp;
                rz_factors(g.problo[0], g.dx[0], C[0], rm, rp);
                bface *= A.side ? rp : rm;
            }
            A.rhs[(long long)c * g.N + cell] -= -A.beta * bface * gc / (h * h);
        }
    }
}

// ---------------------------------------------------------------------------
// Coefficient and rhs of photoionization term sph
// (Source/PhotoIonizationSolve.cpp:237-296): the case's
// get_photoion_acoeff (after a = 0) and get_photoion_rhs (after rhs = 0) on
// the Sborder state, one cell at a time; wave maxima of |rhs|, a and -a.
struct PhotoArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state;
    int sph;
    Real* rhs;            // N
    double* partials;     // 3 * nblocks * kCellWaves
};

__global__ void __launch_bounds__(256) photo_coef_kernel(PhotoArgs A)
{
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    double mr = 0.0, ma = -1e308, mn = -1e308;
    if (in)
    {
        const long long o = g.idx(i, j, k);
        const int gi = i + g.off[0], gj = j + g.off[1], gk = k + g.off[2];
        Real av = 0.0, rv = 0.0;
        const amrex::Dim3 lo{gi, gj, gk}, hi{gi + 1, gj + 1, gk + 1};
        amrex::Array4<Real> aa(&av, lo, hi, 1), ra(&rv, lo, hi, 1);
        const amrex::Array4<Real> sv = state_view(A.state, g);
        user_transport::get_photoion_acoeff(gi, gj, gk, A.sph, sv, aa, garr(g.problo), garr(g.probhi), garr(g.dx),
                                            A.P.time, *A.pp, A.P.Tg, A.P.Pg);
        user_transport::get_photoion_rhs(gi, gj, gk, A.sph, sv, ra, garr(g.problo), garr(g.probhi), garr(g.dx),
                                         A.P.time, *A.pp, A.P.Tg, A.P.Pg);
        A.rhs[o] = rv;
        mr = fabs(rv);
        ma = av;
        mn = -av;
    }
    const long long nw = (long long)gridDim.x * gridDim.y * gridDim.z * kCellWaves;
    wave_max_store(mr, A.partials);
    wave_max_store(ma, A.partials + nw);
    wave_max_store(mn, A.partials + 2 * nw);
}

// ---------------------------------------------------------------------------
// Poisson rhs: -(e/eps0) sum_k z_k n_k (+ user sources), / potscale,
// Source/PotentialSolve.cpp:250-277.
struct PoissonRhsArgs
{
    PhysParams P;
    const ProbParm* pp;
    Real* state_old;
    Real inv_potscale;
    int do_spacechrg;
    Real* rhs;
    double* partials;
};

__global__ void __launch_bounds__(256) poisson_rhs_kernel(PoissonRhsArgs A)
{
    const dim3 bI_ = xcd_block();
    __shared__ double scratch[16];
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    double m = 0.0;
    if (in)
    {
        const long long o = g.idx(i, j, k);
        Real r = 0.0;
        if (A.do_spacechrg)
        {
#pragma unroll
            for (int sp = 0; sp < NUM_SPECIES; ++sp)
            {
                const int z = species_charge(sp);
                if (z != 0) r += z * A.state_old[(long long)sp * g.N + o];
            }
            r *= (-ECHARGE / EPS0);
            const int gi = i + g.off[0], gj = j + g.off[1], gk = k + g.off[2];
            amrex::Array4<Real> rv(&r, amrex::Dim3{gi, gj, gk}, amrex::Dim3{gi + 1, gj + 1, gk + 1}, 1);
            user_sources::add_user_potential_sources(gi, gj, gk, state_view(A.state_old, g), rv, garr(g.problo),
                                                     garr(g.probhi), garr(g.dx), A.P.time, *A.pp);
        }
        r *= A.inv_potscale;
        A.rhs[o] = r;
        m = fabs(r);
    }
    block_max_store(m, A.partials + flat_block_id(), scratch);
}

// Cell E from face gradients of the potential (getGradSolution +
// average_face_to_cellcenter), E/N in Td and the potential itself,
// Source/PotentialSolve.cpp:500-542.
struct EFieldArgs
{
    PhysParams P;
    const Real* sol;
    Real* state;
    Real potscale;
    Real dep0[3][2];
    Real dep1[3][2];            // ghost dependence on the second cell (coarse/fine sides)
    const Real* gconst[3][2];   // per side slabs (1 comp)
    int eidx;
};

__global__ void __launch_bounds__(256) efield_kernel(EFieldArgs A)
{
    const dim3 bI_ = xcd_block();
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    const long long N = g.N;
    const long long o = g.idx(i, j, k);
    const int I[3] = {i, j, k};
    const Real xc = A.sol[o];
    Real Ecell[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < SD; ++d)
    {
        const long long st = (d == 0) ? 1 : (d == 1 ? (long long)g.n[0] : (long long)g.n[0] * g.n[1]);
        const int n = g.n[d];
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        const long long so = (long long)I[t1] + (long long)((SD > t1) ? g.n[t1] : 1) * (t2 < SD ? I[t2] : 0);
        Real xm, xp;
        // inhomogeneous ghost = dep0*x0 + gconst (Dirichlet order 2: -1, 2g;
        // inhomogeneous Neumann: 1, -+g h; Robin: model of the solver; Neumann 1, 0)
        if (I[d] > 0) xm = A.sol[o - st];
        else if (g.periodic[d]) xm = A.sol[o + (long long)(n - 1) * st];
        else xm = A.dep0[d][0] * xc + A.gconst[d][0][so] + A.dep1[d][0] * ((n > 1) ? A.sol[o + st] : xc);
        if (I[d] < n - 1) xp = A.sol[o + st];
        else if (g.periodic[d]) xp = A.sol[o - (long long)(n - 1) * st];
        else xp = A.dep0[d][1] * xc + A.gconst[d][1][so] + A.dep1[d][1] * ((n > 1) ? A.sol[o - st] : xc);
        Real gl = (xc - xm) / g.dx[d];
        gl *= -1.0 * A.potscale;
        Real gh = (xp - xc) / g.dx[d];
        gh *= -1.0 * A.potscale;
        Ecell[d] = 0.5 * (gl + gh);
        A.state[(long long)(EFX_ID + d) * N + o] = Ecell[d];
    }
    Real nd = 0.0;
#pragma unroll
    for (int sp = 0; sp < NUM_SPECIES; ++sp) nd += A.state[(long long)sp * N + o];
    nd = nd - A.state[(long long)A.eidx * N + o];
    Real Esum = 0.0;
#pragma unroll
    for (int d = 0; d < SD; ++d) Esum += Ecell[d] * Ecell[d];
    A.state[(long long)REF_ID * N + o] = (sqrt(Esum) / nd) / 1.0e-21;
    A.state[(long long)POT_ID * N + o] = xc * A.potscale;
}

// ---------------------------------------------------------------------------
// Electron time scales: max |mu_e||E|, max D_e, max |mu_e| n_e,
// Source/TimeStep.cpp:77-110.
struct TimeScaleArgs
{
    Grid g;
    const Real* state;
    Real Tg;
    int eidx;
    double* partials;          // 3 * nblocks
};

__global__ void __launch_bounds__(256) time_scales_kernel(TimeScaleArgs A, int kz)
{
    const dim3 bI_ = xcd_block();
    __shared__ double scratch[16];
    const Grid& g = A.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const bool in = (i < g.n[0] && j < g.n[1]);
    double mv = 0.0, md = 0.0, mm = 0.0;
    // kz planes per thread: the three workgroup reductions paid once per kz
    // cells of a column
    const int k0 = (int)bI_.z * kz, k1 = min(k0 + kz, g.n[2]);
    for (int k = k0; in && k < k1; ++k)
    {
        const long long N = g.N, o = g.idx(i, j, k);
        const Real te = A.state[(long long)ETEMP_ID * N + o];
        Real Esum = 0.0;
#pragma unroll
        for (int d = 0; d < SD; ++d)
        {
            const Real e = A.state[(long long)(EFX_ID + d) * N + o];
            Esum += e * e;
        }
        const Real emag = sqrt(Esum);
        Real nd = 0.0;
#pragma unroll
        for (int sp = 0; sp < NUM_SPECIES; ++sp) nd += A.state[(long long)sp * N + o];
        const Real dc = specDiff(A.eidx, te, nd, emag, A.Tg);
        const Real mu = specMob(A.eidx, te, nd, emag, A.Tg);
        mv = vmax(mv, fabs(mu) * emag);
        md = vmax(md, fabs(dc));
        mm = vmax(mm, fabs(fabs(mu) * A.state[(long long)A.eidx * N + o]));
    }
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z, bid = flat_block_id();
    block_max_store(mv, A.partials + bid, scratch);
    block_max_store(md, A.partials + nb + bid, scratch);
    block_max_store(mm, A.partials + 2 * nb + bid, scratch);
}

// planes per thread of time_scales_kernel
inline int tscale_kz(int n2) { return n2 >= 64 ? 8 : 1; }

// ---------------------------------------------------------------------------
// Electrode current diagnostics (Source/CurrentDens.cpp, no IB: every
// interface is regular and the level mask is 1).

// One-sided face gradient at a domain face with outward normal sgn,
// Source/HelperFuncs.H:13-52 (cells inside the domain; needs >= 3 cells).
__device__ inline Real onesided_grad(const amrex::Array4<Real>& a, const int F[3], int sgn, int idim, int comp,
                                     Real h)
{
    int c0[3] = {F[0], F[1], F[2]};
    c0[idim] += (sgn == 1) ? -1 : 0;
    int c1[3] = {c0[0], c0[1], c0[2]}, c2[3] = {c0[0], c0[1], c0[2]};
    c1[idim] = c0[idim] - sgn;
    c2[idim] = c1[idim] - sgn;
    Real g1, g2;
    if (sgn == -1)
    {
        g1 = (a(c1[0], c1[1], c1[2], comp) - a(c0[0], c0[1], c0[2], comp)) / h;
        g2 = (a(c2[0], c2[1], c2[2], comp) - a(c1[0], c1[1], c1[2], comp)) / h;
    } else
    {
        g1 = (a(c0[0], c0[1], c0[2], comp) - a(c1[0], c1[1], c1[2], comp)) / h;
        g2 = (a(c1[0], c1[1], c1[2], comp) - a(c2[0], c2[1], c2[2], comp)) / h;
    }
    return 2.0 * g1 - g2;
}

// Electron and ion current densities at the face of direction idim with
// nodal index F (compute_current_density_at_level,
// Source/CurrentDens.cpp:171-338): face averages of the two cells (ghost
// cells by the state's ghost policy), central gradients inside and one-sided
// gradients on domain faces.
__device__ inline void face_current(const amrex::Array4<Real>& sv, const Grid& g, int idim, const int F[3], Real Tg,
                                    int eidx, Real& ej, Real& ij)
{
    int L[3] = {F[0], F[1], F[2]};
    L[idim] -= 1;
    const int fi = F[idim];
    const bool lphys = (fi == 0), rphys = (fi == g.n[idim]);
    const Real etemp = (sv(L[0], L[1], L[2], ETEMP_ID) + sv(F[0], F[1], F[2], ETEMP_ID)) / 2;
    Real nsp[NUM_SPECIES], ndens = 0.0;
#pragma unroll
    for (int sp = 0; sp < NUM_SPECIES; ++sp)
    {
        nsp[sp] = (sv(L[0], L[1], L[2], sp) + sv(F[0], F[1], F[2], sp)) / 2;
        ndens += nsp[sp];
    }
    Real Esum = 0.0;
#pragma unroll
    for (int d = 0; d < SD; ++d)
    {
        const Real ef = (sv(L[0], L[1], L[2], EFX_ID + d) + sv(F[0], F[1], F[2], EFX_ID + d)) / 2;
        Esum += ef * ef;
    }
    const Real emag = sqrt(Esum);
    const Real h = g.dx[idim];
    Real fe;
    Real grad[NUM_SPECIES];
    if (!lphys && !rphys)
    {
        fe = -1.0 * (sv(F[0], F[1], F[2], POT_ID) - sv(L[0], L[1], L[2], POT_ID)) / h;
#pragma unroll
        for (int sp = 0; sp < NUM_SPECIES; ++sp) grad[sp] = (sv(F[0], F[1], F[2], sp) - sv(L[0], L[1], L[2], sp)) / h;
    } else
    {
        const int sg = lphys ? -1 : 1;
        fe = -1.0 * onesided_grad(sv, F, sg, idim, POT_ID, h);
#pragma unroll
        for (int sp = 0; sp < NUM_SPECIES; ++sp) grad[sp] = onesided_grad(sv, F, sg, idim, sp, h);
    }
    Real icur = 0.0;
#pragma unroll
    for (int sp = 0; sp < NUM_SPECIES; ++sp)
    {
        const Real chrg = species_charge(sp);
        if (fabs(chrg) > 0 && sp != eidx)
        {
            const Real mu = specMob(sp, etemp, ndens, emag, Tg);
            const Real dc = specDiff(sp, etemp, ndens, emag, Tg);
            icur += chrg * ECHARGE * (mu * nsp[sp] * fe - dc * grad[sp]);
        }
    }
    const Real mue = specMob(eidx, etemp, ndens, emag, Tg);
    const Real dce = specDiff(eidx, etemp, ndens, emag, Tg);
    ij = icur;
    ej = -ECHARGE * (mue * nsp[eidx] * fe - dce * grad[eidx]);
}

// Cell currents ECUR/ICUR (average of the two faces per direction,
// compute_current_den, Source/CurrentDens.cpp:66-127) and the displacement
// current DCUR = eps0 (E - E_old)/dt (compute_disp_current_den, :17-63).
struct CurrentArgs
{
    Grid g;
    Real* state;
    const Real* eold;    // SD * N: E at the start of the step
    Real Tg, time, dt;
    int eidx;
};

__global__ void __launch_bounds__(256) current_den_kernel(CurrentArgs A)
{
    const Grid& g = A.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    const long long N = g.N, o = g.idx(i, j, k);
    const amrex::Array4<Real> sv = state_view(A.state, g);
    const bool live = (A.state[(long long)CMASK_ID * N + o] == 1.0 && A.time > 0.0);
#pragma unroll
    for (int d = 0; d < SD; ++d)
    {
        int F[3] = {i, j, k};
        Real el, il, eh, ih;
        face_current(sv, g, d, F, A.Tg, A.eidx, el, il);
        F[d] += 1;
        face_current(sv, g, d, F, A.Tg, A.eidx, eh, ih);
        A.state[(long long)(ECURX_ID + d) * N + o] = 0.5 * (el + eh);
        A.state[(long long)(ICURX_ID + d) * N + o] = 0.5 * (il + ih);
        A.state[(long long)(DCURX_ID + d) * N + o] =
            live ? EPS0 * (A.state[(long long)(EFX_ID + d) * N + o] - A.eold[(long long)d * N + o]) / A.dt : 0.0;
    }
}

// Per-workgroup partial sums of the conduction current, displacement
// current and collector area of every current location
// (compute_integrated_currents, Source/CurrentDens.cpp:343-456).
struct IntCurrentArgs
{
    Grid g;
    const Real* state;
    int nloc;
    int surf[MAX_CURRENT_LOCS];
    double* partials;    // 3 * nloc * nblocks
};

__global__ void __launch_bounds__(256) integrated_currents_kernel(IntCurrentArgs A)
{
    __shared__ double scratch[16];
    const Grid& g = A.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z, bid = flat_block_id();
    Real vol = 1.0;
    for (int d = 0; d < SD; ++d) vol *= g.dx[d];
    amrex::GpuArray<int, SD> dlo, dhi;
    for (int d = 0; d < SD; ++d)
    {
        dlo[d] = 0;
        dhi[d] = g.n[d] - 1;
    }
    for (int loc = 0; loc < A.nloc; ++loc)
    {
        const int surf = A.surf[loc];
        const Real normal = (surf % 2 == 0) ? -1.0 : 1.0;
        const int dir = surf / 2;
        const Real area = vol / g.dx[dir];
        Real c = 0.0, dsp = 0.0, ar = 0.0;
        if (in)
        {
            const long long o = g.idx(i, j, k);
            amrex::IntVect cid(AMREX_D_DECL(i, j, k));
            const Real cv = user_transport::current_collector_value(cid, loc, surf, garr(g.problo), garr(g.probhi),
                                                                    dlo, dhi, garr(g.dx));
            c = area * normal * (A.state[(long long)(ECURX_ID + dir) * N + o] + A.state[(long long)(ICURX_ID + dir) * N + o]) *
                1 * cv;
            dsp = area * normal * A.state[(long long)(DCURX_ID + dir) * N + o] * 1 * cv;
            ar = area * 1 * cv;
        }
        block_sum_store(c, A.partials + (long long)(3 * loc + 0) * nb + bid, scratch);
        block_sum_store(dsp, A.partials + (long long)(3 * loc + 1) * nb + bid, scratch);
        block_sum_store(ar, A.partials + (long long)(3 * loc + 2) * nb + bid, scratch);
    }
}

// ---------------------------------------------------------------------------
// Solution post-processing of the batched species solve: unscale, clamp
// (Source/ScalarSolve.cpp:894-946).
struct PostArgs
{
    Grid g;
    Real* state;
    int nt;
    int tsp[vd::kMaxComp];
    Real scale[vd::kMaxComp];
    Real floor[vd::kMaxComp];
    int clamp;
    unsigned active;
};

__global__ void __launch_bounds__(256) scale_comps_kernel(PostArgs A)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.g.N) return;
    for (int c = 0; c < A.nt; ++c)
    {
        if (!((A.active >> c) & 1u)) continue;
        Real& v = A.state[(long long)A.tsp[c] * A.g.N + q];
        Real x = v * A.scale[c];
        if (A.clamp && x < A.floor[c]) x = A.floor[c];
        v = x;
    }
}

// LFA electron temperature (Source/ElecEnergySource.cpp:43-69), split in
// two: Te from the pre-solve state, then the energy density from the
// post-solve electron density.
struct LfaArgs
{
    Grid g;
    Real* state;
    Real Tg;
    Real minetemp;
    int eidx;
    const Real* spec;     // Sborder species densities (stride N)
};

__global__ void __launch_bounds__(256) lfa_te_kernel(LfaArgs A)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.g.N) return;
    const long long N = A.g.N;
    if (!int(A.state[(long long)CMASK_ID * N + q])) return;
    Real sp[NUM_SPECIES];
    for (int s = 0; s < NUM_SPECIES; ++s) sp[s] = A.spec[(long long)s * N + q];
    Real te = etempLFA(A.Tg, sp, A.state[(long long)REF_ID * N + q]);
    if (te < A.minetemp) te = A.minetemp;
    A.state[(long long)ETEMP_ID * N + q] = te;
}

__global__ void __launch_bounds__(256) lfa_een_kernel(LfaArgs A, const Real* cmask_src)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.g.N) return;
    const long long N = A.g.N;
    if (!int(cmask_src[q])) return;
    A.state[(long long)EEN_ID * N + q] = 1.5 * K_B * A.state[(long long)A.eidx * N + q] * A.minetemp;
}

// ---------------------------------------------------------------------------
// Electron Joule heating and the rhs of the energy solve (one pass),
// compute_elecenergy_source, Source/ElecEnergySource.cpp:72-294 (no IB:
// every interface is regular), then the rhs of implicit_solve_scalar for the
// energy (Source/ScalarSolve.cpp:626-628, 826-844) and its boundary folding.
// Runs after the electron solve: interior face gradients of n_e come from
// the new (still scaled) electron solution in the state; everything else
// reads the Sborder data (densities from the snapshot sb, E and Te from the
// state, which the species solves do not change).
struct JouleArgs
{
    PhysParams P;
    Real* state;
    const Real* sb;           // Sborder species densities, NUM_SPECIES * N
    const Real* sb_old_een;   // Sborder_old energy density (N)
    const Real* expl;         // energy explicit source without the heating (N)
    const Real* psol;         // scaled potential solution (N)
    Real potscale;
    Real pdep0[3][2];
    const Real* pgconst[3][2];
    Real escale;              // electron density scale of the solution
    int floor_jh;
    Real rscale;              // timescale / elecenergy_scale
    int c_een, nc;
    int kind[3][2];
    const Real* gconst[3][2]; // species slabs (nc components)
    Real beta;
    const Real* iD;           // energy inverse coefficient (N)
    Real* rhs;                // energy rhs (N)
    double* partials;         // per-workgroup maxima of |rhs|
    // fine levels: box faces are coarse/fine faces.  Sborder beyond the box
    // comes from the ghosted copy (components [0, EFZ_ID], ghost width sng);
    // the potential's ghost is pdep0*x0 + pdep1*x1 + pgconst and the new
    // electron solution's ghost kCfD0*x0 + kCfD1*x1 + egconst.
    int cf;
    const Real* sbg;
    int sng;
    Real pdep1[3][2];
    const Real* egconst[3][2];
};

__global__ void __launch_bounds__(256) joule_energy_kernel(JouleArgs A)
{
    const dim3 bI_ = xcd_block();
    __shared__ double scratch[16];
    const Grid& g = A.P.g;
    const int i = bI_.x * blockDim.x + threadIdx.x;
    const int j = bI_.y * blockDim.y + threadIdx.y;
    const int k = bI_.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    double m = 0.0;
    if (in)
    {
        const long long o = g.idx(i, j, k);
        const int I[3] = {i, j, k};
        const int ei = A.P.eidx;
        const Real charge = species_charge(ei) * ECHARGE;
        const Real xc = A.psol[o];
        Real jh = 0.0;
        for (int d = 0; d < SD; ++d)
        {
            const long long st = (d == 0) ? 1 : (d == 1 ? (long long)g.n[0] : (long long)g.n[0] * g.n[1]);
            const int n = g.n[d];
            const int per = g.periodic[d];
            const long long base = o - (long long)I[d] * st;
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            const long long so = (long long)I[t1] + (long long)((SD > t1) ? g.n[t1] : 1) * (t2 < SD ? I[t2] : 0);
            for (int f = 0; f < 2; ++f)
            {
                // face index I[d] + f between cells lc = face-1 and rc = face
                const int fi = I[d] + f;
                const long long lo_ = base + (long long)ghost_map(fi - 1, n, per) * st;
                const long long ro_ = base + (long long)ghost_map(fi, n, per) * st;
                // Sborder component at the face's left (side 0) or right cell:
                // inside the box from the state (E, Te) or the density
                // snapshot, beyond a coarse/fine face from the ghosted copy
                const bool lout = A.cf && (fi - 1 < 0), rout = A.cf && (fi > n - 1);
                long long go_ = 0, gN = 0;
                if (A.cf && (lout || rout))
                {
                    const int q = A.sng;
                    const long long gx = g.n[0] + 2 * q, gy = (SD > 1) ? g.n[1] + 2 * q : 1, gz = (SD > 2) ? g.n[2] + 2 * q : 1;
                    gN = gx * gy * gz;
                    int G[3] = {I[0] + q, SD > 1 ? I[1] + q : 0, SD > 2 ? I[2] + q : 0};
                    G[d] = lout ? q - 1 : n + q;
                    go_ = (long long)G[0] + gx * ((long long)G[1] + gy * G[2]);
                }
                auto sbv = [&](int comp, int side) -> Real {
                    if ((side == 0 && lout) || (side == 1 && rout)) return A.sbg[(long long)comp * gN + go_];
                    const long long cc = side ? ro_ : lo_;
                    return (comp < NUM_SPECIES) ? A.sb[(long long)comp * N + cc] : A.state[(long long)comp * N + cc];
                };
                const Real etemp = (sbv(ETEMP_ID, 0) * 1 + sbv(ETEMP_ID, 1) * 1) / 2.0;
                Real Esum = 0.0;
                for (int q = 0; q < SD; ++q)
                {
                    const Real ef = (sbv(EFX_ID + q, 0) * 1 + sbv(EFX_ID + q, 1) * 1) / 2.0;
                    Esum += ef * ef;
                }
                const Real emag = sqrt(Esum);
                const Real ne = (sbv(ei, 0) * 1 + sbv(ei, 1) * 1) / 2.0;
                // face E: -grad of the potential solution (getGradSolution),
                // ghosts at domain faces as in the E-field kernel
                Real efield_face;
                if (f == 0)
                {
                    Real xm;
                    if (I[d] > 0) xm = A.psol[o - st];
                    else if (per) xm = A.psol[o + (long long)(n - 1) * st];
                    else xm = A.pdep0[d][0] * xc + A.pgconst[d][0][so] + A.pdep1[d][0] * ((n > 1) ? A.psol[o + st] : xc);
                    efield_face = (xc - xm) / g.dx[d];
                } else
                {
                    Real xp;
                    if (I[d] < n - 1) xp = A.psol[o + st];
                    else if (per) xp = A.psol[o - (long long)(n - 1) * st];
                    else xp = A.pdep0[d][1] * xc + A.pgconst[d][1][so] + A.pdep1[d][1] * ((n > 1) ? A.psol[o - st] : xc);
                    efield_face = (xp - xc) / g.dx[d];
                }
                efield_face *= -1.0 * A.potscale;
                Real nd = 0.0;
#pragma unroll
                for (int sp = 0; sp < NUM_SPECIES; ++sp) nd += (sbv(sp, 0) * 1 + sbv(sp, 1) * 1) / 2.0;
                const Real mu = specMob(ei, etemp, nd, emag, A.P.Tg);
                const Real dc = specDiff(ei, etemp, nd, emag, A.P.Tg);
                Real gne;
                if (A.cf && (fi == 0 || fi == n))
                {
                    // coarse/fine face: gradient of the new electron solution
                    // with its C/F ghost
                    const long long on = (long long)ei * N + o;
                    const Real x0 = A.state[on], x1 = (n > 1) ? A.state[on + (fi == 0 ? st : -st)] : x0;
                    const Real xg = kCfD0 * x0 + kCfD1 * x1 + A.egconst[d][fi == 0 ? 0 : 1][so];
                    gne = (fi == 0) ? (x0 - xg) / g.dx[d] : (xg - x0) / g.dx[d];
                    gne *= A.escale;
                } else if (fi == 0 || fi == n)
                {
                    // get_onesided_grad, Source/HelperFuncs.H:13-52 (outward sgn)
                    const int sgn = (fi == 0) ? -1 : 1;
                    const int cell = (sgn == 1) ? fi - 1 : fi;
                    const int c1 = cell - sgn, c2 = c1 - sgn;
                    const Real a0 = A.sb[(long long)ei * N + base + (long long)ghost_map(cell, n, per) * st];
                    const Real a1 = A.sb[(long long)ei * N + base + (long long)ghost_map(c1, n, per) * st];
                    const Real a2 = A.sb[(long long)ei * N + base + (long long)ghost_map(c2, n, per) * st];
                    Real g1, g2;
                    if (sgn == -1)
                    {
                        g1 = (a1 - a0) / g.dx[d];
                        g2 = (a2 - a1) / g.dx[d];
                    } else
                    {
                        g1 = (a0 - a1) / g.dx[d];
                        g2 = (a1 - a2) / g.dx[d];
                    }
                    gne = 2.0 * g1 - g2;
                } else
                {
                    const Real xr = A.state[(long long)ei * N + ro_], xl = A.state[(long long)ei * N + lo_];
                    gne = (xr - xl) / g.dx[d];
                    gne *= A.escale;
                }
                const Real cur = charge * (mu * ne * efield_face - dc * gne);
                jh += cur * efield_face;
            }
        }
        jh *= 0.5;
        if (A.floor_jh && jh < 0.0) jh = 0.0;
        A.state[(long long)EJH_ID * N + o] = jh;
        Real r = ((1.0 / A.P.dt) * A.sb_old_een[o] + 1.0 * (A.expl[o] + jh)) * A.rscale;
        m = fabs(r);
        // fold the inhomogeneous boundary data (as the side kernel does for
        // the other components)
        for (int d = 0; d < SD; ++d)
            for (int s = 0; s < 2; ++s)
            {
                const int kd = A.kind[d][s];
                if (g.periodic[d] || !(kd == BK_DIR || kd == BK_INHNEU || kd == BK_ROBIN)) continue;
                if (I[d] != (s ? g.n[d] - 1 : 0)) continue;
                const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                const int n1 = (SD > t1) ? g.n[t1] : 1, n2 = (SD > t2) ? g.n[t2] : 1;
                const long long so = (long long)I[t1] + (long long)n1 * (t2 < SD ? I[t2] : 0);
                const Real gc = A.gconst[d][s][(long long)A.c_een * n1 * n2 + so];
                const Real h = g.dx[d];
                r -= -A.beta * (1.0 / A.iD[o]) * gc / (h * h);
            }
        A.rhs[o] = r;
    }
    block_max_store(m, A.partials + flat_block_id(), scratch);
}

// Electron temperature from the solved energy and the Sborder electron
// density, floored (Source/ScalarSolve.cpp:955-987).
__global__ void __launch_bounds__(256) energy_te_kernel(Grid g, Real* __restrict__ state, const Real* __restrict__ sb_ne,
                                                        Real minetemp, int eidx)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.N) return;
    const long long N = g.N;
    Real te = twothird / K_B * state[(long long)EEN_ID * N + q] / sb_ne[q];
    if (te < minetemp)
    {
        te = minetemp;
        state[(long long)EEN_ID * N + q] = 1.5 * K_B * state[(long long)eidx * N + q] * minetemp;
    }
    state[(long long)ETEMP_ID * N + q] = te;
}

} // namespace vd
#endif
