// Discrete adjoint of one step of the vidyut3d path, hand-written, for the
// gradient of the time-integrated electrode current with respect to the
// case's run-time parameters (reaction prefactors, electron mobility and
// diffusivity) and the electrode's voltage amplitude.
//
// The forward step being differentiated is vd_ctx::step() with upwind
// fluxes, one corrector, table-driven chemistry (vd_rxn_table.H) and
// state-independent transport coefficients:
//   rho   = -(e/eps0) sum_k z_k n_k^n / ps, Dirichlet data folded in
//   L x   = rho                       (Poisson, symmetric)
//   E_d   = 0.5 (g_lo + g_hi),  g = -ps dx(x) / h;  EN = |E| / (sum n - n_e) / 1e-21
//   F     = n_L v_L (1+s)/2 + n_R v_R (1-s)/2,  v = mu E_d,  s = sign((v_L+v_R)/2)
//   expl  = sum_d (F_lo - F_hi)/h_d + N_A wdot(n/N_A, EN)
//   H y   = ((1/dt) n^n + expl) ts     (species Helmholtz, symmetric)
// and the objective per step is dt * (conduction + displacement current)
// through the electrode of current location 0 (Source/CurrentDens.cpp:17-456).
//
// Reverse step n, given lambda^{n+1} = dJ/dn^{n+1} (transported species):
//   A  objective terms of step n -> lambda^{n+1}, dJ/dPOT^n, dJ/dE^n;
//      species rhs of the adjoint solve
//   solve H mu = lambda^{n+1}
//   C  lambda^n from the rhs, the flux, the reactions and E/N; dJ/dE^n;
//      parameter terms (prefactors, mu_e, D_e)
//   D  dJ/dx from E and POT -> Poisson adjoint rhs; voltage terms
//   solve L y = xbar
//   E  lambda^n from the charge density; voltage terms of the folded rhs
// Every kernel is a gather over the cell's own stencil roles, so each output
// has one writer and every sum has a fixed order.
#ifndef VD_ADJOINT_H_
#define VD_ADJOINT_H_

#include "vd_physics.H"

namespace vd {

constexpr int kAdjPar = VD_NUM_CASE_PARAMS + 1;   // case parameters + voltage amplitude

__device__ inline int map_index(int I, int n, int per)
{
    return ghost_map(I, n, per);
}

__device__ inline long long along(const Grid& g, long long o, int d, int I_from, int I_to)
{
    const long long st = (d == 0) ? 1 : (d == 1 ? (long long)g.n[0] : (long long)g.n[0] * g.n[1]);
    return o + (long long)(I_to - I_from) * st;
}

struct AdjCommon
{
    Grid g;
    const Real* s0;       // state at the start of step n (densities n^n)
    const Real* s1;       // state after step n (n^{n+1}, POT^n, E^n, REF^n)
    int nt;
    int tsp[kMaxComp];
    Real rscale[kMaxComp];
    Real dt, ts, ps, Tg;
    int eidx;
    double* partials;     // kAdjPar * nblocks
};

// Collector weight of a cell for location 0: area * outward normal * value.
__device__ inline Real collector_weight(const Grid& g, int surf, const int C[3])
{
    const int dir = surf / 2;
    if (C[dir] < 0 || C[dir] >= g.n[dir]) return 0.0;
    Real vol = 1.0;
    for (int d = 0; d < SD; ++d) vol *= g.dx[d];
    const Real area = vol / g.dx[dir];
    const Real normal = (surf % 2 == 0) ? -1.0 : 1.0;
    amrex::GpuArray<int, SD> dlo, dhi;
    for (int d = 0; d < SD; ++d)
    {
        dlo[d] = 0;
        dhi[d] = g.n[d] - 1;
    }
    amrex::IntVect cid(AMREX_D_DECL(C[0], C[1], C[2]));
    return area * normal *
           user_transport::current_collector_value(cid, 0, surf, garr(g.problo), garr(g.probhi), dlo, dhi, garr(g.dx));
}

// ---------------------------------------------------------------------------
// A: objective of step n.
struct AdjObjArgs
{
    AdjCommon c;
    int surf;             // current_loc_surfaces[0]
    int live_n;           // displacement current of step n counted (time > 0)
    int has_next;         // step n+1 is part of the objective
    const Real* lam;      // lambda^{n+1} (NUM_SPECIES * N)
    Real* f0;             // out: species adjoint rhs (nt * N, solver layout)
    Real* ebar;           // out: dJ/dE^n (SD * N), initialised here
    Real* potbar;         // out: dJ/dPOT^n (N)
    double* bmax;         // nt * nblocks maxima of |f0|
};

__global__ void __launch_bounds__(256) adj_objective_kernel(AdjObjArgs A)
{
    __shared__ double scratch[16];
    const Grid& g = A.c.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z, bid = flat_block_id();
    const int dir = A.surf / 2;
    const int n = g.n[dir];
    const Real h = g.dx[dir];
    Real ladd[NUM_SPECIES];
    for (int s = 0; s < NUM_SPECIES; ++s) ladd[s] = 0.0;
    Real pbar_mu = 0.0, pbar_D = 0.0;
    long long o = 0;
    if (in)
    {
        o = g.idx(i, j, k);
        const int C[3] = {i, j, k};
        const int I = C[dir];
        const amrex::Array4<Real> sv = state_view(const_cast<Real*>(A.c.s1), g);
        // displacement current: E^n enters disp^n (+) and disp^{n+1} (-)
        const Real wc = collector_weight(g, A.surf, C);
        const bool cm = (A.c.s1[(long long)CMASK_ID * N + o] == 1.0);
        for (int d = 0; d < SD; ++d) A.ebar[(long long)d * N + o] = 0.0;
        if (cm) A.ebar[(long long)dir * N + o] = wc * EPS0 * (Real(A.live_n) - Real(A.has_next));
        // conduction current: faces of this cell's stencil roles
        int faces[4];
        int nf = 0;
        auto add_face = [&](int F) {
            for (int q = 0; q < nf; ++q)
                if (faces[q] == F) return;
            faces[nf++] = F;
        };
        const int per = g.periodic[dir];
        add_face(I);
        add_face(I + 1);
        if (I <= 2) add_face(0);
        if (I >= n - 3) add_face(n);
        if (per && I == n - 1) add_face(0);
        if (per && I == 0) add_face(n);
        Real pb = 0.0;
        for (int q = 0; q < nf; ++q)
        {
            const int F = faces[q];
            int CL[3] = {C[0], C[1], C[2]}, CR[3] = {C[0], C[1], C[2]};
            CL[dir] = F - 1;
            CR[dir] = F;
            const Real W = A.c.dt * 0.5 * (collector_weight(g, A.surf, CL) + collector_weight(g, A.surf, CR));
            if (W == 0.0) continue;
            const int L = map_index(F - 1, n, per), R = map_index(F, n, per);
            // gradient weights of this cell in the face gradient
            Real gw = 0.0;
            if (F > 0 && F < n)
            {
                if (I == F) gw += 1.0 / h;
                if (I == F - 1) gw -= 1.0 / h;
            } else if (F == 0)
            {
                if (I == 0) gw += -2.0 / h;
                if (I == 1) gw += 3.0 / h;
                if (I == 2) gw += -1.0 / h;
            } else
            {
                if (I == n - 1) gw += 2.0 / h;
                if (I == n - 2) gw += -3.0 / h;
                if (I == n - 3) gw += 1.0 / h;
            }
            const Real aw = 0.5 * Real(I == L) + 0.5 * Real(I == R);
            // face values (forward face_current)
            int FF[3] = {C[0], C[1], C[2]};
            FF[dir] = F;
            int LL[3] = {FF[0], FF[1], FF[2]};
            LL[dir] -= 1;
            const Real etemp = (sv(LL[0], LL[1], LL[2], ETEMP_ID) + sv(FF[0], FF[1], FF[2], ETEMP_ID)) / 2;
            Real nsp[NUM_SPECIES], ndens = 0.0;
            for (int sp = 0; sp < NUM_SPECIES; ++sp)
            {
                nsp[sp] = (sv(LL[0], LL[1], LL[2], sp) + sv(FF[0], FF[1], FF[2], sp)) / 2;
                ndens += nsp[sp];
            }
            Real Esum = 0.0;
            for (int d = 0; d < SD; ++d)
            {
                const Real ef = (sv(LL[0], LL[1], LL[2], EFX_ID + d) + sv(FF[0], FF[1], FF[2], EFX_ID + d)) / 2;
                Esum += ef * ef;
            }
            const Real emag = sqrt(Esum);
            Real fe, ge;
            if (F > 0 && F < n)
            {
                fe = -1.0 * (sv(FF[0], FF[1], FF[2], POT_ID) - sv(LL[0], LL[1], LL[2], POT_ID)) / h;
                ge = (sv(FF[0], FF[1], FF[2], A.c.eidx) - sv(LL[0], LL[1], LL[2], A.c.eidx)) / h;
            } else
            {
                const int sg = (F == 0) ? -1 : 1;
                fe = -1.0 * onesided_grad(sv, FF, sg, dir, POT_ID, h);
                ge = onesided_grad(sv, FF, sg, dir, A.c.eidx, h);
            }
            Real dJdE = 0.0;
            for (int sp = 0; sp < NUM_SPECIES; ++sp)
            {
                const int z = species_charge(sp);
                if (z == 0) continue;
                const Real Q = Real(z) * ECHARGE;
                const Real mu = specMob(sp, etemp, ndens, emag, A.c.Tg);
                const Real dc = specDiff(sp, etemp, ndens, emag, A.c.Tg);
                ladd[sp] += W * (Q * mu * fe * aw - Q * dc * gw);
                dJdE += Q * mu * nsp[sp];
            }
            pb += W * dJdE * (-gw);
            // parameter terms: each face counted by one cell
            const bool owner = (F < n) ? (I == F) : (I == n - 1);
            if (owner)
            {
                const Real Qe = -ECHARGE;
                pbar_mu += W * Qe * nsp[A.c.eidx] * fe;
                pbar_D += W * (-Qe) * ge;
            }
        }
        A.potbar[o] = pb;
    }
    // species adjoint rhs = lambda^{n+1} + objective terms
    for (int c = 0; c < A.c.nt; ++c)
    {
        double m = 0.0;
        if (in)
        {
            const int sp = A.c.tsp[c];
            const Real r = A.lam[(long long)sp * N + o] + ladd[sp];
            A.f0[(long long)c * N + o] = r;
            m = fabs(r);
        }
        block_max_store(m, A.bmax + (long long)c * nb + bid, scratch);
    }
    for (int p = 0; p < kAdjPar; ++p)
    {
        Real v = 0.0;
        if (p == VD_PARAM_MU_E) v = pbar_mu;
        if (p == VD_PARAM_D_E) v = pbar_D;
        block_sum_store(v, A.c.partials + (long long)p * nb + bid, scratch);
    }
}

// ---------------------------------------------------------------------------
// C: species rhs, flux, reactions and E/N.
struct AdjCellArgs
{
    AdjCommon c;
    const Real* mu;       // adjoint species solution (NUM_SPECIES * N, species index)
    Real* lam;            // out: lambda^n (NUM_SPECIES * N)
    Real* ebar;           // in/out: dJ/dE^n (SD * N)
    Real Dspec[kMaxComp]; // diffusivity of each transported species
};

__global__ void __launch_bounds__(256) adj_cell_kernel(AdjCellArgs A)
{
    __shared__ double scratch[16];
    const Grid& g = A.c.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z, bid = flat_block_id();
    Real pbar[kAdjPar];
    for (int p = 0; p < kAdjPar; ++p) pbar[p] = 0.0;
    if (in)
    {
        const long long o = g.idx(i, j, k);
        const int C[3] = {i, j, k};
        const Real* s0 = A.c.s0;
        const Real* s1 = A.c.s1;
        Real nbar[NUM_SPECIES], rbar[NUM_SPECIES];
        for (int s = 0; s < NUM_SPECIES; ++s)
        {
            nbar[s] = 0.0;
            rbar[s] = 0.0;
        }
        Real Eb[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < A.c.nt; ++c)
        {
            const int sp = A.c.tsp[c];
            const Real m = A.mu[(long long)sp * N + o];
            nbar[sp] += m * A.c.rscale[c] / A.c.dt;   // rhs = (n/dt + expl) ts
            rbar[sp] = m * A.c.rscale[c];              // dJ/d expl = dJ/d rxn
            // D_sp: rhs side -mu^T (dH/dD) y with y = n^{n+1}
            if (sp == A.c.eidx)
            {
                Real lap = 0.0;
                for (int d = 0; d < SD; ++d)
                {
                    const int n = g.n[d], per = g.periodic[d], I = C[d];
                    const Real y = s1[(long long)sp * N + o];
                    const Real ym = s1[(long long)sp * N + along(g, o, d, I, map_index(I - 1, n, per))];
                    const Real yp = s1[(long long)sp * N + along(g, o, d, I, map_index(I + 1, n, per))];
                    lap += (yp - 2.0 * y + ym) / (g.dx[d] * g.dx[d]);
                }
                pbar[VD_PARAM_D_E] += A.c.ts * m * lap;
            }
        }
        // flux: every (face, role) of this cell along each direction
        for (int d = 0; d < SD; ++d)
        {
            const int n = g.n[d], per = g.periodic[d], I = C[d];
            const Real h = g.dx[d];
            int fF[4], fR[4];   // face index, role (0 = L, 1 = R)
            int nf = 0;
            fF[nf] = I;
            fR[nf++] = 1;
            fF[nf] = I + 1;
            fR[nf++] = 0;
            if (!per && I == 0) { fF[nf] = 0; fR[nf++] = 0; }
            if (!per && I == n - 1) { fF[nf] = n; fR[nf++] = 1; }
            if (per && I == n - 1) { fF[nf] = 0; fR[nf++] = 0; }
            if (per && I == 0) { fF[nf] = n; fR[nf++] = 1; }
            for (int q = 0; q < nf; ++q)
            {
                const int F = fF[q];
                const int L = map_index(F - 1, n, per), R = map_index(F, n, per);
                const long long oL = along(g, o, d, I, L), oR = along(g, o, d, I, R);
                const Real EL = s1[(long long)(EFX_ID + d) * N + oL], ER = s1[(long long)(EFX_ID + d) * N + oR];
                for (int c = 0; c < A.c.nt; ++c)
                {
                    const int sp = A.c.tsp[c];
                    // dJ/dF = (ebar of the cell whose lo face F is - ebar of the cell whose hi face it is)/h
                    Real Fb = 0.0;
                    if (F < n) Fb += A.mu[(long long)sp * N + along(g, o, d, I, F)] * A.c.rscale[c];
                    if (F >= 1) Fb -= A.mu[(long long)sp * N + along(g, o, d, I, F - 1)] * A.c.rscale[c];
                    Fb /= h;
                    if (Fb == 0.0) continue;
                    const Real mob = specMob(sp, 0.0, 0.0, 0.0, A.c.Tg);
                    const Real vL = mob * EL, vR = mob * ER;
                    const Real sg = copysign(1.0, 0.5 * (vL + vR));
                    if (fR[q] == 0)
                    {
                        const Real nL = s0[(long long)sp * N + oL];
                        const Real wgt = Fb * 0.5 * (1.0 + sg);
                        nbar[sp] += wgt * vL;
                        Eb[d] += wgt * nL * mob;
                        if (sp == A.c.eidx) pbar[VD_PARAM_MU_E] += wgt * nL * EL;
                    } else
                    {
                        const Real nR = s0[(long long)sp * N + oR];
                        const Real wgt = Fb * 0.5 * (1.0 - sg);
                        nbar[sp] += wgt * vR;
                        Eb[d] += wgt * nR * mob;
                        if (sp == A.c.eidx) pbar[VD_PARAM_MU_E] += wgt * nR * ER;
                    }
                }
            }
        }
        // reactions at this cell: densities n^n, E/N of step n
        {
            Real nn[NUM_SPECIES];
            for (int s = 0; s < NUM_SPECIES; ++s) nn[s] = s0[(long long)s * N + o];
            const Real EN = s1[(long long)REF_ID * N + o];
            Real pr[VD_NUM_RXN];
            for (int r = 0; r < VD_NUM_RXN; ++r) pr[r] = 0.0;
            const Real enbar = vd_table_ckwc_vjp(nn, EN, rbar, nbar, pr);
            for (int r = 0; r < VD_NUM_RXN; ++r) pbar[r] += pr[r];
            // EN = |E| / nd / 1e-21, nd = sum n - n_e
            Real nd = 0.0;
            for (int s = 0; s < NUM_SPECIES; ++s) nd += nn[s];
            nd = nd - nn[A.c.eidx];
            Real Esum = 0.0;
            Real Ev[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < SD; ++d)
            {
                Ev[d] = s1[(long long)(EFX_ID + d) * N + o];
                Esum += Ev[d] * Ev[d];
            }
            const Real emag = sqrt(Esum);
            if (enbar != 0.0)
            {
                if (emag > 0.0)
                    for (int d = 0; d < SD; ++d) Eb[d] += enbar * (Ev[d] / emag) / nd / 1.0e-21;
                const Real dnd = -enbar * (emag / (nd * nd)) / 1.0e-21;
                for (int s = 0; s < NUM_SPECIES; ++s)
                    if (s != A.c.eidx) nbar[s] += dnd;
            }
        }
        for (int s = 0; s < NUM_SPECIES; ++s) A.lam[(long long)s * N + o] = nbar[s];
        for (int d = 0; d < SD; ++d) A.ebar[(long long)d * N + o] += Eb[d];
    }
    for (int p = 0; p < kAdjPar; ++p) block_sum_store(pbar[p], A.c.partials + (long long)p * nb + bid, scratch);
}

// ---------------------------------------------------------------------------
// D: dJ/dx (scaled potential) from E and POT; voltage terms of the E ghosts.
struct AdjEfieldArgs
{
    AdjCommon c;
    const Real* ebar;     // dJ/dE^n (SD * N)
    const Real* potbar;   // dJ/dPOT^n (N)
    Real dep0[3][2];      // homogeneous ghost factor per side (Dirichlet -1, Neumann 1)
    int surf;             // electrode of the voltage parameter
    Real dVdA;            // d(applied voltage)/d(amplitude) at the step's time
    Real* f0;             // out: Poisson adjoint rhs (N)
    double* bmax;         // nblocks maxima of |f0|
};

__global__ void __launch_bounds__(256) adj_efield_kernel(AdjEfieldArgs A)
{
    __shared__ double scratch[16];
    const Grid& g = A.c.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z, bid = flat_block_id();
    const Real ps = A.c.ps;
    double m = 0.0;
    Real vbar = 0.0;
    if (in)
    {
        const long long o = g.idx(i, j, k);
        const int C[3] = {i, j, k};
        Real xb = ps * A.potbar[o];
        for (int d = 0; d < SD; ++d)
        {
            const int n = g.n[d], per = g.periodic[d], I = C[d];
            const Real h = g.dx[d];
            const Real* Eb = A.ebar + (long long)d * N;
            // this cell is x_m of cell I+1 and x_p of cell I-1
            if (I + 1 < n) xb += 0.5 * ps / h * Eb[along(g, o, d, I, I + 1)];
            else if (per) xb += 0.5 * ps / h * Eb[along(g, o, d, I, 0)];
            if (I - 1 >= 0) xb -= 0.5 * ps / h * Eb[along(g, o, d, I, I - 1)];
            else if (per) xb -= 0.5 * ps / h * Eb[along(g, o, d, I, n - 1)];
            if (!per)
            {
                if (I == 0) xb += 0.5 * ps * A.dep0[d][0] / h * Eb[o];
                if (I == n - 1) xb -= 0.5 * ps * A.dep0[d][1] / h * Eb[o];
                // ghost constant gc = 2 V/ps on the electrode side
                const int ed = A.surf / 2, es = A.surf % 2;
                if (d == ed && es == 0 && I == 0) vbar += (0.5 * ps / h) * Eb[o] * 2.0 * A.dVdA / ps;
                if (d == ed && es == 1 && I == n - 1) vbar += (-0.5 * ps / h) * Eb[o] * 2.0 * A.dVdA / ps;
            }
        }
        A.f0[o] = xb;
        m = fabs(xb);
    }
    block_max_store(m, A.bmax + bid, scratch);
    for (int p = 0; p < kAdjPar; ++p)
        block_sum_store(p == VD_NUM_CASE_PARAMS ? vbar : 0.0, A.c.partials + (long long)p * nb + bid, scratch);
}

// ---------------------------------------------------------------------------
// E: charge density and the folded Dirichlet data of the Poisson rhs.
struct AdjRhoArgs
{
    AdjCommon c;
    const Real* y;        // Poisson adjoint solution (N)
    Real* lam;            // in/out: lambda^n
    int surf;
    Real dVdA;
    int do_spacechrg;
};

__global__ void __launch_bounds__(256) adj_rho_kernel(AdjRhoArgs A)
{
    __shared__ double scratch[16];
    const Grid& g = A.c.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const bool in = (i < g.n[0] && j < g.n[1]);
    const long long N = g.N;
    const long long nb = (long long)gridDim.x * gridDim.y * gridDim.z, bid = flat_block_id();
    Real vbar = 0.0;
    if (in)
    {
        const long long o = g.idx(i, j, k);
        const int C[3] = {i, j, k};
        const Real yv = A.y[o];
        if (A.do_spacechrg)
            for (int c = 0; c < A.c.nt; ++c)
            {
                const int sp = A.c.tsp[c];
                const int z = species_charge(sp);
                if (z != 0) A.lam[(long long)sp * N + o] += yv * (-ECHARGE / EPS0) * Real(z) / A.c.ps;
            }
        // rhs -= gc/h^2 in the electrode's boundary cells, gc = 2 V/ps
        const int ed = A.surf / 2, es = A.surf % 2;
        const int I = C[ed], n = g.n[ed];
        if ((es == 0 && I == 0) || (es == 1 && I == n - 1))
        {
            const Real h = g.dx[ed];
            vbar += -yv / (h * h) * 2.0 * A.dVdA / A.c.ps;
        }
    }
    for (int p = 0; p < kAdjPar; ++p)
        block_sum_store(p == VD_NUM_CASE_PARAMS ? vbar : 0.0, A.c.partials + (long long)p * nb + bid, scratch);
}

} // namespace vd
#endif
