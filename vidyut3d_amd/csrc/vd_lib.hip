// vidyut3d_amd per-case library: host step driver and C ABI over the gfx950
// kernels.  Built once per case directory, like the reference's one
// executable per case:
//   hipcc --offload-arch=gfx950 -DAMREX_SPACEDIM=<d> -I<case> -I compat
//         -I core vd_lib.hip <case>/Chemistry.cpp -shared -o libvd_<case>.so
#include <AMReX_compat.H>
#include <Chemistry.H>
#include <VarDefines.H>
#include <UnivConstants.H>
#include <ProbParm.H>
#include <Prob.H>
#include <UserFunctions.H>
#include <UserSources.H>

#include "vd_grid.H"
#include "vd_io.H"
#include "vd_reduce.H"
#include "vd_mg.H"
#include "vd_ktimer.H"
#include "vd_physics.H"
#if defined(VD_NUM_RXN) && defined(VD_NUM_CASE_PARAMS) && defined(VD_PARAM_MU_E) && defined(VD_PARAM_D_E)
#define VD_HAS_ADJOINT 1
#include "vd_adjoint.H"
#endif
#include "vd_operator.H"
#include "../../include/vidyut3d_amd.h"

#include <string>
#include <functional>
#include <limits>
#include <vector>
#include <memory>
#include <sstream>
#include <cstring>
#include <mutex>

using namespace vd;
using amrex::Real;

namespace {

std::string g_global_error;

// An argument or option the call cannot accept (VD_E_BADARG at the ABI).
struct BadArg : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

void load_inputs(const char* text)
{
    auto& T = amrex::ParmTable::get().entries;
    T.clear();
    std::istringstream is(text ? text : "");
    std::string line;
    while (std::getline(is, line))
    {
        auto h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = line.substr(0, eq);
        key.erase(0, key.find_first_not_of(" \t"));
        key.erase(key.find_last_not_of(" \t") + 1);
        std::istringstream vs(line.substr(eq + 1));
        std::vector<std::string> vals;
        std::string tok;
        while (vs >> tok) vals.push_back(tok);
        if (!key.empty()) T[key] = vals;
    }
}

int bc_kind(int code)
{
    switch (code)
    {
    case PERBC: return BK_PER;
    case DIRCBC: return BK_DIR;
    case HNEUBC: return BK_NEU;
    case IHNEUBC: return BK_INHNEU;
    case ROBINBC: return BK_ROBIN;
    case AXISBC: return BK_NEU;
    }
    throw std::runtime_error("unknown boundary code " + std::to_string(code));
}

__global__ void init_state_kernel(Real* s, Grid g, amrex::GeometryData gd, ProbParm* pp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    // global level indices (a fine box starts at g.off)
    amrex::Array4<Real> a(s, amrex::Dim3{g.off[0], g.off[1], g.off[2]},
                          amrex::Dim3{g.off[0] + g.n[0], g.off[1] + g.n[1], g.off[2] + g.n[2]}, NVAR);
    amrex::IntVect lo, hi;
    lo[0] = hi[0] = i + g.off[0];
#if AMREX_SPACEDIM > 1
    lo[1] = hi[1] = j + g.off[1];
#endif
#if AMREX_SPACEDIM > 2
    lo[2] = hi[2] = k + g.off[2];
#endif
    initdomaindata(amrex::Box(lo, hi), a, gd, pp);
}

__global__ void average_kernel(long long n, Real* __restrict__ dst, const Real* __restrict__ a)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = 0.5 * a[q] + 0.5 * dst[q];
}

__global__ void fill_kernel(long long n, Real* __restrict__ p, Real v)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

__global__ void fill_float_kernel(long long n, float* __restrict__ p)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0.0f;
}

__global__ void bg_update_kernel(long long N, Real* __restrict__ state, const Real* __restrict__ rxn, int sp, Real dt,
                                 int bound, Real floor)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    Real& v = state[(long long)sp * N + q];
    v += rxn[(long long)sp * N + q] * dt;
    if (v < floor && bound) v = floor;
}

} // namespace

// One static refinement level (a box strictly inside its parent level):
// its state and face arrays, a ghosted Sborder copy for its stencils, and
// one-level solvers whose box sides are coarse/fine sides.
struct FineLevel
{
    Grid g{};
    LevelMap m{};
    int ng = 1;                        // ghost width of the Sborder copy
    DevBuf<Real> state, phi_sol, sbg, sbsnap, een_expl, rxn, pfraw, sfraw, rho;
    DevBuf<Real> flux[3];
    MGSolver<SD, Real> pois, spec;
    CfSlabs sl{};
    // C/F data: u_int of the parent potential (pslab) and species (sslab)
    // for residual folding and reflux, of the parent correction (tslab);
    // (8/15) u_int of the final parent potential (pgc, E field) and of the
    // parent's new electron solution (eslab, Joule heating)
    DevBuf<double> pslab, sslab, tslab, pgc, eslab;
    DevBuf<double> part, part_id, part_e, partn;
    DevBuf<unsigned long long> red, red2, nrm;
    long long nblk = 0;
};

// Scratch of vd_explicit_rhs_noalloc (the face fluxes of one call): a
// file-scope device array, so that entry point allocates nothing.  Each GPU
// the library runs on holds its own copy until the process ends; calls that
// overlap on one GPU share it and race.  Host code never passes the array's
// name to a kernel (on the host it only stands for the symbol): the device
// address comes from hipGetSymbolAddress, once per device.
constexpr long long kNoAllocScratchDoubles = 8LL << 20;   // 64 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

double* noalloc_scratch_ptr()
{
    static std::mutex m;
    static std::vector<double*> per_device;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(m);
    if ((int)per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev])
    {
        double* p = nullptr;
        hip_check(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        per_device[dev] = p;
    }
    return per_device[dev];
}

struct vd_ctx;
// context whose run-time case parameters are in the process-wide copies
static const vd_ctx* g_param_owner = nullptr;

struct vd_ctx
{
    std::string err;
    hipStream_t stream = nullptr;
    Grid g{};
    // parameters (defaults: Source/Vidyut.H:221-401)
    Real fixed_dt = 1e-12;
    int num_correctors = 1;
    int hyp_order = 1, weno_scheme = 3;
    Real gas_temperature = T_NTP, gas_pressure = P_NTP;
    std::vector<int> bg_species;
    int pot_bc_lo[3] = {ROBINBC, ROBINBC, ROBINBC}, pot_bc_hi[3] = {ROBINBC, ROBINBC, ROBINBC};
    int eden_bc_lo[3] = {ROBINBC, ROBINBC, ROBINBC}, eden_bc_hi[3] = {ROBINBC, ROBINBC, ROBINBC};
    int eenrg_bc_lo[3] = {ROBINBC, ROBINBC, ROBINBC}, eenrg_bc_hi[3] = {ROBINBC, ROBINBC, ROBINBC};
    int ion_bc_lo[3] = {ROBINBC, ROBINBC, ROBINBC}, ion_bc_hi[3] = {ROBINBC, ROBINBC, ROBINBC};
    int neu_bc_lo[3] = {ROBINBC, ROBINBC, ROBINBC}, neu_bc_hi[3] = {ROBINBC, ROBINBC, ROBINBC};
    int voltage_profile = 0;
    Real vamp_lo[3] = {0, 0, 0}, vamp_hi[3] = {0, 0, 0};
    Real vfreq = 1.0, vdur = 0.0, vcen = 0.0;
    Real reltol = 1e-10, abstol = 1e-10;
    int pre_smooth = 2, post_smooth = 2, maxiter = 100, max_coarsening = 10;
    // vidyut.mg_fused_smoothing: plane-marching fused smoothing passes on
    // levels of >= 4M cells (1), on every eligible level (2), off (0)
    int mg_fused = 1;
    // vidyut.mg_fused_min_cells: smallest level (cells x components) that
    // runs the fused passes when mg_fused_smoothing = 1
    long long mg_fused_min = 1LL << 22;
    // vidyut.mg_graphs: replay the coarse part of a one-component solver's
    // V-cycle from a captured HIP graph when its finest level has at most 4M
    // cells (1), on every level size (2), or launch it eagerly (0)
    int mg_graphs = 1;
    // vidyut.flux_march: WENO5 faces by the marching kernels (1) or the
    // per-face kernel (0)
    int flux_march = 1;
    // vidyut.mg_smooth_iters: species components whose operator is
    // diagonally dominant enough (Jacobi bound <= 1e-2) iterate smoothing
    // passes instead of V-cycles, single-level runs (1), or always V-cycles (0)
    int mg_smooth_iters = 1;
    // vidyut.mg_direct_bottom: the V-cycle solves the first coarse level
    // whose extents are all at most this directly (separable operators; LDS
    // tiled products up to 64, rocBLAS GEMMs above); 0: red-black sweeps
    // down to the coarsest level
    int mg_direct_bottom = 128;
    // vidyut.mg_direct_finest: on single-level runs the finest level is a
    // candidate for the direct solve too (its extents all at most
    // mg_direct_bottom, a separable operator): such a solve is one direct
    // solve of the residual equation and the usual stopping test (0: the
    // V-cycle's first coarse level at the earliest)
    int mg_direct_finest = 0;
    int bound_specden = 0;
    Real min_species_density = 1e-20, min_electron_density = 1e12, min_electron_temp = 300.0;
    int using_LFA = 0, elecenergy_solve = 0, floor_jh = 0;
    int do_reactions = 1, do_transport = 1, do_spacechrg = 1;
    int user_defined_potential = 0, user_defined_species = 0, user_defined_vel = 0;
    int do_bg_reactions = 0;
    Real potscale = 1.0, eescale = 1.0, timescale = 1.0;
    Real specscales[NUM_SPECIES];
    int eidx = 0;
    Real cur_time = 0.0;
    int istep = 0;
    // plotfiles and checkpoints (amr.* keys, Source/Vidyut.H:371-381,
    // Source/Vidyut.cpp:368-382), written as InitData and Evolve write them
    std::string plot_file = "plt", chk_file = "chk", restart_chkfile;
    int plot_int = -1, plot_int_old = -1, chk_int = -1, chk_int_old = -1;
    Real plot_time = -1.0, chk_time = -1.0;
    int plotfilenum = 0, chkfilenum = 0;
    Real plottime = 0.0, chktime = 0.0;
    // electrode currents (Source/Vidyut.H:290,395; Source/Vidyut.cpp:442,512-533)
    int track_current_den = 0, track_integrated_currents = 0, ncurrent_locs = 0;
    int current_loc_surfaces[MAX_CURRENT_LOCS] = {0};
    double int_cond[MAX_CURRENT_LOCS] = {0}, int_disp[MAX_CURRENT_LOCS] = {0}, int_area[MAX_CURRENT_LOCS] = {0};
    // time integral of conduction + displacement current per location:
    // sum over steps of dt * current (the adjoint's objective)
    double int_time[MAX_CURRENT_LOCS] = {0};

    // transported species (electron first, then the others in index order)
    std::vector<int> tsp;

    // device data
    // the state the context advances: its own buffer, or a caller's device
    // buffer of the same layout attached with vd_attach_state (the caller
    // keeps ownership of that one)
    struct StateView
    {
        Real* p = nullptr;
        size_t n = 0;
    } state;
    DevBuf<Real> state_own, state_old, phi_sol, flux[3], rxn;
    double* scratch_dev = nullptr;    // this device's copy of the no-alloc scratch
    DevBuf<ProbParm> d_pp;
    DevBuf<unsigned long long> red;   // [0..31] species bnorm, [32] poisson bnorm, [33..35] time scales
    DevBuf<double> part;              // workgroup partial maxima
    DevBuf<double> part_id;           // workgroup maxima of 1/D and -1/D
    DevBuf<double> part_split;        // first-stage maxima of part / part_id rows
    DevBuf<unsigned long long> red2;  // [0..31] max 1/D, [32..63] max -1/D
    DevBuf<Real> eold;                // SD * N: E at the start of a step (displacement current)
    // electron energy: Sborder densities kept across the species solves,
    // the energy's explicit source and its rhs maxima
    DevBuf<Real> sbsnap, een_expl;
    DevBuf<double> part_e;
    DevBuf<double> cur_part, cur_dev; // integrated-current partial sums and results
    long long nblk = 0;               // workgroups of a cell launch
    struct Slab
    {
        DevBuf<Real> g, a, b, f, c;
        int n1 = 0, n2 = 0;
    };
    Slab pslab[3][2], sslab[3][2];
    MGSolver<SD, Real> pois, spec;     // FP64 multigrid
    // three-term photoionization (vidyut.do_photoionization,
    // Source/Evolve.cpp:756-799, Source/PhotoIonizationSolve.cpp): one
    // Helmholtz solve per term, solutions kept for the cell pass
    int do_photo = 0, photoion_sp = -1, user_defined_photoionization = 0;
    int photo_bc_lo[3] = {ROBINBC, ROBINBC, ROBINBC}, photo_bc_hi[3] = {ROBINBC, ROBINBC, ROBINBC};
    MGSolver<SD, Real> photo;
    Slab phslab[3][2];
    DevBuf<Real> photo_sol;            // 3 * N
    DevBuf<double> photo_part;
    DevBuf<unsigned long long> photo_red;
    int last_photo_iters = 0;
    // the step adds the photoionization terms to the reaction source (the
    // member entry points of update_rxnsrc/update_explsrc do not)
    bool photo_in_step = false;
    // mixed form: FP32 Poisson V-cycles inside FP64 iterative refinement.  The
    // species solves stay FP64: their operator is so diagonally dominant
    // that one FP64 V-cycle reaches the tolerance, where an FP32
    // preconditioner needs two refinement steps.
    MGSolver<SD, float> pois32;
    int mixed = 0;                      // vidyut.mixed_precision
    DevBuf<float> flux32[3];
    KTimer kt;

    // static refinement: levels 1..max_level (ratio 2, each box the central
    // half of its parent per direction), no subcycling
    int max_level = 0;
    std::vector<std::unique_ptr<FineLevel>> fine;
    int flo[8][3], fnn[8][3];          // box origin (global level index) and extents per level

    // accessors of the active solvers' FP64 level-0 data
    double* pf0() { return mixed ? pois32.f0.p : pois.f0.p; }
    double* sf0() { return spec.f0.p; }
    double* siD0() { return spec.iD0.p; }
    SideSpec& pside(int c, int d, int s) { return mixed ? pois32.side(c, d, s) : pois.side(c, d, s); }
    SideSpec& sside(int c, int d, int s) { return spec.side(c, d, s); }

    // last-step diagnostics
    int last_pot_iters = 0, last_spec_iters = 0;
    double last_ts[3] = {0, 0, 0};
    double last_pot_r0 = 0, last_pot_r = 0, last_pot_b = 0;
    double last_spec_r0 = 0, last_spec_r = 0, last_spec_b = 0;

    PhysParams phys(Real t, Real dt) const
    {
        PhysParams P;
        P.g = g;
        P.Tg = gas_temperature;
        P.Pg = gas_pressure;
        P.time = t;
        P.dt = dt;
        P.timescale = timescale;
        P.user_vel = user_defined_vel;
        P.user_species = user_defined_species;
        P.user_potential = user_defined_potential;
        P.user_photo = user_defined_photoionization;
        P.do_transport = do_transport;
        P.do_reactions = do_reactions;
        P.eidx = eidx;
        return P;
    }

    void read_params()
    {
        amrex::ParmParse pg("geometry");
        std::vector<double> lo(SD, 0.0), hi(SD, 1.0);
        std::vector<int> per(SD, 0), ncell(SD, 1);
        pg.queryarr("prob_lo", lo, 0, SD);
        pg.queryarr("prob_hi", hi, 0, SD);
        pg.queryarr("is_periodic", per, 0, SD);
        int coord = 0;
        pg.query("coord_sys", coord);
        if (coord != 0 && !(coord == 1 && SD == 2))
            throw std::runtime_error("geometry.coord_sys must be 0 (Cartesian) or, in 2D, 1 (axisymmetric RZ)");
        g.coord = coord;
        amrex::ParmParse pa("amr");
        pa.queryarr("n_cell", ncell, 0, SD);
        pa.query("plot_file", plot_file);
        pa.query("plot_int", plot_int);
        plot_int_old = plot_int;
        pa.query("plot_int_old", plot_int_old);
        pa.query("plot_time", plot_time);
        pa.query("chk_file", chk_file);
        pa.query("chk_int", chk_int);
        chk_int_old = chk_int;
        pa.query("chk_int_old", chk_int_old);
        pa.query("chk_time", chk_time);
        pa.query("restart", restart_chkfile);
        pa.query("max_level", max_level);
        if (max_level < 0 || max_level > 7) throw std::runtime_error("amr.max_level must be in [0, 7]");
        if (max_level > 0)
        {
            std::vector<int> rr(1, 2);
            pa.queryarr("ref_ratio", rr);
            for (int r : rr)
                if (r != 2) throw std::runtime_error("only amr.ref_ratio = 2 is supported");
        }
        for (int d = 0; d < 3; ++d)
        {
            g.n[d] = (d < SD) ? ncell[d] : 1;
            g.problo[d] = (d < SD) ? lo[d] : 0.0;
            g.probhi[d] = (d < SD) ? hi[d] : 1.0;
            g.periodic[d] = (d < SD) ? per[d] : 0;
            g.dx[d] = (d < SD) ? (hi[d] - lo[d]) / ncell[d] : 1.0;
            if (d < SD && ncell[d] < 1) throw std::runtime_error("amr.n_cell must be positive");
        }
        g.dim = SD;
        g.N = (long long)g.n[0] * g.n[1] * g.n[2];
        // static nesting: level l covers the central half of level l-1 per
        // direction; its parent region keeps >= 4 parent cells to the parent
        // box (ghost fills, slopes and C/F stencils stay inside the parent)
        for (int d = 0; d < 3; ++d)
        {
            flo[0][d] = 0;
            fnn[0][d] = g.n[d];
        }
        for (int l = 1; l <= max_level; ++l)
            for (int d = 0; d < 3; ++d)
            {
                if (d >= SD)
                {
                    flo[l][d] = 0;
                    fnn[l][d] = 1;
                    continue;
                }
                const int pn = fnn[l - 1][d];
                if (pn % 4 != 0 || pn / 4 < 4)
                    throw std::runtime_error("static refinement needs parent extents divisible by 4 and >= 16");
                flo[l][d] = 2 * (flo[l - 1][d] + pn / 4);
                fnn[l][d] = pn;
            }

        amrex::ParmParse pp_("vidyut");
        pp_.query("dt", fixed_dt);
        int adaptive = 0;
        pp_.query("adaptive_dt", adaptive);
        if (adaptive) throw std::runtime_error("adaptive_dt is not supported (fixed dt only)");
        pp_.query("num_timestep_correctors", num_correctors);
        pp_.query("hyp_order", hyp_order);
        pp_.query("weno_scheme", weno_scheme);
        pp_.query("gas_temperature", gas_temperature);
        pp_.query("gas_pressure", gas_pressure);
        pp_.queryarr("bg_species_ids", bg_species);
        auto qa = [&](const char* nm, int* a) {
            std::vector<int> v(a, a + SD);
            pp_.queryarr(nm, v, 0, SD);
            for (int d = 0; d < SD; ++d) a[d] = v[d];
        };
        qa("pot_bc_lo", pot_bc_lo);
        qa("pot_bc_hi", pot_bc_hi);
        qa("eden_bc_lo", eden_bc_lo);
        qa("eden_bc_hi", eden_bc_hi);
        qa("eenrg_bc_lo", eenrg_bc_lo);
        qa("eenrg_bc_hi", eenrg_bc_hi);
        qa("ion_bc_lo", ion_bc_lo);
        qa("ion_bc_hi", ion_bc_hi);
        qa("neutral_bc_lo", neu_bc_lo);
        qa("neutral_bc_hi", neu_bc_hi);
        pp_.query("voltage_profile", voltage_profile);
        {
            std::vector<double> v(vamp_lo, vamp_lo + SD), w(vamp_hi, vamp_hi + SD);
            pp_.queryarr("voltage_amp_lo", v, 0, SD);
            pp_.queryarr("voltage_amp_hi", w, 0, SD);
            for (int d = 0; d < SD; ++d)
            {
                vamp_lo[d] = v[d];
                vamp_hi[d] = w[d];
            }
        }
        if (voltage_profile == 1) pp_.get("voltage_freq", vfreq);
        if (voltage_profile == 2)
        {
            pp_.get("voltage_dur", vdur);
            pp_.get("voltage_center", vcen);
        }
        pp_.query("linsolve_reltol", reltol);
        pp_.query("linsolve_abstol", abstol);
        pp_.query("linsolve_num_pre_smooth", pre_smooth);
        pp_.query("linsolve_num_post_smooth", post_smooth);
        pp_.query("linsolve_maxiter", maxiter);
        pp_.query("linsolve_max_coarsening_level", max_coarsening);
        pp_.query("mg_fused_smoothing", mg_fused);
        {
            double fm = (double)mg_fused_min;
            pp_.query("mg_fused_min_cells", fm);
            if (!(fm >= 0.0)) throw BadArg("vidyut.mg_fused_min_cells must be >= 0");
            mg_fused_min = (long long)fm;
        }
        pp_.query("mg_graphs", mg_graphs);
        pp_.query("flux_march", flux_march);
        pp_.query("mg_smooth_iters", mg_smooth_iters);
        pp_.query("mg_direct_bottom", mg_direct_bottom);
        pp_.query("mg_direct_finest", mg_direct_finest);
        if (mg_direct_bottom < 0) throw BadArg("vidyut.mg_direct_bottom must be >= 0");
        pp_.query("bound_specden", bound_specden);
        pp_.query("min_species_density", min_species_density);
        pp_.query("min_electron_density", min_electron_density);
        pp_.query("min_electron_temp", min_electron_temp);
        pp_.query("using_LFA", using_LFA);
        pp_.query("elecenergy_solve", elecenergy_solve);
        pp_.query("floor_jh", floor_jh);
        pp_.query("do_reactions", do_reactions);
        pp_.query("do_transport", do_transport);
        pp_.query("do_spacechrg", do_spacechrg);
        pp_.query("user_defined_potential", user_defined_potential);
        pp_.query("user_defined_species", user_defined_species);
        pp_.query("user_defined_vel", user_defined_vel);
        pp_.query("do_bg_reactions", do_bg_reactions);
        pp_.query("track_current_den", track_current_den);
        pp_.query("track_integrated_currents", track_integrated_currents);
        if (track_integrated_currents)
        {
            std::vector<int> locs;
            pp_.queryarr("current_loc_surfaces", locs);
            if ((int)locs.size() > MAX_CURRENT_LOCS) throw std::runtime_error("Reduce the number of ncurrent locs");
            ncurrent_locs = (int)locs.size();
            for (int i = 0; i < ncurrent_locs; ++i)
            {
                if (locs[i] < 0 || locs[i] >= 2 * SD) throw std::runtime_error("bad current_loc_surfaces entry");
                current_loc_surfaces[i] = locs[i];
            }
        }
        int ib = 0;
        pp_.query("using_ib", ib);
        if (ib) throw std::runtime_error("using_ib is not supported");
        int coupling = 0;
        pp_.query("electron_poisson_coupling", coupling);
        if (coupling) throw std::runtime_error("electron_poisson_coupling is not supported");
        pp_.query("do_photoionization", do_photo);
        pp_.query("photoion_ID", photoion_sp);
        pp_.query("user_defined_photoionization", user_defined_photoionization);
        qa("photoion_bc_lo", photo_bc_lo);
        qa("photoion_bc_hi", photo_bc_hi);
        if (do_photo && (photoion_sp < 0 || photoion_sp >= NUM_SPECIES))
            throw std::runtime_error("do_photoionization needs vidyut.photoion_ID, a species index");
        pp_.query("mixed_precision", mixed);
        pp_.query("potential_scale", potscale);
        pp_.query("elecenergy_scale", eescale);
        pp_.query("time_scale", timescale);
        for (int i = 0; i < NUM_SPECIES; ++i) specscales[i] = 1.0;
        std::vector<int> sid;
        std::vector<double> sval;
        pp_.queryarr("scaling_species_ids", sid);
        pp_.queryarr("scale_values", sval);
        if (sid.size() != sval.size()) throw std::runtime_error("scaling lists dont have the same length");
        for (size_t i = 0; i < sid.size(); ++i) specscales[sid[i]] = sval[i];
        if (!(hyp_order == 1 || hyp_order == 2 || hyp_order == 5))
            throw std::runtime_error("Specified hyp_order not implemented yet");
        if (hyp_order == 5 && !(weno_scheme >= 1 && weno_scheme <= 3)) throw std::runtime_error("Unknown WENO scheme");
        for (int d = 0; d < SD; ++d)
            if (g.periodic[d] && (g.n[d] < 2 || g.n[d] % 2 != 0))
                throw std::runtime_error("a periodic direction needs an even number of cells (red-black smoothing)");
        {
            const int pm = (g.periodic[0] ? 1 : 0) | (g.periodic[1] ? 2 : 0) | (g.periodic[2] ? 4 : 0);
            int cn[3];
            mg_coarsest_extents(g.n, SD, pm, max_coarsening, cn);
            for (int d = 0; d < SD; ++d)
                if (cn[d] > kMaxCoarsestExtent)
                    throw std::runtime_error("the coarsest multigrid level would have " + std::to_string(cn[d]) +
                                             " cells in direction " + std::to_string(d) + " (at most " +
                                             std::to_string(kMaxCoarsestExtent) +
                                             "): choose extents with more factors of 2 or allow more coarsening");
        }
        if (track_current_den)
            for (int d = 0; d < SD; ++d)
                if (g.n[d] < 3) throw std::runtime_error("track_current_den needs >= 3 cells per direction");
        std::vector<std::string> names;
        CKSYMS_STR(names);
        eidx = -1;
        for (const char* nm : {"E", "E-", "e"})
            for (size_t q = 0; q < names.size(); ++q)
                if (eidx < 0 && names[q] == nm) eidx = int(q);
        if (eidx < 0) throw std::runtime_error("electron species not found");
        tsp.clear();
        tsp.push_back(eidx);
        for (int sp = 0; sp < NUM_SPECIES; ++sp)
        {
            bool bg = std::find(bg_species.begin(), bg_species.end(), sp) != bg_species.end();
            if (!bg && sp != eidx) tsp.push_back(sp);
        }
        // the electron energy is the last component of the batched solver
        if (elecenergy_solve) tsp.push_back(EEN_ID);
        if (elecenergy_solve)
            for (int d = 0; d < SD; ++d)
                if (g.n[d] < 3) throw std::runtime_error("elecenergy_solve needs >= 3 cells per direction");
        if ((int)tsp.size() > kMaxComp) throw std::runtime_error("too many transported species");
        if (max_level > 0)
        {
            if (num_correctors != 1) throw std::runtime_error("static refinement needs num_timestep_correctors = 1");
            if (g.coord) throw std::runtime_error("static refinement runs Cartesian geometry (coord_sys = 0)");
            if (do_photo) throw std::runtime_error("static refinement runs without photoionization");
            if (mixed) throw std::runtime_error("static refinement runs the FP64 form");
            if (track_current_den) throw std::runtime_error("static refinement does not track electrode currents");
        }
    }

    const int* bclo_of(int sp) const
    {
        if (sp == EEN_ID) return eenrg_bc_lo;
        return sp == eidx ? eden_bc_lo : (species_charge(sp) != 0 ? ion_bc_lo : neu_bc_lo);
    }
    const int* bchi_of(int sp) const
    {
        if (sp == EEN_ID) return eenrg_bc_hi;
        return sp == eidx ? eden_bc_hi : (species_charge(sp) != 0 ? ion_bc_hi : neu_bc_hi);
    }
    // solver scale of a transported component (Source/ScalarSolve.cpp:826-844)
    Real comp_scale(int sp) const { return sp == EEN_ID ? eescale : specscales[sp]; }
    // index of the energy component (-1 without the energy equation)
    int een_comp() const { return elecenergy_solve ? (int)tsp.size() - 1 : -1; }

    void setup()
    {
        hip_check(hipStreamCreate(&stream), "hipStreamCreate");
        kt.stream = stream;
        const long long N = g.N;
        state_own.alloc(N * NVAR);
        state.p = state_own.p;
        state.n = state_own.n;
        scratch_dev = noalloc_scratch_ptr();
        if (num_correctors > 1) state_old.alloc(N * NVAR);
        phi_sol.alloc(N);
        const int nt = (int)tsp.size();
        for (int d = 0; d < SD; ++d)
        {
            long long nf = (long long)(g.n[0] + (d == 0)) * (g.n[1] + (d == 1)) * (g.n[2] + (d == 2));
            if (mixed) flux32[d].alloc(nf * nt);
            else flux[d].alloc(nf * nt);
        }
        if (do_bg_reactions) rxn.alloc(N * (NUM_SPECIES + 1));
        red.alloc(64);
        {
            Launch3 L = cell_launch(g.n);
            nblk = (long long)L.grid.x * L.grid.y * L.grid.z;
            // per-wave partials of the cell-source kernel: kCellWaves per workgroup
            part.alloc((size_t)nblk * kCellWaves * std::max<size_t>(tsp.size(), 3));
            part_id.alloc((size_t)nblk * kCellWaves * 2 * tsp.size());
            part_split.alloc((size_t)3 * kMaxComp * kMaxSplit);
            red2.alloc(64);
            if (track_current_den) eold.alloc((size_t)SD * N);
            if (elecenergy_solve)
            {
                sbsnap.alloc((size_t)NUM_SPECIES * N);
                een_expl.alloc((size_t)N);
                part_e.alloc((size_t)nblk);
            }
            if (track_current_den && track_integrated_currents && ncurrent_locs > 0)
            {
                cur_part.alloc((size_t)nblk * 3 * ncurrent_locs);
                cur_dev.alloc(3 * MAX_CURRENT_LOCS);
            }
        }
        d_pp.alloc(1);
        ProbParm h;
        amrex_probinit(h, *d_pp.p);
        // boundary slabs
        for (int d = 0; d < SD; ++d)
        {
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            const int n1 = (SD > t1) ? g.n[t1] : 1, n2 = (SD > t2) ? g.n[t2] : 1;
            for (int s = 0; s < 2; ++s)
            {
                for (Slab* sl : {&pslab[d][s], &sslab[d][s]})
                {
                    const int nc = (sl == &pslab[d][s]) ? 1 : nt;
                    sl->n1 = n1;
                    sl->n2 = n2;
                    const long long sz = (long long)n1 * n2 * nc;
                    sl->g.alloc(sz);
                    sl->a.alloc(sz);
                    sl->b.alloc(sz);
                    sl->f.alloc(sz);
                    sl->c.alloc(sz);
                    fill_kernel<<<(unsigned)((sz + 255) / 256), 256, 0, stream>>>(sz, sl->c.p, 0.0);
                }
            }
        }
        int per_mask = (g.periodic[0] ? 1 : 0) | (g.periodic[1] ? 2 : 0) | (g.periodic[2] ? 4 : 0);
        if (mixed) cfg_pois(pois32, per_mask);
        else cfg_pois(pois, per_mask);
        cfg_spec(spec, per_mask);
        if (do_photo)
        {
            for (int d = 0; d < SD; ++d)
            {
                const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                const int n1 = (SD > t1) ? g.n[t1] : 1, n2 = (SD > t2) ? g.n[t2] : 1;
                for (int s2 = 0; s2 < 2; ++s2)
                {
                    Slab& sl = phslab[d][s2];
                    sl.n1 = n1;
                    sl.n2 = n2;
                    const long long sz = (long long)n1 * n2;
                    sl.g.alloc(sz);
                    sl.a.alloc(sz);
                    sl.b.alloc(sz);
                    sl.f.alloc(sz);
                    sl.c.alloc(sz);
                    fill_kernel<<<(unsigned)((sz + 255) / 256), 256, 0, stream>>>(sz, sl.c.p, 0.0);
                }
            }
            photo_sol.alloc((size_t)3 * g.N);
            const Launch3 L = cell_launch(g.n);
            photo_part.alloc((size_t)3 * L.grid.x * L.grid.y * L.grid.z * kCellWaves);
            photo_red.alloc(3);
            cfg_photo(photo, per_mask);
        }
        for (int l = 1; l <= max_level; ++l) setup_fine(l);
        hip_check(hipStreamSynchronize(stream), "setup sync");
    }

    // Buffers and one-level solvers of fine level l (its box sides are all
    // coarse/fine sides).
    void setup_fine(int l)
    {
        fine.emplace_back(new FineLevel);
        FineLevel& F = *fine.back();
        const int nt = (int)tsp.size();
        const FineLevel* P = (l > 1) ? fine[l - 2].get() : nullptr;
        F.g = g;
        for (int d = 0; d < 3; ++d)
        {
            F.g.n[d] = fnn[l][d];
            F.g.dx[d] = (d < SD) ? g.dx[d] / double(1 << l) : g.dx[d];
            F.g.periodic[d] = 0;
            F.g.off[d] = flo[l][d];
        }
        F.g.N = (long long)F.g.n[0] * F.g.n[1] * F.g.n[2];
        F.m.dim = SD;
        for (int d = 0; d < 3; ++d)
        {
            F.m.n[d] = F.g.n[d];
            F.m.pn[d] = fnn[l - 1][d];
            F.m.plo[d] = (d < SD) ? flo[l][d] / 2 - flo[l - 1][d] : 0;
        }
        F.m.N = F.g.N;
        F.m.pN = P ? P->g.N : g.N;
        F.ng = (hyp_order == 1) ? 1 : (hyp_order == 2 ? 2 : 3);
        const long long N = F.g.N;
        F.state.alloc(N * NVAR);
        F.phi_sol.alloc(N);
        {
            long long gN = 1;
            for (int d = 0; d < SD; ++d) gN *= F.g.n[d] + 2 * F.ng;
            F.sbg.alloc(gN * (EFZ_ID + 1));
        }
        for (int d = 0; d < SD; ++d)
            F.flux[d].alloc((long long)(F.g.n[0] + (d == 0)) * (F.g.n[1] + (d == 1)) * (F.g.n[2] + (d == 2)) * nt);
        F.pfraw.alloc(N);
        F.sfraw.alloc(N * nt);
        F.rho.alloc(N * nt);
        if (do_bg_reactions) F.rxn.alloc(N * (NUM_SPECIES + 1));
        long long off = 0;
        for (int d = 0; d < 3; ++d)
        {
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            F.sl.n1[d] = (SD > t1) ? F.g.n[t1] : 1;
            F.sl.n2[d] = (SD > t2) ? F.g.n[t2] : 1;
            for (int s2 = 0; s2 < 2; ++s2)
            {
                F.sl.off[d][s2] = off;
                if (d < SD) off += (long long)F.sl.n1[d] * F.sl.n2[d];
            }
        }
        F.sl.tot = off;
        F.pslab.alloc(off);
        F.pgc.alloc(off);
        F.eslab.alloc(off);
        F.sslab.alloc(off * nt);
        F.tslab.alloc(off * nt);
        Launch3 L = cell_launch(F.g.n);
        F.nblk = (long long)L.grid.x * L.grid.y * L.grid.z;
        F.part.alloc((size_t)F.nblk * kCellWaves * std::max(nt, 3));
        F.part_id.alloc((size_t)F.nblk * kCellWaves * 2 * nt);
        F.partn.alloc((size_t)1024 * nt);
        F.red.alloc(64);
        F.red2.alloc(64);
        F.nrm.alloc(kMaxComp);
        if (elecenergy_solve)
        {
            F.sbsnap.alloc((size_t)NUM_SPECIES * N);
            F.een_expl.alloc((size_t)N);
            F.part_e.alloc((size_t)F.nblk);
        }
        for (MGSolver<SD, Real>* M : {&F.pois, &F.spec})
        {
            const bool sp = (M == &F.spec);
            M->stream = stream;
            M->kt = &kt;
            M->kbase = sp ? K_SGSRB0 : K_PGSRB0;
            M->kfused = sp ? K_SPRE0 : K_PPRE0;
            M->fused_on = mg_fused != 0;
            M->coarse_graphs = mg_graphs != 0;
            M->graph_max_cells = (mg_graphs == 2) ? (1LL << 62) : (1LL << 22);
            M->fused_min_cells = (mg_fused == 2) ? 0 : mg_fused_min;
            M->build(F.g.n, F.g.dx, sp ? nt : 1, sp ? CK_CELLINV : CK_CONST, 0, 0);
            if (M->fused_on) M->enable_fine_fused();
            M->beta = sp ? timescale : 1.0;
            M->alpha_a = 0.0;
            M->bconst = 2.0 * (-1.0) * (-1.0) / ((-1.0) + (-1.0));
            M->maxorder = sp ? 3 : 2;
            M->nu1 = pre_smooth;
            M->nu2 = post_smooth;
            M->maxiter = maxiter;
            for (int c = 0; c < (sp ? nt : 1); ++c)
                for (int d = 0; d < SD; ++d)
                    for (int s2 = 0; s2 < 2; ++s2) M->side(c, d, s2) = SideSpec{BK_CF, 0.0, 1.0};
            M->upload_bc();
            if (sp) M->set_compmap(tsp);
            else M->set_compmap({0});
        }
    }

    // potential: A=1, a=0, B=1, b=-1 (harmonic face mean of -1 is -1), maxorder 2
    template <class MG>
    void cfg_pois(MG& P, int per_mask)
    {
        P.stream = stream;
        P.kt = &kt;
        P.kbase = K_PGSRB0;
        P.kfused = K_PPRE0;
        P.fused_on = mg_fused != 0;
        P.coarse_graphs = mg_graphs != 0;
        P.graph_max_cells = (mg_graphs == 2) ? (1LL << 62) : (1LL << 22);
        P.fused_min_cells = (mg_fused == 2) ? 0 : mg_fused_min;
        P.build(g.n, g.dx, 1, CK_CONST, max_coarsening, per_mask);
        P.rz = g.coord == 1;
        P.rz_lo = g.problo[0];
        P.alpha_a = 1.0 * 0.0;
        P.beta = 1.0;
        P.bconst = 2.0 * (-1.0) * (-1.0) / ((-1.0) + (-1.0));
        P.maxorder = 2;
        P.nu1 = pre_smooth;
        P.nu2 = post_smooth;
        P.maxiter = maxiter;
        P.direct_max_n = mg_direct_bottom;
        P.direct_finest = mg_direct_finest && max_level == 0;
        for (int d = 0; d < SD; ++d)
        {
            P.side(0, d, 0) = SideSpec{bc_kind(pot_bc_lo[d]), 1.0, 0.0};
            P.side(0, d, 1) = SideSpec{bc_kind(pot_bc_hi[d]), 1.0, 0.0};
        }
        P.upload_bc();
        P.set_compmap({0});
    }

    // photoionization: A=1, a = the case's get_photoion_acoeff (uniform),
    // B=1, face b = 1, maxorder 2 (Source/PhotoIonizationSolve.cpp:199-212)
    void cfg_photo(MGSolver<SD, Real>& P, int per_mask)
    {
        P.stream = stream;
        P.kt = &kt;
        P.kbase = K_PGSRB0;
        P.kfused = K_PPRE0;
        P.fused_on = mg_fused != 0;
        P.coarse_graphs = mg_graphs != 0;
        P.graph_max_cells = (mg_graphs == 2) ? (1LL << 62) : (1LL << 22);
        P.fused_min_cells = (mg_fused == 2) ? 0 : mg_fused_min;
        P.build(g.n, g.dx, 1, CK_CONST, max_coarsening, per_mask);
        P.rz = g.coord == 1;
        P.rz_lo = g.problo[0];
        P.alpha_a = 0.0;
        P.beta = 1.0;
        P.bconst = 1.0;
        P.maxorder = 2;
        P.nu1 = pre_smooth;
        P.nu2 = post_smooth;
        P.maxiter = maxiter;
        P.direct_max_n = mg_direct_bottom;
        P.direct_finest = mg_direct_finest && max_level == 0;
        for (int d = 0; d < SD; ++d)
        {
            P.side(0, d, 0) = SideSpec{bc_kind(photo_bc_lo[d]), 1.0, 0.0};
            P.side(0, d, 1) = SideSpec{bc_kind(photo_bc_hi[d]), 1.0, 0.0};
        }
        P.upload_bc();
        P.set_compmap({0});
    }

    // solve_three_term_photoionization (Source/Evolve.cpp:756-799): for each
    // term, solve_photoionization (Source/PhotoIonizationSolve.cpp:19-463) on
    // the Sborder state: coefficient and rhs from the case hooks, boundary
    // data from photoionization_bc folded in, zero initial guess.
    void photo_solve(Real time)
    {
        PhysParams P = phys(time, fixed_dt);
        const Launch3 L = cell_launch(g.n);
        const long long nw = (long long)L.grid.x * L.grid.y * L.grid.z * kCellWaves;
        last_photo_iters = 0;
        for (int t = 0; t < 3; ++t)
        {
            PhotoArgs A{P, d_pp.p, state.p, t, photo.f0.p, photo_part.p};
            {
                KScope ks(&kt, K_PRHS);
                photo_coef_kernel<<<L.grid, L.block, 0, stream>>>(A);
                finalize_max_kernel<<<3, 1024, 0, stream>>>(photo_part.p, nw, photo_red.p);
            }
            for (int d = 0; d < SD; ++d)
            {
                if (g.periodic[d]) continue;
                SidePair SP{};
                for (int s2 = 0; s2 < 2; ++s2)
                {
                    SideArgs& B = SP.s[s2];
                    B.P = P;
                    B.pp = d_pp.p;
                    B.state = state.p;
                    B.d = d;
                    B.side = s2;
                    B.n1 = phslab[d][s2].n1;
                    B.n2 = phslab[d][s2].n2;
                    B.nc = 1;
                    B.kind[0] = photo.side(0, d, s2).kind;
                    B.potential = 2;
                    B.vscale = 1.0;
                    B.maxorder = 2;
                    B.beta = 1.0;
                    B.bconst = 1.0;
                    B.slab_g = phslab[d][s2].g.p;
                    B.slab_a = phslab[d][s2].a.p;
                    B.slab_b = phslab[d][s2].b.p;
                    B.slab_f = phslab[d][s2].f.p;
                    B.gconst = phslab[d][s2].c.p;
                    B.rhs = photo.f0.p;
                }
                dim3 blk(64, 1, 1), grd((SP.s[0].n1 + 63) / 64, SP.s[0].n2, 2);
                KScope ks(&kt, K_SIDE);
                side_bc_kernel<<<grd, blk, 0, stream>>>(SP);
            }
            refresh_robin(photo, phslab, 1, true);
            const std::vector<double> m = read_bits(photo_red.p, 3);   // |rhs|, max a, max -a
            if (!(m[1] == -m[2]))
                throw BadArg("photoionization: get_photoion_acoeff must be uniform over the domain");
            photo.alpha_a = 1.0 * m[1];
            Real* sol = photo_sol.p + (size_t)t * g.N;
            fill_kernel<<<(unsigned)((g.N + 255) / 256), 256, 0, stream>>>(g.N, sol, 0.0);
            std::vector<double> b(1, m[0]);
            std::vector<int> its;
            std::vector<double> rn, r0;
            const int it = photo.solve(sol, g.N, b, reltol, abstol, &its, &rn, &r0);
            if (it < 0) throw std::runtime_error("MLMG failed to converge (photoionization)");
            last_photo_iters = std::max(last_photo_iters, it);
        }
        hip_check(hipGetLastError(), "photoionization");
    }

    // species: A=1, a=1/dt*time_scale, B=time_scale, harmonic face D, maxorder 3
    template <class MG>
    void cfg_spec(MG& P, int per_mask)
    {
        const int nt = (int)tsp.size();
        P.stream = stream;
        P.kt = &kt;
        P.kbase = K_SGSRB0;
        P.kfused = K_SPRE0;
        P.fused_on = mg_fused != 0;
        P.coarse_graphs = mg_graphs != 0;
        P.graph_max_cells = (mg_graphs == 2) ? (1LL << 62) : (1LL << 22);
        P.fused_min_cells = (mg_fused == 2) ? 0 : mg_fused_min;
        P.build(g.n, g.dx, nt, CK_CELLINV, max_coarsening, per_mask);
        P.rz = g.coord == 1;
        P.rz_lo = g.problo[0];
        P.beta = timescale;
        P.maxorder = 3;
        P.nu1 = pre_smooth;
        P.nu2 = post_smooth;
        P.maxiter = maxiter;
        P.direct_max_n = mg_direct_bottom;
        P.direct_finest = mg_direct_finest && max_level == 0;
        for (int c = 0; c < nt; ++c)
            for (int d = 0; d < SD; ++d)
            {
                P.side(c, d, 0) = SideSpec{bc_kind(bclo_of(tsp[c])[d]), 0.0, 1.0};
                P.side(c, d, 1) = SideSpec{bc_kind(bchi_of(tsp[c])[d]), 0.0, 1.0};
            }
        P.upload_bc();
        P.set_compmap(tsp);
    }

    void init_state()
    {
        amrex::GeometryData gd;
        for (int d = 0; d < SD; ++d)
        {
            gd.prob_lo[d] = g.problo[d];
            gd.prob_hi[d] = g.probhi[d];
            gd.dx[d] = g.dx[d];
            gd.is_periodic[d] = g.periodic[d];
        }
        amrex::IntVect lo, hi;
        for (int d = 0; d < SD; ++d)
        {
            lo[d] = 0;
            hi[d] = g.n[d] - 1;
        }
        gd.domain = amrex::Box(lo, hi);
        Launch3 L = cell_launch(g.n);
        init_state_kernel<<<L.grid, L.block, 0, stream>>>(state.p, g, gd, d_pp.p);
        hip_check(hipGetLastError(), "init_state");
        // each level from initdomaindata at its own resolution
        // (MakeNewLevelFromScratch, Source/GridWork.cpp:97-106)
        for (int l = 1; l <= (int)fine.size(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            amrex::GeometryData gf = gd;
            amrex::IntVect dlo, dhi;
            for (int d = 0; d < SD; ++d)
            {
                gf.dx[d] = F.g.dx[d];
                dlo[d] = 0;
                dhi[d] = g.n[d] * (1 << l) - 1;
            }
            gf.domain = amrex::Box(dlo, dhi);
            Launch3 Lf = cell_launch(F.g.n);
            init_state_kernel<<<Lf.grid, Lf.block, 0, stream>>>(F.state.p, F.g, gf, d_pp.p);
            hip_check(hipGetLastError(), "init_state (fine)");
        }
        cur_time = 0.0;
        istep = 0;
    }

    // ---------------------------------------------------------------------
    // Plotfiles and checkpoints (Source/Plot.cpp:32-225).

    // allvarnames: the case's species names, then the plasma variables
    // (Source/Vidyut.cpp:34-74, without the EB variables)
    std::vector<std::string> varnames() const
    {
        std::vector<std::string> v;
        CKSYMS_STR(v);
        v.resize(NUM_SPECIES);
        static const char* plasma[] = {"Electron_energy", "Electron_Temp", "Potential", "Efieldx", "Efieldy",
                                       "Efieldz", "Electron_Jheat", "Electron_inelasticHeat", "Electron_elasticHeat",
                                       "ReducedEF", "SurfaceCharge", "PhotoIon_Src", "cellmask", "Ecurden_X",
                                       "Ecurden_Y", "Ecurden_Z", "Icurden_X", "Icurden_Y", "Icurden_Z",
                                       "Dcurden_X", "Dcurden_Y", "Dcurden_Z"};
        static_assert(sizeof(plasma) / sizeof(plasma[0]) == NUM_PLASMAVARS, "plasma variable names");
        for (const char* n : plasma) v.push_back(n);
        return v;
    }

    // every level's state on the host, with its box and domain
    void host_levels(std::vector<std::vector<double>>& buf, std::vector<io::Level>& lv)
    {
        const int nl = 1 + (int)fine.size();
        buf.resize(nl);
        lv.resize(nl);
        hip_check(hipStreamSynchronize(stream), "sync");
        for (int l = 0; l < nl; ++l)
        {
            const Grid& G = lgrid(l);
            buf[l].resize((size_t)(G.N * NVAR));
            hip_check(hipMemcpy(buf[l].data(), lstate(l), sizeof(double) * G.N * NVAR, hipMemcpyDeviceToHost), "d2h");
            io::Level& L = lv[l];
            for (int d = 0; d < 3; ++d)
            {
                L.box.lo[d] = flo[l][d];
                L.box.hi[d] = flo[l][d] + fnn[l][d] - 1;
                L.domain.lo[d] = 0;
                L.domain.hi[d] = (d < SD) ? g.n[d] * (1 << l) - 1 : 0;
                L.dx[d] = G.dx[d];
            }
            L.data = buf[l].data();
        }
    }

    // Vidyut::WritePlotFile: every level, time t_new[0], level steps istep
    void write_plotfile(const std::string& dir)
    {
        std::vector<std::vector<double>> buf;
        std::vector<io::Level> lv;
        host_levels(buf, lv);
        io::write_plotfile(dir, varnames(), SD, cur_time, g.problo, g.probhi, lv,
                           std::vector<int>(lv.size(), istep), g.coord);
    }

    // Vidyut::WriteCheckpointFile
    void write_checkpoint(const std::string& dir)
    {
        std::vector<std::vector<double>> buf;
        std::vector<io::Level> lv;
        host_levels(buf, lv);
        const size_t nl = lv.size();
        io::write_checkpoint(dir, NVAR, SD, lv, std::vector<int>(nl, istep), std::vector<double>(nl, fixed_dt),
                             std::vector<double>(nl, cur_time));
    }

    // Vidyut::ReadCheckpointFile: the levels must be this context's (static
    // refinement from the inputs); the boxes of each level in the file,
    // however many, must cover the context's box of that level
    void read_checkpoint(const std::string& dir)
    {
        const io::CheckpointHeader H = io::read_checkpoint_header(dir, SD);
        const int nl = 1 + (int)fine.size();
        if (H.finest_level + 1 != nl)
            throw BadArg("checkpoint " + dir + " has " + std::to_string(H.finest_level + 1) +
                         " levels, the context " + std::to_string(nl) + " (amr.max_level)");
        std::vector<std::vector<double>> buf(nl);
        for (int l = 0; l < nl; ++l)
        {
            io::IBox b;
            for (int d = 0; d < 3; ++d)
            {
                b.lo[d] = flo[l][d];
                b.hi[d] = flo[l][d] + fnn[l][d] - 1;
            }
            buf[l].resize((size_t)(b.cells() * NVAR));
            io::read_vismf(dir + "/Level_" + std::to_string(l) + "/phi", NVAR, SD, b, buf[l].data());
        }
        hip_check(hipStreamSynchronize(stream), "sync");
        for (int l = 0; l < nl; ++l)
            hip_check(hipMemcpy(lstate(l), buf[l].data(), sizeof(double) * buf[l].size(), hipMemcpyHostToDevice),
                      "h2d");
        cur_time = H.t_new[0];
        istep = H.istep[0];
    }

    // InitData (Source/Vidyut.cpp:226-256): from scratch, or from the
    // checkpoint named by amr.restart; checkpoint 0 and the first plotfile
    // when output is on; then the output counters of Evolve
    // (Source/Evolve.cpp:20-40)
    void init_data()
    {
        if (restart_chkfile.empty())
        {
            init_state();
            if (chk_int > 0 || chk_time > 0.0) write_checkpoint(io::concatenate(chk_file, 0));
        } else
            read_checkpoint(restart_chkfile);
        if (plot_int > 0 || plot_time > 0.0)
            write_plotfile(io::concatenate(plot_file, plot_int > 0 ? (int)std::floor(double(istep) / plot_int) : 0));
        reset_output_counters();
    }
    void reset_output_counters()
    {
        plottime = chktime = 0.0;
        plotfilenum = plot_int_old > 0 ? (int)std::floor(double(istep) / plot_int_old) : 0;
        chkfilenum = chk_int_old > 0 ? (int)std::floor(double(istep) / chk_int_old) : 0;
        if (plot_time > 0.0) plotfilenum = (int)std::floor(cur_time / plot_time);
        if (chk_time > 0.0) chkfilenum = (int)std::floor(cur_time / chk_time);
    }
    // end of a coarse step (Source/Evolve.cpp:635-676)
    void output_after_step(Real dt)
    {
        if (plot_int <= 0 && plot_time <= 0.0 && chk_int <= 0 && chk_time <= 0.0) return;
        plottime += dt;
        chktime += dt;
        if (plot_time > 0.0)
        {
            if (plottime > plot_time)
            {
                write_plotfile(io::concatenate(plot_file, ++plotfilenum));
                plottime = 0.0;
            }
        } else if (plot_int > 0 && istep % plot_int == 0)
            write_plotfile(io::concatenate(plot_file, ++plotfilenum));
        if (chk_time > 0.0)
        {
            if (chktime > chk_time)
            {
                write_checkpoint(io::concatenate(chk_file, ++chkfilenum));
                chktime = 0.0;
            }
        } else if (chk_int > 0 && istep % chk_int == 0)
            write_checkpoint(io::concatenate(chk_file, ++chkfilenum));
    }

    void zero_red(int first, int n)
    {
        zero_u64_kernel<<<1, 64, 0, stream>>>(red.p + first, n);
    }
    std::vector<double> read_red(int first, int n)
    {
        std::vector<unsigned long long> h(n);
        d2h_sync(h.data(), red.p + first, sizeof(unsigned long long) * n, stream, "red d2h");
        std::vector<double> o(n);
        for (int i = 0; i < n; ++i) o[i] = bits_to_double(h[i]);
        require_finite(o.data(), n, "norm or maximum");
        return o;
    }

    // find_time_scales (maxima on the device; the three scales on the host)
    void time_scales(double out[3], const Real* st = nullptr)
    {
        TimeScaleArgs A{g, st ? st : state.p, gas_temperature, eidx, part.p};
        const int kz = tscale_kz(g.n[2]);
        int ce[3] = {g.n[0], g.n[1], (g.n[2] + kz - 1) / kz};
        Launch3 L = cell_launch(ce);
        const long long nb = (long long)L.grid.x * L.grid.y * L.grid.z;
        {
            KScope ks(&kt, K_TSCALE);
            time_scales_kernel<<<L.grid, L.block, 0, stream>>>(A, kz);
            finalize_max_kernel<<<3, 1024, 0, stream>>>(part.p, nb, red.p + 33);
        }
        auto m = read_red(33, 3);
        const double big = std::numeric_limits<double>::max();
        out[0] = out[1] = out[2] = big;
        if (m[0] > 0)
            for (int d = 0; d < SD; ++d) out[0] = std::min(out[0], g.dx[d] / m[0]);
        if (m[1] > 0)
            for (int d = 0; d < SD; ++d) out[1] = std::min(out[1], 0.5 * g.dx[d] * g.dx[d] / m[1] / SD);
        if (m[2] > 0) out[2] = EPS0 / ECHARGE / m[2];
    }

    // Robin coefficients of a side from the first boundary face (the
    // solver model takes them uniform over a side).
    template <class MG>
    bool refresh_robin(MG& mg, Slab (&sl)[3][2], int nc, bool user_hook)
    {
        bool changed = false;
        for (int c = 0; c < nc; ++c)
            for (int d = 0; d < SD; ++d)
                for (int s = 0; s < 2; ++s)
                {
                    SideSpec& sp = mg.side(c, d, s);
                    if (sp.kind != BK_ROBIN || !user_hook) continue;
                    const long long slab = (long long)sl[d][s].n1 * sl[d][s].n2;
                    double a = 0, b = 0;
                    hip_check(hipMemcpyAsync(&a, sl[d][s].a.p + c * slab, sizeof(double), hipMemcpyDeviceToHost, stream),
                              "robin");
                    hip_check(hipMemcpyAsync(&b, sl[d][s].b.p + c * slab, sizeof(double), hipMemcpyDeviceToHost, stream),
                              "robin");
                    hip_check(hipStreamSynchronize(stream), "robin sync");
                    if (a != sp.ra || b != sp.rb)
                    {
                        sp.ra = a;
                        sp.rb = b;
                        changed = true;
                    }
                }
        if (changed) mg.upload_bc();
        return changed;
    }

    // integrated currents handed to the bc hooks,
    // Source/PotentialSolve.cpp:42-52 and Source/ScalarSolve.cpp:404-414
    void hook_currents(SideArgs& A) const
    {
        if (!track_integrated_currents) return;
        for (int i = 0; i < ncurrent_locs; ++i)
        {
            A.ic[i] = int_cond[i];
            A.ia[i] = int_area[i];
            A.is[i] = current_loc_surfaces[i];
        }
    }

    // Source/Evolve.cpp:582-593: current densities of the new state, then
    // the integrated currents (read back: the bc hooks of the next solve
    // receive them by value).
    void track_currents(Real dt)
    {
        CurrentArgs C{g, state.p, eold.p, gas_temperature, cur_time, dt, eidx};
        Launch3 L = cell_launch(g.n);
        {
            KScope ks(&kt, K_OTHER);
            current_den_kernel<<<L.grid, L.block, 0, stream>>>(C);
        }
        hip_check(hipGetLastError(), "current_den");
        if (!track_integrated_currents || ncurrent_locs == 0) return;
        IntCurrentArgs I{};
        I.g = g;
        I.state = state.p;
        I.nloc = ncurrent_locs;
        for (int q = 0; q < ncurrent_locs; ++q) I.surf[q] = current_loc_surfaces[q];
        I.partials = cur_part.p;
        {
            KScope ks(&kt, K_OTHER);
            integrated_currents_kernel<<<L.grid, L.block, 0, stream>>>(I);
            finalize_sum_kernel<<<3 * ncurrent_locs, 1024, 0, stream>>>(cur_part.p, nblk, cur_dev.p, 0);
        }
        double h[3 * MAX_CURRENT_LOCS];
        d2h_sync(h, cur_dev.p, sizeof(double) * 3 * ncurrent_locs, stream, "currents d2h");
        for (int q = 0; q < ncurrent_locs; ++q)
        {
            int_cond[q] = h[3 * q];
            int_disp[q] = h[3 * q + 1];
            int_area[q] = h[3 * q + 2];
        }
    }

    // solve_potential, Source/PotentialSolve.cpp:19-554
    int solve_potential(Real time, const Real* sb_old, vd_solve_info* info)
    {
        const double bn = pot_rhs0(time, sb_old);
        std::vector<double> b(1, bn);
        std::vector<int> its;
        std::vector<double> rn, r0;
        int it = mixed ? pois32.solve(phi_sol.p, g.N, b, reltol, abstol, &its, &rn, &r0)
                       : pois.solve(phi_sol.p, g.N, b, reltol, abstol, &its, &rn, &r0);
        last_pot_r0 = r0[0];
        last_pot_r = rn[0];
        last_pot_b = b[0];
        if (info)
        {
            info->iters = it;
            info->bnorm = b[0];
            info->resid0 = r0[0];
            info->resid = rn[0];
        }
        if (it < 0) throw std::runtime_error("MLMG failed to converge (potential)");
        last_pot_iters = it;
        pot_efield0(time);
        return it;
    }

    // level-0 Poisson rhs (time-n charge density) with the domain boundary
    // data folded in; zero initial guess; returns |rhs|_inf
    double pot_rhs0(Real time, const Real* sb_old)
    {
        PhysParams P = phys(time, fixed_dt);
        PoissonRhsArgs R{P, d_pp.p, const_cast<Real*>(sb_old), 1.0 / potscale, do_spacechrg, pf0(), part.p};
        Launch3 L = cell_launch(g.n);
        {
            KScope ks(&kt, K_PRHS);
            poisson_rhs_kernel<<<L.grid, L.block, 0, stream>>>(R);
            finalize_max_kernel<<<1, 1024, 0, stream>>>(part.p, nblk, red.p + 32);
        }
        for (int d = 0; d < SD; ++d)
        {
            if (g.periodic[d]) continue;
            SidePair SP{};
            for (int s = 0; s < 2; ++s)
            {
                SideArgs& A = SP.s[s];
                A.P = P;
                A.pp = d_pp.p;
                A.state = state.p;
                A.d = d;
                A.side = s;
                A.n1 = pslab[d][s].n1;
                A.n2 = pslab[d][s].n2;
                A.nc = 1;
                A.kind[0] = pside(0, d, s).kind;
                A.potential = 1;
                A.app_voltage = applied_voltage(time, s ? 1 : -1, voltage_profile, vamp_lo[d], vamp_hi[d], vfreq, vdur,
                                                vcen);
                A.vscale = 1.0 / potscale;
                A.maxorder = 2;
                A.beta = 1.0;
                A.bconst = -1.0;
                A.slab_g = pslab[d][s].g.p;
                A.slab_a = pslab[d][s].a.p;
                A.slab_b = pslab[d][s].b.p;
                A.slab_f = pslab[d][s].f.p;
                A.gconst = pslab[d][s].c.p;
                A.rhs = pf0();
                hook_currents(A);
            }
            dim3 blk(64, 1, 1), grd((SP.s[0].n1 + 63) / 64, SP.s[0].n2, 2);
            KScope ks(&kt, K_SIDE);
            side_bc_kernel<<<grd, blk, 0, stream>>>(SP);
        }
        if (mixed) refresh_robin(pois32, pslab, 1, user_defined_potential != 0);
        else refresh_robin(pois, pslab, 1, user_defined_potential != 0);
        auto b = read_red(32, 1);
        fill_kernel<<<(unsigned)((g.N + 255) / 256), 256, 0, stream>>>(g.N, phi_sol.p, 0.0);
        return b[0];
    }

    // level-0 E field, E/N and potential from the solution
    void pot_efield0(Real time)
    {
        PhysParams P = phys(time, fixed_dt);
        Launch3 L = cell_launch(g.n);
        EFieldArgs E{};
        E.P = P;
        E.sol = phi_sol.p;
        E.state = state.p;
        E.potscale = potscale;
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s)
            {
                E.gconst[d][s] = (d < SD) ? pslab[d][s].c.p : nullptr;
                int kd = (d < SD) ? pside(0, d, s).kind : BK_NEU;
                double dep = 1.0;
                if (kd == BK_DIR) dep = -1.0;
                else if (kd == BK_ROBIN)
                {
                    const SideSpec& sp = pside(0, d, s);
                    double h = g.dx[d];
                    dep = -(0.5 * sp.ra - sp.rb / h) / (0.5 * sp.ra + sp.rb / h);
                }
                E.dep0[d][s] = dep;
            }
        E.eidx = eidx;
        {
            KScope ks(&kt, K_EFIELD);
            efield_kernel<<<L.grid, L.block, 0, stream>>>(E);
        }
        hip_check(hipGetLastError(), "efield");
    }

    template <int HYP, int SCHEME, class TF>
    void launch_flux_t(const FluxArgs& base, const Grid& G, void* const* Fd)
    {
        for (int d = 0; d < SD; ++d)
        {
            FluxArgs A = base;
            A.F = Fd[d];
            A.nf = (long long)(G.n[0] + (d == 0)) * (G.n[1] + (d == 1)) * (G.n[2] + (d == 2));
            int e[3] = {G.n[0] + (d == 0), G.n[1] + (d == 1), G.n[2] + (d == 2)};
            Launch3 L;
            L.block = dim3(256, 1, 1);
            L.grid = dim3((unsigned)(((long long)e[0] * e[1] + 255) / 256), 1, e[2]);
            KScope ks(&kt, K_FLUX);
            // WENO5 faces with default velocities on a level with
            // index-mapped ghosts: the marching kernels (y, z: lanes along
            // x, FP64 or the mixed form's FP32; x: rows staged through LDS,
            // FP64 only -- with FP32 faces its fixed cost of staging FP64
            // cells at two wavefronts per SIMD made it slower than the
            // per-face kernel, 1.33 against 0.94 ms at 256^3)
            const bool march = HYP == 5 && !A.P.user_vel && A.sg == nullptr && flux_march;
            if constexpr (HYP == 5 && std::is_same<TF, double>::value)
                if (d == 0 && march)
                {
                    dim3 grid((unsigned)(A.nt * ((G.n[1] * G.n[2] + 63) / 64)), 1, 1);
                    face_flux_march_x_kernel<SCHEME, TF><<<grid, 64, 0, stream>>>(A);
                    continue;
                }
            if (d == 0) face_flux_kernel<0, HYP, SCHEME, TF><<<L.grid, L.block, 0, stream>>>(A);
#if AMREX_SPACEDIM > 1
            if constexpr (HYP == 5)
                if (d > 0 && march)
                {
                    const int nf = G.n[d] + 1;
                    // columns in chunks of about 86 faces (3 at 256^3): each
                    // chunk pays a six-cell warm-up, and more chunks only add
                    // waves (32-face chunks measured 1-2% slower)
                    const int nch0 = (nf + 85) / 86, chunk = (nf + nch0 - 1) / nch0, nch = (nf + chunk - 1) / chunk;
                    dim3 grid((unsigned)(A.nt * ((G.n[0] + 63) / 64)), (unsigned)G.n[d == 1 ? 2 : 1], (unsigned)nch);
                    if (d == 1) face_flux_march_kernel<1, SCHEME, TF><<<grid, 64, 0, stream>>>(A, chunk);
#if AMREX_SPACEDIM > 2
                    else face_flux_march_kernel<2, SCHEME, TF><<<grid, 64, 0, stream>>>(A, chunk);
#endif
                    continue;
                }
            if (d == 1) face_flux_kernel<1, HYP, SCHEME, TF><<<L.grid, L.block, 0, stream>>>(A);
#endif
#if AMREX_SPACEDIM > 2
            if (d == 2) face_flux_kernel<2, HYP, SCHEME, TF><<<L.grid, L.block, 0, stream>>>(A);
#endif
        }
    }

    // face fluxes of all directions on one level in precision TF
    template <class TF>
    void launch_flux_any(const FluxArgs& A, const Grid& G, void* const* Fd)
    {
        if (hyp_order == 1) launch_flux_t<1, 1, TF>(A, G, Fd);
        else if (hyp_order == 2) launch_flux_t<2, 1, TF>(A, G, Fd);
        else if (weno_scheme == 1) launch_flux_t<5, 1, TF>(A, G, Fd);
        else if (weno_scheme == 2) launch_flux_t<5, 2, TF>(A, G, Fd);
        else launch_flux_t<5, 3, TF>(A, G, Fd);
    }

    // FP64 fluxes of a fine level (zero without transport)
    void launch_flux_level(FluxArgs A, const Grid& G, DevBuf<Real>* fl)
    {
        A.dxmin = G.dx[0];
        for (int d = 1; d < SD; ++d) A.dxmin = std::min(A.dxmin, G.dx[d]);
        void* Fd[3] = {fl[0].p, SD > 1 ? fl[1].p : nullptr, SD > 2 ? fl[2].p : nullptr};
        if (do_transport)
            launch_flux_any<double>(A, G, Fd);
        else
            for (int d = 0; d < SD; ++d)
                fill_kernel<<<(unsigned)((fl[d].n + 255) / 256), 256, 0, stream>>>(fl[d].n, fl[d].p, 0.0);
    }

    // Flux arguments of the transported species (compute_scalar_transport_flux,
    // Source/ScalarSolve.cpp:218-329) on level 0.
    FluxArgs flux_args(Real time, Real dt)
    {
        const int nt = (int)tsp.size();
        FluxArgs F{};
        F.P = phys(time, dt);
        F.pp = d_pp.p;
        F.state = state.p;
        F.nt = nt;
        for (int c = 0; c < nt; ++c)
        {
            F.tsp[c] = tsp[c];
            for (int d = 0; d < 3; ++d)
            {
                F.bclo[c][d] = (d < SD) ? bclo_of(tsp[c])[d] : 0;
                F.bchi[c][d] = (d < SD) ? bchi_of(tsp[c])[d] : 0;
            }
        }
        F.weno_scheme = weno_scheme;
        F.dxmin = g.dx[0];
        for (int d = 1; d < SD; ++d) F.dxmin = std::min(F.dxmin, g.dx[d]);
        return F;
    }

    void species_flux0(Real time, Real dt)
    {
        FluxArgs F = flux_args(time, dt);
        if (do_transport)
        {
            if (mixed)
            {
                void* Fd[3] = {flux32[0].p, flux32[1].p, flux32[2].p};
                launch_flux_any<float>(F, g, Fd);
            } else
            {
                void* Fd[3] = {flux[0].p, flux[1].p, flux[2].p};
                launch_flux_any<double>(F, g, Fd);
            }
        } else
        {
            for (int d = 0; d < SD; ++d)
            {
                if (mixed)
                    fill_float_kernel<<<(unsigned)((flux32[d].n + 255) / 256), 256, 0, stream>>>(flux32[d].n, flux32[d].p);
                else
                    fill_kernel<<<(unsigned)((flux[d].n + 255) / 256), 256, 0, stream>>>(flux[d].n, flux[d].p, 0.0);
            }
        }
    }

    // Cell-kernel arguments (divergence + reactions + implicit rhs) on level 0.
    CellArgs cell_args(Real time, Real dt, const Real* sb_old)
    {
        const int nt = (int)tsp.size();
        CellArgs C{};
        C.P = phys(time, dt);
        C.pp = d_pp.p;
        C.state = state.p;
        C.state_old = sb_old;
        C.nt = nt;
        for (int c = 0; c < nt; ++c)
        {
            C.tsp[c] = tsp[c];
            C.rscale[c] = timescale / comp_scale(tsp[c]);
        }
        C.defer_c = een_comp();
        C.defer_out = een_expl.p;
        C.write_eih = elecenergy_solve;
        if (photo_in_step)
        {
            for (int t = 0; t < 3; ++t) C.photo[t] = photo_sol.p + (size_t)t * g.N;
            C.photo_sp = photoion_sp;
        }
        for (int d = 0; d < 3; ++d)
        {
            C.F[d] = (d < SD) ? (mixed ? (const void*)flux32[d].p : (const void*)flux[d].p) : nullptr;
            C.nf[d] = (d < SD) ? (long long)(mixed ? flux32[d].n : flux[d].n) / nt : 0;
        }
        C.rhs = sf0();
        C.iD = siD0();
        C.rxn_out = do_bg_reactions ? rxn.p : nullptr;
        C.partials = part.p;
        C.partials_id = part_id.p;
        return C;
    }

    // single level: the components whose 1/D the last solve found uniform
    // (predicted uniform again, so the cell pass does not store it) and
    // those the last cell pass skipped
    unsigned iD_pred = 0u, iD_skipped = 0u;
    void species_cell0(Real time, Real dt, const Real* sb_old, Real* expl_out, Real* rxn_out)
    {
        const int nt = (int)tsp.size();
        CellArgs C = cell_args(time, dt, sb_old);
        C.expl_out = expl_out;
        if (rxn_out) C.rxn_out = rxn_out;
        // the electron energy's 1/D is also read by the Joule heating
        const int ce = een_comp();
        C.skip_iD = (nlev() == 1) ? (iD_pred & ~(ce >= 0 ? 1u << ce : 0u)) : 0u;
        iD_skipped = C.skip_iD;
        Launch3 L = cell_launch(g.n);
        {
            KScope ks(&kt, K_CELL);
            if (mixed) cell_source_kernel<float><<<L.grid, L.block, 0, stream>>>(C);
            else cell_source_kernel<double><<<L.grid, L.block, 0, stream>>>(C);
            const long long nw = nblk * kCellWaves;
            // the per-wave partials (one per 64 cells) in two stages: a
            // single workgroup per row spent about 50 us at 256^3
            max_split3_kernel<<<dim3(kMaxSplit, 3 * nt), 256, 0, stream>>>(part.p, part_id.p, nw, nt, part_split.p);
            finalize_max3_kernel<<<3 * nt, 1024, 0, stream>>>(part_split.p, part_split.p + (size_t)nt * kMaxSplit,
                                                             kMaxSplit, nt, red.p, red2.p, red2.p + 32);
        }
        hip_check(hipGetLastError(), "cell_source");
    }

    // boundary hooks of the level-0 domain sides folded into the rhs
    void species_side0(Real time, Real dt)
    {
        PhysParams P = phys(time, dt);
        const int nt = (int)tsp.size();
        for (int d = 0; d < SD; ++d)
        {
            if (g.periodic[d]) continue;
            SidePair SP{};
            for (int s = 0; s < 2; ++s)
            {
                SideArgs& A = SP.s[s];
                A.P = P;
                A.pp = d_pp.p;
                A.state = state.p;
                A.d = d;
                A.side = s;
                A.n1 = sslab[d][s].n1;
                A.n2 = sslab[d][s].n2;
                A.nc = nt;
                for (int c = 0; c < nt; ++c)
                {
                    A.tsp[c] = tsp[c];
                    A.kind[c] = sside(c, d, s).kind;
                    A.cscale[c] = 1.0 / comp_scale(tsp[c]);
                }
                A.nofold = (een_comp() >= 0) ? (1u << een_comp()) : 0u;
                A.potential = 0;
                A.maxorder = 3;
                A.beta = timescale;
                A.iD = siD0();
                A.skip_iD = iD_skipped;
                A.slab_g = sslab[d][s].g.p;
                A.slab_a = sslab[d][s].a.p;
                A.slab_b = sslab[d][s].b.p;
                A.slab_f = sslab[d][s].f.p;
                A.gconst = sslab[d][s].c.p;
                A.rhs = sf0();
                hook_currents(A);
            }
            dim3 blk(64, 1, 1), grd((SP.s[0].n1 + 63) / 64, SP.s[0].n2, 2);
            KScope ks(&kt, K_SIDE);
            side_bc_kernel<<<grd, blk, 0, stream>>>(SP);
        }
        hip_check(hipGetLastError(), "side_bc");
    }

    // Explicit sources, implicit rhs, coefficients and boundary data for all
    // transported species (update_explsrc + update_rxnsrc + the input side of
    // implicit_solve_scalar).  expl_out/rxn_out: optional device arrays.
    void species_inputs(Real time, Real dt, const Real* sb_old, Real* expl_out, Real* rxn_out)
    {
        species_flux0(time, dt);
        species_cell0(time, dt, sb_old, expl_out, rxn_out);
        species_side0(time, dt);
    }

    // ---------------------------------------------------------------------
    // Operator entry points: the MLMG operators of the potential and species
    // solves on caller device buffers, boundary data given per side.
    DevBuf<double> op_part, op_gc[3][2];
    DevBuf<unsigned long long> op_nrm;

    void check_op_bc(const vd_bc& bc) const
    {
        for (int d = 0; d < SD; ++d)
            for (int s = 0; s < 2; ++s)
            {
                const int k = bc.kind[d][s];
                if (k < 0 || k > 4 || ((k == PERBC) != (g.periodic[d] != 0)))
                    throw BadArg("bc kind " + std::to_string(k) + " in direction " + std::to_string(d) +
                                 " (0 exactly on periodic directions, 1-4 elsewhere)");
                if (k == ROBINBC && !(0.5 * bc.robin_a[d][s] + bc.robin_b[d][s] / g.dx[d] != 0.0))
                    throw BadArg("Robin side with a/2 + b/h = 0");
            }
    }

    // Solve  alpha u - beta div(b grad u) = rhs  for ncomp components on the
    // solver M (Poisson: a = 0, b = bconst; species: b = harmonic face mean
    // of bcell).  sol: initial guess in, solution out.  grad[d]: scale *
    // du/dx_d on the faces of direction d (nullable).
    int op_solve(MGSolver<SD, Real>& M, bool cell_coef, int ncomp, const double* rhs, double* sol, double alpha,
                 double beta, const double* bcell, const vd_bc& bc, double rtol, double atol, double* const grad[3],
                 double scale, vd_solve_info* info)
    {
        if (ncomp < 1 || ncomp > M.nc)
            throw BadArg("ncomp must be in [1, " + std::to_string(M.nc) + "]");
        if (!rhs || !sol || (cell_coef && !bcell)) throw BadArg("rhs, sol and the coefficients are required");
        if (!(rtol >= 0.0) || !(atol >= 0.0)) throw BadArg("tolerances must be non-negative");
        check_op_bc(bc);
        const long long N = g.N;
        // the context's own setup, restored below
        std::vector<SideSpec> saved;
        for (int c = 0; c < M.nc; ++c)
            for (int d = 0; d < 3; ++d)
                for (int s = 0; s < 2; ++s) saved.push_back(M.side(c, d, s));
        const double saved_alpha = M.alpha_a, saved_beta = M.beta;
        struct Restore
        {
            vd_ctx* x;
            MGSolver<SD, Real>& M;
            std::vector<SideSpec>& saved;
            double a, b;
            std::vector<double> sb;
            ~Restore()
            {
                size_t q = 0;
                for (int c = 0; c < M.nc; ++c)
                    for (int d = 0; d < 3; ++d)
                        for (int s = 0; s < 2; ++s) M.side(c, d, s) = saved[q++];
                M.alpha_a = a;
                M.beta = b;
                M.smooth_bmax.swap(sb);
                try
                {
                    M.upload_bc();
                    if (&M == &x->spec) x->spec.set_compmap(x->tsp);
                } catch (...)
                {
                }
            }
        } restore{this, M, saved, saved_alpha, saved_beta, {}};
        // the operator entry points iterate V-cycles, as MLMG does
        restore.sb.swap(M.smooth_bmax);
        for (int c = 0; c < M.nc; ++c)
            for (int d = 0; d < SD; ++d)
                for (int s = 0; s < 2; ++s)
                    M.side(c, d, s) = SideSpec{bc_kind(bc.kind[d][s]), bc.robin_a[d][s], bc.robin_b[d][s]};
        M.upload_bc();
        M.alpha_a = alpha;
        M.beta = beta;
        // components: identity map into sol (stride N); unused slots point at 0
        std::vector<int> idm(M.nc, 0);
        for (int c = 0; c < ncomp; ++c) idm[c] = c;
        M.set_compmap(idm);
        if (cell_coef)
        {
            inv_coef_kernel<<<(unsigned)((N * ncomp + 255) / 256), 256, 0, stream>>>(N * ncomp, bcell, M.iD0.p);
            M.set_uniform(0u, std::vector<double>(M.nc, 0.0));
            M.coarsen_coeffs();
        }
        // |rhs| per component (unfolded), then the rhs with the side data
        if (op_part.n < (size_t)1024 * kMaxComp) op_part.alloc((size_t)1024 * kMaxComp);
        if (op_nrm.n < (size_t)kMaxComp) op_nrm.alloc(kMaxComp);
        const unsigned mask = (ncomp >= 32) ? 0xffffffffu : ((1u << ncomp) - 1u);
        {
            dim3 grd(1024, ncomp, 1);
            amr_absmax_kernel<<<grd, 256, 0, stream>>>(rhs, N, mask, op_part.p, 1024);
            finalize_max_kernel<<<ncomp, 1024, 0, stream>>>(op_part.p, 1024, op_nrm.p);
        }
        std::vector<double> bn = read_bits(op_nrm.p, ncomp);
        bn.resize(M.nc, 0.0);
        hip_check(hipMemcpyAsync(M.f0.p, rhs, sizeof(double) * N * ncomp, hipMemcpyDeviceToDevice, stream), "rhs");
        for (int d = 0; d < SD; ++d)
        {
            if (g.periodic[d]) continue;
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            const int n1 = (SD > t1) ? g.n[t1] : 1, n2 = (SD > t2) ? g.n[t2] : 1;
            for (int s = 0; s < 2; ++s)
            {
                DevBuf<double>& gcb = op_gc[d][s];
                if (gcb.n < (size_t)n1 * n2 * M.nc) gcb.alloc((size_t)n1 * n2 * M.nc);
                BcFoldArgs F{};
                F.g = g;
                F.d = d;
                F.side = s;
                F.nc = ncomp;
                F.kind = bc_kind(bc.kind[d][s]);
                F.maxorder = M.maxorder;
                F.value = bc.value[d][s];
                F.ra = bc.robin_a[d][s];
                F.rb = bc.robin_b[d][s];
                F.beta = beta;
                F.bconst = M.bconst;
                F.iD = cell_coef ? M.iD0.p : nullptr;
                F.rhs = M.f0.p;
                F.gout = gcb.p;
                F.n1 = n1;
                F.n2 = n2;
                dim3 grd((unsigned)((n1 + 63) / 64), (unsigned)n2, (unsigned)ncomp);
                bc_fold_kernel<<<grd, 64, 0, stream>>>(F);
            }
        }
        hip_check(hipGetLastError(), "operator rhs");
        std::vector<int> its;
        std::vector<double> rn, r0;
        const int it = solve_masked(M, mask, bn, its, rn, r0, sol, rtol, atol);
        if (info)
        {
            info->iters = it;
            info->bnorm = info->resid0 = info->resid = 0.0;
            for (int c = 0; c < ncomp; ++c)
            {
                info->bnorm = std::max(info->bnorm, bn[c]);
                info->resid0 = std::max(info->resid0, r0[c]);
                info->resid = std::max(info->resid, rn[c]);
            }
        }
        if (it < 0) throw std::runtime_error("MLMG failed to converge (operator solve)");
        for (int d = 0; d < SD; ++d)
        {
            if (!grad || !grad[d]) continue;
            FaceGradArgs A{};
            A.g = g;
            A.d = d;
            A.nc = ncomp;
            A.u = sol;
            A.scale = scale;
            A.out = grad[d];
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            A.n1 = (SD > t1) ? g.n[t1] : 1;
            A.n2 = (SD > t2) ? g.n[t2] : 1;
            for (int s = 0; s < 2; ++s)
            {
                A.gc[s] = g.periodic[d] ? nullptr : op_gc[d][s].p;
                const int k = bc_kind(bc.kind[d][s]);
                double d0 = 1.0, d1 = 0.0;
                if (k == BK_DIR)
                {
                    const int m = std::min(g.n[d] + 1, M.maxorder);
                    d0 = (m >= 3) ? -2.0 : -1.0;
                    d1 = (m >= 3) ? 1.0 / 3.0 : 0.0;
                } else if (k == BK_ROBIN)
                {
                    const double h = g.dx[d], a = bc.robin_a[d][s], b = bc.robin_b[d][s];
                    d0 = -(0.5 * a - b / h) / (0.5 * a + b / h);
                }
                A.dep0[s] = d0;
                A.dep1[s] = d1;
            }
            const int e[3] = {g.n[0] + (d == 0), g.n[1] + (d == 1), g.n[2] + (d == 2)};
            dim3 grd((unsigned)((e[0] + 255) / 256), (unsigned)e[1], (unsigned)(e[2] * ncomp));
            face_grad_kernel<<<grd, 256, 0, stream>>>(A);
        }
        hip_check(hipGetLastError(), "operator gradients");
        return it;
    }

    // update_explsrc_at_all_levels for species [start, start + num) with the
    // side codes bclo/bhi (Source/ScalarSolve.cpp:81-154) on a caller's
    // state st (device, this context's grid and layout): face fluxes into
    // Fd (num components each, FP64), then compute_dsdt with the reaction
    // source rxn (nullptr: evaluated in the same pass) into expl.
    void explicit_rhs(const Real* st, const Real* rxn, int start, int num, const int* bclo, const int* bchi,
                      Real time, Real dt, Real* expl, Real* const Fd[3])
    {
        FluxArgs F{};
        F.P = phys(time, dt);
        F.pp = d_pp.p;
        F.state = const_cast<Real*>(st);
        F.nt = num;
        for (int c = 0; c < num; ++c)
        {
            F.tsp[c] = start + c;
            for (int d = 0; d < 3; ++d)
            {
                F.bclo[c][d] = (d < SD) ? bclo[d] : 0;
                F.bchi[c][d] = (d < SD) ? bchi[d] : 0;
            }
        }
        F.weno_scheme = weno_scheme;
        F.dxmin = g.dx[0];
        for (int d = 1; d < SD; ++d) F.dxmin = std::min(F.dxmin, g.dx[d]);
        DsdtArgs D{};
        for (int d = 0; d < SD; ++d)
        {
            D.nf[d] = (long long)(g.n[0] + (d == 0)) * (g.n[1] + (d == 1)) * (g.n[2] + (d == 2));
            D.F[d] = Fd[d];
        }
        if (do_transport)
        {
            void* Fv[3] = {Fd[0], SD > 1 ? Fd[1] : nullptr, SD > 2 ? Fd[2] : nullptr};
            launch_flux_any<double>(F, g, Fv);
        } else
            for (int d = 0; d < SD; ++d)
                fill_kernel<<<(unsigned)((D.nf[d] * num + 255) / 256), 256, 0, stream>>>(D.nf[d] * num, Fd[d], 0.0);
        D.P = F.P;
        D.pp = d_pp.p;
        D.state = F.state;
        D.rxn = rxn;
        D.start = start;
        D.num = num;
        D.expl = expl;
        Launch3 L = cell_launch(g.n);
        {
            KScope ks(&kt, K_CELL);
            dsdt_kernel<<<L.grid, L.block, 0, stream>>>(D);
        }
        hip_check(hipGetLastError(), "explicit_rhs");
    }

    // update_rxnsrc_at_all_levels on a caller's state (device pointers)
    void reaction_src(const Real* st, Real time, Real* rxn)
    {
        RxnArgs R{phys(time, fixed_dt), d_pp.p, const_cast<Real*>(st), rxn};
        Launch3 L = cell_launch(g.n);
        {
            KScope ks(&kt, K_CELL);
            reaction_source_kernel<<<L.grid, L.block, 0, stream>>>(R);
        }
        hip_check(hipGetLastError(), "reaction_src");
    }

    // checks of a species range and its side codes (the reference's
    // Vector<int> bc_lo/bc_hi of the species class)
    void check_range(int start, int num, const int* bclo, const int* bchi) const
    {
        if (start < 0 || num < 1 || start + num > NUM_SPECIES + 1 || num > kMaxComp)
            throw BadArg("species range outside [0, NUM_SPECIES] or longer than " + std::to_string(kMaxComp));
        if (!bclo || !bchi) throw BadArg("bc_lo and bc_hi are required");
        for (int d = 0; d < SD; ++d)
            for (int c : {bclo[d], bchi[d]})
                if (c < 0 || c > 4 || ((c == PERBC) != (g.periodic[d] != 0)))
                    throw BadArg("side code " + std::to_string(c) + " in direction " + std::to_string(d) +
                                 " (0 exactly on periodic directions, 1-4 elsewhere)");
    }
    long long faces_of(int d) const
    {
        return (long long)(g.n[0] + (d == 0)) * (g.n[1] + (d == 1)) * (g.n[2] + (d == 2));
    }

    // Per-step setup of the species operator: Robin data, a = 1/dt, the
    // uniform-coefficient components and the coarse coefficients.
    template <class MG>
    void spec_prepare_t(MG& spec, Real dt)
    {
        const int nt = (int)tsp.size();
        refresh_robin(spec, sslab, nt, user_defined_species != 0);
        spec.alpha_a = 1.0 * ((1.0 / dt) * timescale);
        // components with a spatially uniform coefficient (max 1/D == min 1/D)
        // use the constant-coefficient kernels with the same face value
        // 2/(1/D + 1/D) the harmonic-mean kernels would form
        {
            std::vector<unsigned long long> h(64);
            d2h_sync(h.data(), red2.p, sizeof(unsigned long long) * 64, stream, "red2 d2h");
            unsigned um = 0;
            std::vector<double> vals(nt, 0.0), bmax(nt, 0.0);
            for (int c = 0; c < nt; ++c)
            {
                const double mx = bits_to_double(h[c]), mn = -bits_to_double(h[32 + c]);
                if (mx == mn && !(mx != mx))
                {
                    um |= 1u << c;
                    vals[c] = 2.0 / (mx + mx);
                }
                // the largest face coefficient: a harmonic mean of two cell
                // values is at most the larger one, 1/min(1/D)
                bmax[c] = (mn > 0.0) ? 1.0 / mn : std::numeric_limits<double>::infinity();
            }
            spec.set_uniform(um, vals);
            if (mg_smooth_iters && max_level == 0) spec.smooth_bmax = bmax;
            else spec.smooth_bmax.clear();
            if ((void*)&spec == (void*)&this->spec)
            {
                // a component predicted uniform that is not: store its 1/D
                // now, from the state the cell pass read
                if (const unsigned miss = iD_skipped & ~um)
                {
                    CellArgs C = cell_args(0.0, dt, state.p);
                    Launch3 L = cell_launch(g.n);
                    KScope ks(&kt, K_CELL);
                    idiff_fill_kernel<<<L.grid, L.block, 0, stream>>>(C, miss);
                }
                iD_pred = um;
                iD_skipped = 0u;
            }
        }
        spec.coarsen_coeffs();
    }
    void spec_prepare(Real dt) { spec_prepare_t(spec, dt); }

    // Batched implicit species solve over the components in 'mask'
    // (implicit_solve_scalar per species, Source/ScalarSolve.cpp:380-1000).
    int species_solve(Real dt, unsigned mask, vd_solve_info* info)
    {
        return species_solve_t(spec, dt, mask, info);
    }

    // 'between' runs after the linear solve, before the solution is unscaled
    // and clamped (the electron energy reads the electron solution there).
    // 'prepared': the coefficients were set up by an earlier call of this step.
    template <class MG>
    int species_solve_t(MG& spec, Real dt, unsigned mask, vd_solve_info* info,
                        const std::function<void()>& between = nullptr, bool prepared = false,
                        const std::vector<double>* bnorm_in = nullptr)
    {
        const int nt = (int)tsp.size();
        if (!prepared) spec_prepare_t(spec, dt);
        auto b = bnorm_in ? *bnorm_in : read_red(0, nt);
        bool unit_scale = true;
        for (int c = 0; c < nt; ++c)
            if (((mask >> c) & 1u) && comp_scale(tsp[c]) != 1.0) unit_scale = false;
        PostArgs S{};
        S.g = g;
        S.state = state.p;
        S.nt = nt;
        S.active = mask;
        for (int c = 0; c < nt; ++c)
        {
            S.tsp[c] = tsp[c];
            S.scale[c] = 1.0 / comp_scale(tsp[c]);
            // the energy is never clamped (Source/ScalarSolve.cpp:910-943)
            S.floor[c] = (tsp[c] == EEN_ID) ? -std::numeric_limits<double>::max()
                                            : ((tsp[c] == eidx) ? min_electron_density : min_species_density);
        }
        const unsigned blocks = (unsigned)((g.N + 255) / 256);
        if (!unit_scale)
        {
            S.clamp = 0;
            scale_comps_kernel<<<blocks, 256, 0, stream>>>(S);
        }
        // components outside 'mask' are excluded through a zero bnorm/target:
        // solve() decides activity from the residuals, so mask them by
        // temporarily treating them as converged.
        std::vector<int> its;
        std::vector<double> rn, r0;
        std::vector<double> bm = b;
        int it = solve_masked(spec, mask, bm, its, rn, r0);
        if (info)
        {
            info->iters = it;
            info->bnorm = 0;
            info->resid0 = 0;
            info->resid = 0;
            for (int c = 0; c < nt; ++c)
                if ((mask >> c) & 1u)
                {
                    info->bnorm = std::max(info->bnorm, b[c]);
                    info->resid0 = std::max(info->resid0, r0[c]);
                    info->resid = std::max(info->resid, rn[c]);
                }
        }
        if (it < 0) throw std::runtime_error("MLMG failed to converge (species)");
        if (between) between();
        if (!unit_scale || bound_specden)
        {
            for (int c = 0; c < nt; ++c) S.scale[c] = comp_scale(tsp[c]);
            S.clamp = bound_specden;
            scale_comps_kernel<<<blocks, 256, 0, stream>>>(S);
        }
        last_spec_iters = std::max(last_spec_iters, it);
        return it;
    }

    // MG solve restricted to the components of 'mask'.
    template <class MG>
    // sol (default: the state) holds component c at sol + compmap[c] * N;
    // rtol/atol default to the inputs' linsolve tolerances.
    int solve_masked(MG& spec, unsigned mask, const std::vector<double>& bnorm, std::vector<int>& its,
                     std::vector<double>& rn, std::vector<double>& r0, double* sol = nullptr, double rtol = -1.0,
                     double atol = -1.0)
    {
        if (!sol) sol = state.p;
        const double reltol = rtol >= 0.0 ? rtol : this->reltol, abstol = atol >= 0.0 ? atol : this->abstol;
        return spec.solve(sol, g.N, bnorm, reltol, abstol, &its, &rn, &r0, mask);
    }

    // Species solves, then the electron energy (Source/Evolve.cpp:274-280,
    // 355-372): the batched solve of the transported species runs first (the
    // ion and neutral solves read only Sborder, so the reference's order
    // electron -> energy -> others is kept where it matters); between the
    // linear solve and the unscaling the Joule kernel takes the new electron
    // gradients and completes the energy rhs; then the energy solve and the
    // electron temperature update.
    void species_energy_solve(Real t, Real dt, const Real* sb_old)
    {
        const int nt = (int)tsp.size();
        const int ce = een_comp();
        const long long N = g.N;
        const unsigned all = (nt >= 32) ? 0xffffffffu : ((1u << nt) - 1u);
        const unsigned emask = 1u << ce, smask = all & ~emask;
        // Sborder densities for the Joule heating and the temperature update
        hip_check(hipMemcpyAsync(sbsnap.p, state.p, sizeof(Real) * NUM_SPECIES * N, hipMemcpyDeviceToDevice, stream),
                  "sborder snapshot");
        std::vector<double> bn;
        auto joule = [&]() { launch_joule(t, dt, sb_old); };
        species_solve_t(spec, dt, smask, nullptr, joule);
        bn = read_red(0, nt);
        species_solve_t(spec, dt, emask, nullptr, nullptr, true, &bn);
        const unsigned blocks = (unsigned)((N + 255) / 256);
        {
            KScope ks(&kt, K_OTHER);
            energy_te_kernel<<<blocks, 256, 0, stream>>>(g, state.p, sbsnap.p + (long long)eidx * N, min_electron_temp,
                                                        eidx);
        }
        hip_check(hipGetLastError(), "energy_te");
    }

    void launch_joule(Real t, Real dt, const Real* sb_old)
    {
        const int nt = (int)tsp.size();
        const int ce = een_comp();
        const long long N = g.N;
        JouleArgs J{};
        J.P = phys(t, dt);
        J.state = state.p;
        J.sb = sbsnap.p;
        J.sb_old_een = sb_old + (long long)EEN_ID * N;
        J.expl = een_expl.p;
        J.psol = phi_sol.p;
        J.potscale = potscale;
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s)
            {
                J.pgconst[d][s] = (d < SD) ? pslab[d][s].c.p : nullptr;
                J.gconst[d][s] = (d < SD) ? sslab[d][s].c.p : nullptr;
                int kd = (d < SD) ? pside(0, d, s).kind : BK_NEU;
                double dep = 1.0;
                if (kd == BK_DIR) dep = -1.0;
                else if (kd == BK_ROBIN)
                {
                    const SideSpec& sp = pside(0, d, s);
                    const double h = g.dx[d];
                    dep = -(0.5 * sp.ra - sp.rb / h) / (0.5 * sp.ra + sp.rb / h);
                }
                J.pdep0[d][s] = dep;
                J.kind[d][s] = (d < SD) ? sside(ce, d, s).kind : BK_NEU;
            }
        J.escale = specscales[eidx];
        J.floor_jh = floor_jh;
        J.rscale = timescale / eescale;
        J.c_een = ce;
        J.nc = nt;
        J.beta = timescale;
        J.iD = siD0() + (long long)ce * N;
        J.rhs = sf0() + (long long)ce * N;
        J.partials = part_e.p;
        Launch3 L = cell_launch(g.n);
        {
            KScope ks(&kt, K_CELL);
            joule_energy_kernel<<<L.grid, L.block, 0, stream>>>(J);
            finalize_max_kernel<<<1, 1024, 0, stream>>>(part_e.p, nblk, red.p + ce);
        }
        hip_check(hipGetLastError(), "joule_energy");
    }

    bool is_bg(int sp) const
    {
        return std::find(bg_species.begin(), bg_species.end(), sp) != bg_species.end();
    }

    // Run-time parameters: the case's (case_param) followed by the voltage
    // amplitude of the electrode named by current_loc_surfaces[0].  The
    // case's are copied into the process-wide host and constant-memory
    // copies whenever this context becomes the active one.
#if defined(VD_NUM_CASE_PARAMS)
    Real cparams[VD_NUM_CASE_PARAMS];
    bool cparams_init = false;
#endif
    void activate_params()
    {
#if defined(VD_NUM_CASE_PARAMS)
        if (!cparams_init)
        {
            for (int i = 0; i < VD_NUM_CASE_PARAMS; ++i) cparams[i] = vd_case_param_defaults[i];
            cparams_init = true;
        }
        if (g_param_owner == this) return;
        for (int i = 0; i < VD_NUM_CASE_PARAMS; ++i) vd_case_param_host[i] = cparams[i];
        // the null stream waits for the work queued on every context's stream
        hip_check(hipMemcpyToSymbol(HIP_SYMBOL(vd_case_param_dev), vd_case_param_host,
                                    sizeof(amrex::Real) * VD_NUM_CASE_PARAMS),
                  "case parameters");
        g_param_owner = this;
#endif
    }
    Real& electrode_amp()
    {
        if (ncurrent_locs < 1) throw BadArg("run-time parameters need current_loc_surfaces");
        const int surf = current_loc_surfaces[0];
        return (surf % 2) ? vamp_hi[surf / 2] : vamp_lo[surf / 2];
    }
    void get_params(double* p, int n)
    {
#if defined(VD_NUM_CASE_PARAMS)
        if (!p || n != VD_NUM_CASE_PARAMS + 1) throw BadArg("wrong number of parameters");
        for (int i = 0; i < VD_NUM_CASE_PARAMS; ++i) p[i] = cparams[i];
        p[VD_NUM_CASE_PARAMS] = electrode_amp();
#else
        throw BadArg("this case has no run-time parameters");
#endif
    }
    void set_params(const double* p, int n)
    {
#if defined(VD_NUM_CASE_PARAMS)
        if (!p || n != VD_NUM_CASE_PARAMS + 1) throw BadArg("wrong number of parameters");
        Real& amp = electrode_amp();
        hip_check(hipStreamSynchronize(stream), "sync");
        for (int i = 0; i < VD_NUM_CASE_PARAMS; ++i) cparams[i] = p[i];
        g_param_owner = nullptr;
        activate_params();
        amp = p[VD_NUM_CASE_PARAMS];
#else
        throw BadArg("this case has no run-time parameters");
#endif
    }

    // -----------------------------------------------------------------------
    // Objective (time-integrated current through the electrode of current
    // location 0) with step checkpoints, and its discrete adjoint.
    struct Adjoint
    {
        int nsteps = 0, every = 0;
        bool ready = false;
        double J = 0.0;
        std::vector<DevBuf<Real>> ckpt;   // states at steps 0, every, 2 every, ...
        std::vector<double> tstep;        // time at the start of each step (+ final)
        DevBuf<Real> seg;                 // (every + 1) states of the segment being reversed
        DevBuf<Real> final_state;
        DevBuf<Real> lam, mu, ebar, potbar, yadj;
        DevBuf<double> part, bmax, grad;
        DevBuf<unsigned long long> red;
        int last_spec_iters = 0, last_pois_iters = 0;
    } adj;

    void check_adjoint_config()
    {
#if defined(VD_HAS_ADJOINT)
        auto need = [](bool ok, const char* what) {
            if (!ok) throw BadArg(std::string("adjoint: ") + what);
        };
        need(hyp_order == 1, "needs hyp_order = 1 (upwind)");
        need(g.coord == 0, "needs Cartesian geometry (coord_sys = 0)");
        need(num_correctors == 1, "needs num_timestep_correctors = 1");
        need(fine.empty(), "needs a single level");
        need(!elecenergy_solve && !using_LFA, "needs no electron energy equation and no LFA");
        need(!mixed, "needs the FP64 form");
        need(do_transport && do_reactions, "needs do_transport and do_reactions");
        need(!user_defined_vel && !user_defined_species && !user_defined_potential, "needs the default bc/velocity hooks");
        need(!bound_specden && !do_bg_reactions, "needs bound_specden = 0 and do_bg_reactions = 0");
        for (int s = 0; s < NUM_SPECIES; ++s) need(specscales[s] == 1.0, "needs unit species scales");
        need(track_current_den && track_integrated_currents && ncurrent_locs >= 1,
             "needs track_current_den, track_integrated_currents and current_loc_surfaces");
        for (int d = 0; d < SD; ++d)
        {
            if (g.periodic[d]) continue;
            for (int sd = 0; sd < 2; ++sd)
            {
                const int pk = pside(0, d, sd).kind;
                need(pk == BK_DIR || pk == BK_NEU, "needs Dirichlet or homogeneous Neumann potential sides");
                for (int c = 0; c < (int)tsp.size(); ++c)
                    need(sside(c, d, sd).kind == BK_NEU, "needs homogeneous Neumann species sides");
            }
        }
#else
        throw BadArg("this case has no adjoint (needs table chemistry and run-time parameters)");
#endif
    }

    // the checkpoints of the last objective forward no longer describe the
    // state (it, the time or the parameters were changed through the ABI)
    void invalidate_adjoint() { adj.ready = false; }

    void objective_forward(int nsteps, int every, double* J)
    {
        adj.ready = false;
        check_adjoint_config();
        if (nsteps < 1 || every < 1) throw BadArg("nsteps and checkpoint interval must be >= 1");
        const long long SZ = g.N * NVAR;
        const int M = (nsteps + every - 1) / every;
        if ((int)adj.ckpt.size() != M)
        {
            adj.ckpt.clear();
            adj.ckpt.resize(M);
        }
        for (auto& c : adj.ckpt)
            if (c.n != (size_t)SZ) c.alloc(SZ);
        if (adj.final_state.n != (size_t)SZ) adj.final_state.alloc(SZ);
        adj.nsteps = nsteps;
        adj.every = every;
        adj.tstep.assign(nsteps + 1, 0.0);
        for (int q = 0; q < MAX_CURRENT_LOCS; ++q) int_time[q] = 0.0;
        for (int n = 0; n < nsteps; ++n)
        {
            if (n % every == 0)
                hip_check(hipMemcpyAsync(adj.ckpt[n / every].p, state.p, sizeof(Real) * SZ, hipMemcpyDeviceToDevice,
                                         stream),
                          "checkpoint");
            adj.tstep[n] = cur_time;
            step();
        }
        adj.tstep[nsteps] = cur_time;
        hip_check(hipMemcpyAsync(adj.final_state.p, state.p, sizeof(Real) * SZ, hipMemcpyDeviceToDevice, stream),
                  "final state");
        hip_check(hipStreamSynchronize(stream), "objective sync");
        adj.J = int_time[0];
        adj.ready = true;
        if (J) *J = adj.J;
    }

#if defined(VD_HAS_ADJOINT)
    std::vector<double> adj_maxima(int ncomp)
    {
        finalize_max_kernel<<<ncomp, 1024, 0, stream>>>(adj.bmax.p, nblk, adj.red.p);
        std::vector<unsigned long long> h(ncomp);
        d2h_sync(h.data(), adj.red.p, sizeof(unsigned long long) * ncomp, stream, "adjoint norms d2h");
        std::vector<double> o(ncomp);
        for (int c = 0; c < ncomp; ++c) o[c] = bits_to_double(h[c]);
        return o;
    }
    void adj_accumulate()
    {
        finalize_sum_kernel<<<kAdjPar, 1024, 0, stream>>>(adj.part.p, nblk, adj.grad.p, 1);
    }

    void reverse_step(int n, const Real* s0, const Real* s1, Real t)
    {
        const long long N = g.N;
        const int nt = (int)tsp.size();
        const int surf = current_loc_surfaces[0];
        AdjCommon C{};
        C.g = g;
        C.s0 = s0;
        C.s1 = s1;
        C.nt = nt;
        for (int c = 0; c < nt; ++c)
        {
            C.tsp[c] = tsp[c];
            C.rscale[c] = timescale / specscales[tsp[c]];
        }
        C.dt = fixed_dt;
        C.ts = timescale;
        C.ps = potscale;
        C.Tg = gas_temperature;
        C.eidx = eidx;
        C.partials = adj.part.p;
        const Real dVdA = applied_voltage(t, (surf % 2) ? 1 : -1, voltage_profile, 1.0, 1.0, vfreq, vdur, vcen);
        Launch3 L = cell_launch(g.n);
        // A: objective of step n, species adjoint rhs
        AdjObjArgs A{};
        A.c = C;
        A.surf = surf;
        A.live_n = (t > 0.0) ? 1 : 0;
        A.has_next = (n + 1 < adj.nsteps) ? 1 : 0;
        A.lam = adj.lam.p;
        A.f0 = sf0();
        A.ebar = adj.ebar.p;
        A.potbar = adj.potbar.p;
        A.bmax = adj.bmax.p;
        adj_objective_kernel<<<L.grid, L.block, 0, stream>>>(A);
        adj_accumulate();
        std::vector<double> bs = adj_maxima(nt);
        fill_kernel<<<(unsigned)((N * NUM_SPECIES + 255) / 256), 256, 0, stream>>>(N * NUM_SPECIES, adj.mu.p, 0.0);
        std::vector<int> its;
        std::vector<double> rn, r0;
        int it = spec.solve(adj.mu.p, N, bs, reltol, 0.0, &its, &rn, &r0);
        if (it < 0) throw std::runtime_error("MLMG failed to converge (species adjoint)");
        adj.last_spec_iters = it;
        // C: rhs, flux, reactions, E/N
        AdjCellArgs Cc{};
        Cc.c = C;
        Cc.mu = adj.mu.p;
        Cc.lam = adj.lam.p;
        Cc.ebar = adj.ebar.p;
        adj_cell_kernel<<<L.grid, L.block, 0, stream>>>(Cc);
        adj_accumulate();
        // D: E and POT -> Poisson adjoint rhs
        AdjEfieldArgs D{};
        D.c = C;
        D.ebar = adj.ebar.p;
        D.potbar = adj.potbar.p;
        for (int d = 0; d < 3; ++d)
            for (int sd = 0; sd < 2; ++sd)
                D.dep0[d][sd] = (d < SD && pside(0, d, sd).kind == BK_DIR) ? -1.0 : 1.0;
        D.surf = surf;
        D.dVdA = dVdA;
        D.f0 = pf0();
        D.bmax = adj.bmax.p;
        adj_efield_kernel<<<L.grid, L.block, 0, stream>>>(D);
        adj_accumulate();
        std::vector<double> bp = adj_maxima(1);
        fill_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(N, adj.yadj.p, 0.0);
        it = pois.solve(adj.yadj.p, N, bp, reltol, 0.0, &its, &rn, &r0);
        if (it < 0) throw std::runtime_error("MLMG failed to converge (potential adjoint)");
        adj.last_pois_iters = it;
        // E: charge density and folded Dirichlet data
        AdjRhoArgs E{};
        E.c = C;
        E.y = adj.yadj.p;
        E.lam = adj.lam.p;
        E.surf = surf;
        E.dVdA = dVdA;
        E.do_spacechrg = do_spacechrg;
        adj_rho_kernel<<<L.grid, L.block, 0, stream>>>(E);
        adj_accumulate();
        hip_check(hipGetLastError(), "adjoint step");
    }
#endif

    void adjoint_gradient(double* grad, int np)
    {
#if defined(VD_HAS_ADJOINT)
        if (!adj.ready) throw BadArg("adjoint: run the objective forward first");
        if (!grad || np != kAdjPar) throw BadArg("adjoint: wrong gradient length");
        check_adjoint_config();
        // a failure part way leaves the state between checkpoints: the
        // forward has to be run again (restored below on success)
        adj.ready = false;
        const long long N = g.N, SZ = N * NVAR;
        const int K = adj.every, M = (int)adj.ckpt.size();
        if (adj.seg.n != (size_t)(SZ * (K + 1))) adj.seg.alloc(SZ * (K + 1));
        if (adj.lam.n != (size_t)(N * NUM_SPECIES))
        {
            adj.lam.alloc(N * NUM_SPECIES);
            adj.mu.alloc(N * NUM_SPECIES);
            adj.ebar.alloc(N * SD);
            adj.potbar.alloc(N);
            adj.yadj.alloc(N);
            adj.part.alloc((size_t)nblk * kAdjPar);
            adj.bmax.alloc((size_t)nblk * tsp.size());
            adj.grad.alloc(kAdjPar);
            adj.red.alloc(64);
        }
        fill_kernel<<<(unsigned)((N * NUM_SPECIES + 255) / 256), 256, 0, stream>>>(N * NUM_SPECIES, adj.lam.p, 0.0);
        fill_kernel<<<1, 64, 0, stream>>>(kAdjPar, adj.grad.p, 0.0);
        // the recomputation steps again: keep what the forward left behind
        double keep[4][MAX_CURRENT_LOCS];
        for (int q = 0; q < MAX_CURRENT_LOCS; ++q)
        {
            keep[0][q] = int_cond[q];
            keep[1][q] = int_disp[q];
            keep[2][q] = int_area[q];
            keep[3][q] = int_time[q];
        }
        const int keep_istep = istep;
        for (int m = M - 1; m >= 0; --m)
        {
            const int n0 = m * K, n1 = std::min(n0 + K, adj.nsteps);
            hip_check(hipMemcpyAsync(state.p, adj.ckpt[m].p, sizeof(Real) * SZ, hipMemcpyDeviceToDevice, stream),
                      "restore checkpoint");
            cur_time = adj.tstep[n0];
            for (int n = n0; n < n1; ++n)
            {
                hip_check(hipMemcpyAsync(adj.seg.p + (n - n0) * SZ, state.p, sizeof(Real) * SZ,
                                         hipMemcpyDeviceToDevice, stream),
                          "segment state");
                step();
            }
            hip_check(hipMemcpyAsync(adj.seg.p + (n1 - n0) * SZ, state.p, sizeof(Real) * SZ, hipMemcpyDeviceToDevice,
                                     stream),
                      "segment state");
            for (int n = n1 - 1; n >= n0; --n)
                reverse_step(n, adj.seg.p + (n - n0) * SZ, adj.seg.p + (n - n0 + 1) * SZ, adj.tstep[n]);
        }
        // leave the context at the end of the forward run
        hip_check(hipMemcpyAsync(state.p, adj.final_state.p, sizeof(Real) * SZ, hipMemcpyDeviceToDevice, stream),
                  "restore final");
        cur_time = adj.tstep[adj.nsteps];
        istep = keep_istep;
        for (int q = 0; q < MAX_CURRENT_LOCS; ++q)
        {
            int_cond[q] = keep[0][q];
            int_disp[q] = keep[1][q];
            int_area[q] = keep[2][q];
            int_time[q] = keep[3][q];
        }
        hip_check(hipMemcpyAsync(grad, adj.grad.p, sizeof(double) * kAdjPar, hipMemcpyDeviceToHost, stream),
                  "gradient");
        hip_check(hipStreamSynchronize(stream), "adjoint sync");
        adj.ready = true;
#else
        (void)grad;
        (void)np;
        throw BadArg("this case has no adjoint (needs table chemistry and run-time parameters)");
#endif
    }

    // One coarse step of Vidyut::Evolve (Source/Evolve.cpp:69-714).
    // -----------------------------------------------------------------------
    // Static refinement (max_level > 0): the multi-level step of
    // Source/Evolve.cpp:69-714 without subcycling (every level advances with
    // the coarse dt, AverageDown at the end), composite potential and
    // species solves over all levels (one MLABecLaplacian over levels
    // 0..finest, Source/PotentialSolve.cpp:445-500, Source/ScalarSolve.cpp:
    // 573-585) and FillPatchTwoLevels ghosts for the fine stencils.
    int nlev() const { return 1 + (int)fine.size(); }
    Grid& lgrid(int l) { return l ? fine[l - 1]->g : g; }
    Real* lstate(int l) { return l ? fine[l - 1]->state.p : state.p; }
    Real* lsol(int l) { return l ? fine[l - 1]->phi_sol.p : phi_sol.p; }
    MGSolver<SD, Real>& lmg(int l, bool sp) { return l ? (sp ? fine[l - 1]->spec : fine[l - 1]->pois) : (sp ? spec : pois); }
    unsigned long long* lred(int l) { return l ? fine[l - 1]->red.p : red.p; }
    // unknowns of a composite solve on level l: the potential solution or
    // the transported components of the state
    Real* usol(int l, bool sp) { return sp ? lstate(l) : lsol(l); }
    CompSet ucomps(bool sp) const
    {
        CompSet c{};
        c.nc = sp ? (int)tsp.size() : 1;
        for (int i = 0; i < c.nc; ++i) c.map[i] = sp ? tsp[i] : 0;
        return c;
    }
    static CompSet idcomps(int nc)
    {
        CompSet c{};
        c.nc = nc;
        for (int i = 0; i < nc; ++i) c.map[i] = i;
        return c;
    }
    std::vector<double> read_bits(const unsigned long long* p, int n)
    {
        std::vector<unsigned long long> h(n);
        d2h_sync(h.data(), p, sizeof(unsigned long long) * n, stream, "d2h");
        std::vector<double> o(n);
        for (int i = 0; i < n; ++i) o[i] = bits_to_double(h[i]);
        return o;
    }

    void cf_interp(int l, const double* parent, long long pstride, const CompSet& cs, double* slabs, double w)
    {
        FineLevel& F = *fine[l - 1];
        int n1 = 1, n2 = 1;
        for (int d = 0; d < SD; ++d)
        {
            n1 = std::max(n1, F.sl.n1[d]);
            n2 = std::max(n2, F.sl.n2[d]);
        }
        dim3 grd((n1 + 63) / 64, n2, 2 * SD * cs.nc);
        amr_cf_interp_kernel<<<grd, 64, 0, stream>>>(F.m, parent, pstride, cs, F.sl, slabs, w);
    }
    void cf_fold(int l, const double* slabs, int nc, double* rhs, const MGSolver<SD, Real>& M, bool sp)
    {
        FineLevel& F = *fine[l - 1];
        for (int d = 0; d < SD; ++d)
        {
            dim3 grd((F.sl.n1[d] + 63) / 64, F.sl.n2[d], 2 * nc);
            amr_cf_fold_kernel<<<grd, 64, 0, stream>>>(F.m, d, slabs, F.sl, nc, rhs, M.beta, M.bconst,
                                                      sp ? M.iD0.p : nullptr, F.g.dx[d]);
        }
    }
    // reflux of fine level l into the parent residual r
    void reflux(int l, const double* fv, const CompSet& fm, const double* slabs, const double* cv, const CompSet& cm,
                double* r, bool sp)
    {
        FineLevel& F = *fine[l - 1];
        RefluxArgs A{};
        A.m = F.m;
        A.fine = fv;
        A.fstride = F.g.N;
        A.fmap = fm;
        A.slabs = slabs;
        A.sl = F.sl;
        A.coarse = cv;
        A.cstride = F.m.pN;
        A.cmap = cm;
        A.r = r;
        A.beta = lmg(l, sp).beta;
        A.bconst = lmg(l, sp).bconst;
        A.iDf = sp ? lmg(l, sp).iD0.p : nullptr;
        A.iDc = sp ? lmg(l - 1, sp).iD0.p : nullptr;
        int m1 = 1, m2 = 1;
        for (int d = 0; d < 3; ++d)
        {
            A.hf[d] = d < SD ? F.g.dx[d] : 1.0;
            A.hc[d] = d < SD ? lgrid(l - 1).dx[d] : 1.0;
        }
        for (int d = 0; d < SD; ++d)
        {
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            m1 = std::max(m1, F.g.n[t1] / 2);
            m2 = std::max(m2, (SD > 2) ? F.g.n[t2] / 2 : 1);
        }
        // every side of the box in one launch
        dim3 grd((m1 + 63) / 64, m2, 2 * SD * fm.nc);
        amr_reflux_kernel<<<grd, 64, 0, stream>>>(A);
    }
    // covered parent cells of level l = mean of the children
    void average(int l, const double* fv, const CompSet& fm, double* cv, const CompSet& cm)
    {
        FineLevel& F = *fine[l - 1];
        const int hz = (SD > 2) ? F.g.n[2] / 2 : 1;
        dim3 grd((F.g.n[0] / 2 + 63) / 64, (SD > 1) ? F.g.n[1] / 2 : 1, hz * fm.nc);
        amr_average_kernel<<<grd, 64, 0, stream>>>(F.m, fv, F.g.N, fm, cv, F.m.pN, cm);
    }
    void prolong(int l, const double* cv, double* fv, int nc, unsigned act)
    {
        FineLevel& F = *fine[l - 1];
        const int hz = (SD > 2) ? F.g.n[2] / 2 : 1;
        dim3 grd((F.g.n[0] / 2 + 63) / 64, (SD > 1) ? F.g.n[1] / 2 : 1, hz * nc);
        amr_prolong_kernel<SD><<<grd, 64, 0, stream>>>(F.m, cv, fv, act);
    }
    // per-component max |v| over nc components of level l's N cells, into
    // words [l * kMaxComp, l * kMaxComp + nc) of ml_nrm (read by the caller)
    void absmax(int l, const double* v, int nc, unsigned act)
    {
        const long long N = lgrid(l).N;
        const long long nb = std::min<long long>(1024, (N + 255) / 256);
        if (!ml_partn.p) ml_partn.alloc((size_t)1024 * kMaxComp);
        if (ml_nrm.n < (size_t)nlev() * kMaxComp) ml_nrm.alloc((size_t)nlev() * kMaxComp);
        double* partials = l ? fine[l - 1]->partn.p : ml_partn.p;
        amr_absmax_kernel<<<dim3((unsigned)nb, nc), 256, 0, stream>>>(v, N, act, partials, nb);
        finalize_max_kernel<<<nc, 1024, 0, stream>>>(partials, nb, ml_nrm.p + (size_t)l * kMaxComp);
    }
    DevBuf<double> ml_partn;
    DevBuf<unsigned long long> ml_nrm;

    // Composite residual into every level's fx[0] (finest first: a level's
    // covered cells take the mean of its child's residual and the cells
    // across the interface the reflux of the child's face gradients);
    // returns the per-component max over all levels.
    // fresh: the fine levels' folded rhs is rebuilt from the raw rhs in full
    // (the first residual of a solve); otherwise only the boundary cells the
    // C/F fold changes are restored before folding again.
    std::vector<double> ml_residual(bool sp, unsigned act, bool fresh)
    {
        const int L = nlev() - 1;
        const CompSet uc = ucomps(sp);
        const int nc = uc.nc;
        for (int l = L; l >= 0; --l)
        {
            auto& M = lmg(l, sp);
            const long long N = lgrid(l).N;
            if (l >= 1)
            {
                FineLevel& F = *fine[l - 1];
                const double* raw = sp ? F.sfraw.p : F.pfraw.p;
                if (fresh)
                    hip_check(hipMemcpyAsync(M.f0.p, raw, sizeof(double) * nc * N, hipMemcpyDeviceToDevice, stream),
                              "raw rhs");
                else
                {
                    int n1 = 1, n2 = 1;
                    for (int d = 0; d < SD; ++d)
                    {
                        n1 = std::max(n1, F.sl.n1[d]);
                        n2 = std::max(n2, F.sl.n2[d]);
                    }
                    amr_shell_copy_kernel<<<dim3((n1 + 63) / 64, n2, 2 * SD * nc), 64, 0, stream>>>(F.m, F.sl, nc, raw,
                                                                                                 M.f0.p);
                }
                double* sl = sp ? F.sslab.p : F.pslab.p;
                cf_interp(l, usol(l - 1, sp), lgrid(l - 1).N, uc, sl, 1.0);
                cf_fold(l, sl, nc, M.f0.p, M, sp);
            }
            M.residual_state(usol(l, sp), N, act, false);
            if (l < L)
            {
                FineLevel& C = *fine[l];
                average(l + 1, lmg(l + 1, sp).fx[0].p, idcomps(nc), M.fx[0].p, idcomps(nc));
                reflux(l + 1, usol(l + 1, sp), uc, sp ? C.sslab.p : C.pslab.p, usol(l, sp), uc, M.fx[0].p, sp);
            }
        }
        std::vector<double> rn(nc, 0.0);
        for (int l = 0; l <= L; ++l) absmax(l, lmg(l, sp).fx[0].p, nc, act);
        const std::vector<double> m = read_bits(ml_nrm.p, (L + 1) * kMaxComp);   // one read for every level
        for (int l = 0; l <= L; ++l)
            for (int c = 0; c < nc; ++c)
                if ((act >> c) & 1u) rn[c] = vmax(rn[c], m[(size_t)l * kMaxComp + c]);
        return rn;
    }

    // One composite V-cycle on the residuals in fx[0] (MLMG over AMR
    // levels): a fine level smooths its correction with zero parent
    // correction and hands the residual of that correction (mean over
    // covered parents, reflux across the interface) to its parent; the
    // coarsest level runs its own V-cycle; on the way up each fine level adds
    // the interpolated parent correction and smooths with the parent
    // correction in its C/F ghosts.  Then u += e on every level and the
    // covered cells are averaged down.
    void ml_vcycle(bool sp, unsigned act)
    {
        const int L = nlev() - 1;
        const int nc = sp ? (int)tsp.size() : 1;
        const CompSet idc = idcomps(nc);
        for (int l = L; l >= 1; --l)
        {
            FineLevel& F = *fine[l - 1];
            auto& M = lmg(l, sp);
            auto& P = lmg(l - 1, sp);
            // constant-coefficient components: the fused pass (2 sweeps
            // from zero, then the residual), the others per colour
            const unsigned fm = M.fused_mask(0, act, true), rm = act & ~fm;
            if (fm) M.fused_pass(0, fm, true, 0u, true, F.rho.p);
            if (rm)
            {
                M.gsrb(0, M.ex[0].p, M.fx[0].p, rm, pre_smooth, true);
                M.residual_plain(M.ex[0].p, M.fx[0].p, F.rho.p, rm);
            }
            average(l, F.rho.p, idc, P.fx[0].p, idc);
            reflux(l, M.ex[0].p, idc, nullptr, nullptr, idc, P.fx[0].p, sp);
        }
        lmg(0, sp).vcycle(0, act);
        for (int l = 1; l <= L; ++l)
        {
            FineLevel& F = *fine[l - 1];
            auto& M = lmg(l, sp);
            auto& P = lmg(l - 1, sp);
            prolong(l, P.ex[0].p, M.ex[0].p, nc, act);
            cf_interp(l, P.ex[0].p, lgrid(l - 1).N, idc, F.tslab.p, 1.0);
            cf_fold(l, F.tslab.p, nc, M.fx[0].p, M, sp);
            const unsigned fm = M.fused_mask(0, act, true), rm = act & ~fm;
            if (rm) M.gsrb(0, M.ex[0].p, M.fx[0].p, rm, post_smooth, false);
            if (fm) M.fused_pass(0, fm, false, rm, true);
        }
        for (int l = 0; l <= L; ++l) lmg(l, sp).add_to_state(usol(l, sp), lgrid(l).N, act);
        const CompSet uc = ucomps(sp);
        for (int l = L; l >= 1; --l) average(l, usol(l, sp), uc, usol(l - 1, sp), uc);
    }

    // Composite solve to the MLMG stopping test over all levels.
    int ml_solve(bool sp, unsigned mask, const std::vector<double>& bnorm, double* r0max, double* rmax)
    {
        const int nc = sp ? (int)tsp.size() : 1;
        std::vector<double> r0 = ml_residual(sp, mask, true), rn = r0;
        require_finite(bnorm.data(), nc, "right-hand side norm", mask);
        require_finite(r0.data(), nc, "residual norm", mask);
        std::vector<double> target(nc, 0.0);
        unsigned active = 0;
        for (int c = 0; c < nc; ++c)
        {
            if (!((mask >> c) & 1u)) continue;
            target[c] = std::max(abstol, reltol * std::max(bnorm[c], r0[c]));
            if (!(r0[c] <= target[c])) active |= 1u << c;
        }
        int used = 0;
        // the coarsest level's direct bottom solve (bases for its operator)
        lmg(0, sp).direct_prepare(active);
        for (int it = 1; it <= maxiter && active; ++it)
        {
            ml_vcycle(sp, active);
            rn = ml_residual(sp, active, false);
            require_finite(rn.data(), nc, "residual norm", active);
            for (int c = 0; c < nc; ++c)
                if ((active >> c) & 1u)
                {
                    used = it;
                    if (rn[c] <= target[c]) active &= ~(1u << c);
                }
        }
        if (r0max) *r0max = *std::max_element(r0.begin(), r0.end());
        if (rmax) *rmax = *std::max_element(rn.begin(), rn.end());
        return active ? -1 : used;
    }

    void time_scales_ml(double out[3])
    {
        const double big = std::numeric_limits<double>::max();
        out[0] = out[1] = out[2] = big;
        for (int l = 0; l < nlev(); ++l)
        {
            Grid& G = lgrid(l);
            double* pt = l ? fine[l - 1]->part.p : part.p;
            const long long nb = l ? fine[l - 1]->nblk : nblk;
            TimeScaleArgs A{G, lstate(l), gas_temperature, eidx, pt};
            const int kz = tscale_kz(G.n[2]);
            int ce[3] = {G.n[0], G.n[1], (G.n[2] + kz - 1) / kz};
            Launch3 L = cell_launch(ce);
            const long long nbk = (long long)L.grid.x * L.grid.y * L.grid.z;
            (void)nb;
            {
                KScope ks(&kt, K_TSCALE);
                time_scales_kernel<<<L.grid, L.block, 0, stream>>>(A, kz);
                finalize_max_kernel<<<3, 1024, 0, stream>>>(pt, nbk, lred(l) + 33);
            }
            auto m = read_bits(lred(l) + 33, 3);
            require_finite(m.data(), 3, "time-scale maximum");
            if (m[0] > 0)
                for (int d = 0; d < SD; ++d) out[0] = std::min(out[0], G.dx[d] / m[0]);
            if (m[1] > 0)
                for (int d = 0; d < SD; ++d) out[1] = std::min(out[1], 0.5 * G.dx[d] * G.dx[d] / m[1] / SD);
            if (m[2] > 0) out[2] = std::min(out[2], EPS0 / ECHARGE / m[2]);
        }
    }

    // potential: level-0 rhs and domain data as in the single-level solve,
    // raw rhs of the fine levels, composite solve, E field per level
    void solve_potential_ml(Real time)
    {
        double b = pot_rhs0(time, state.p);
        for (int l = 1; l < nlev(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            PhysParams P = phys(time, fixed_dt);
            P.g = F.g;
            PoissonRhsArgs R{P, d_pp.p, F.state.p, 1.0 / potscale, do_spacechrg, F.pfraw.p, F.part.p};
            Launch3 L = cell_launch(F.g.n);
            {
                KScope ks(&kt, K_PRHS);
                poisson_rhs_kernel<<<L.grid, L.block, 0, stream>>>(R);
                finalize_max_kernel<<<1, 1024, 0, stream>>>(F.part.p, F.nblk, F.red.p + 32);
            }
            b = vmax(b, read_bits(F.red.p + 32, 1)[0]);
            fill_kernel<<<(unsigned)((F.g.N + 255) / 256), 256, 0, stream>>>(F.g.N, F.phi_sol.p, 0.0);
        }
        double r0 = 0, rn = 0;
        const int it = ml_solve(false, 1u, std::vector<double>(1, b), &r0, &rn);
        last_pot_r0 = r0;
        last_pot_r = rn;
        last_pot_b = b;
        if (it < 0) throw std::runtime_error("MLMG failed to converge (potential)");
        last_pot_iters = it;
        pot_efield0(time);
        for (int l = 1; l < nlev(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            cf_interp(l, lsol(l - 1), lgrid(l - 1).N, idcomps(1), F.pgc.p, kCfW);
            EFieldArgs E{};
            E.P = phys(time, fixed_dt);
            E.P.g = F.g;
            E.sol = F.phi_sol.p;
            E.state = F.state.p;
            E.potscale = potscale;
            for (int d = 0; d < 3; ++d)
                for (int s2 = 0; s2 < 2; ++s2)
                {
                    E.gconst[d][s2] = (d < SD) ? F.pgc.p + F.sl.off[d][s2] : nullptr;
                    E.dep0[d][s2] = kCfD0;
                    E.dep1[d][s2] = kCfD1;
                }
            E.eidx = eidx;
            Launch3 L = cell_launch(F.g.n);
            KScope ks(&kt, K_EFIELD);
            efield_kernel<<<L.grid, L.block, 0, stream>>>(E);
        }
        hip_check(hipGetLastError(), "potential (levels)");
    }

    // Sborder of the fine levels: ghosted copies filled from their parents
    void fill_sborder_ml()
    {
        for (int l = 1; l < nlev(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            const int gx = F.g.n[0] + 2 * F.ng, gy = (SD > 1) ? F.g.n[1] + 2 * F.ng : 1,
                      gz = (SD > 2) ? F.g.n[2] + 2 * F.ng : 1;
            dim3 grd((gx + 63) / 64, gy, gz);
            KScope ks(&kt, K_OTHER);
            amr_fill_ghosted_kernel<<<grd, 64, 0, stream>>>(F.m, F.ng, EFZ_ID + 1, F.state.p, lstate(l - 1), F.sbg.p);
        }
    }

    // explicit sources of every level: face fluxes, fine faces averaged onto
    // the covered coarse faces (average_down_faces), then divergence +
    // reactions + implicit rhs per level and the level-0 domain data
    void species_inputs_ml(Real time, Real dt)
    {
        species_flux0(time, dt);
        for (int l = 1; l < nlev(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            FluxArgs A = flux_args(time, dt);
            A.P.g = F.g;
            A.state = F.state.p;
            A.sg = F.sbg.p;
            A.sng = F.ng;
            // a fine box has no domain faces: no zero-flux wall codes
            for (int c = 0; c < A.nt; ++c)
                for (int d = 0; d < 3; ++d) A.bclo[c][d] = A.bchi[c][d] = PERBC;
            launch_flux_level(A, F.g, F.flux);
        }
        const int nt = (int)tsp.size();
        for (int l = nlev() - 1; l >= 1; --l)
        {
            FineLevel& F = *fine[l - 1];
            for (int d = 0; d < SD; ++d)
            {
                const int em0 = F.g.n[0] / 2 + (d == 0), em1 = (SD > 1) ? F.g.n[1] / 2 + (d == 1) : 1;
                const int em2 = (SD > 2) ? F.g.n[2] / 2 + (d == 2) : 1;
                dim3 grd((em0 + 63) / 64, em1, em2 * nt);
                void* cf = (l == 1) ? (void*)flux[d].p : (void*)fine[l - 2]->flux[d].p;
                KScope ks(&kt, K_OTHER);
                amr_avgdown_faces_kernel<<<grd, 64, 0, stream>>>(F.m, d, nt, F.flux[d].p, cf, 0);
            }
        }
        species_cell0(time, dt, state.p, nullptr, nullptr);
        for (int l = 1; l < nlev(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            CellArgs C = cell_args(time, dt, F.state.p);
            C.P.g = F.g;
            C.state = F.state.p;
            for (int d = 0; d < 3; ++d)
            {
                C.F[d] = (d < SD) ? (const void*)F.flux[d].p : nullptr;
                C.nf[d] = (d < SD) ? (long long)F.flux[d].n / nt : 0;
            }
            C.rhs = F.sfraw.p;
            C.iD = F.spec.iD0.p;
            C.defer_out = F.een_expl.p;
            C.rxn_out = do_bg_reactions ? F.rxn.p : nullptr;
            C.partials = F.part.p;
            C.partials_id = F.part_id.p;
            Launch3 L = cell_launch(F.g.n);
            {
                KScope ks(&kt, K_CELL);
                cell_source_kernel<double><<<L.grid, L.block, 0, stream>>>(C);
                const long long nw = F.nblk * kCellWaves;   // one partial per wave
                // max |rhs|, and max 1/D and max -1/D per component (uniform
                // coefficients)
                finalize_max3_kernel<<<3 * nt, 1024, 0, stream>>>(F.part.p, F.part_id.p, nw, nt, F.red.p, F.red2.p,
                                                                   F.red2.p + 32);
            }
        }
        species_side0(time, dt);
        hip_check(hipGetLastError(), "species inputs (levels)");
    }

    void launch_joule_fine(int l, Real t, Real dt)
    {
        FineLevel& F = *fine[l - 1];
        const int ce = een_comp();
        const long long N = F.g.N;
        cf_interp(l, lstate(l - 1), lgrid(l - 1).N, [&] {
            CompSet c{};
            c.nc = 1;
            c.map[0] = eidx;
            return c;
        }(), F.eslab.p, kCfW);
        JouleArgs J{};
        J.P = phys(t, dt);
        J.P.g = F.g;
        J.state = F.state.p;
        J.sb = F.sbsnap.p;
        J.sb_old_een = F.state.p + (long long)EEN_ID * N;
        J.expl = F.een_expl.p;
        J.psol = F.phi_sol.p;
        J.potscale = potscale;
        for (int d = 0; d < 3; ++d)
            for (int s2 = 0; s2 < 2; ++s2)
            {
                J.pgconst[d][s2] = (d < SD) ? F.pgc.p + F.sl.off[d][s2] : nullptr;
                J.egconst[d][s2] = (d < SD) ? F.eslab.p + F.sl.off[d][s2] : nullptr;
                J.pdep0[d][s2] = kCfD0;
                J.pdep1[d][s2] = kCfD1;
                J.gconst[d][s2] = nullptr;
                J.kind[d][s2] = BK_NEU;   // coarse/fine data enters in the composite solve
            }
        J.escale = specscales[eidx];
        J.floor_jh = floor_jh;
        J.rscale = timescale / eescale;
        J.c_een = ce;
        J.nc = (int)tsp.size();
        J.beta = timescale;
        J.iD = F.spec.iD0.p + (long long)ce * N;
        J.rhs = F.sfraw.p + (long long)ce * N;
        J.partials = F.part_e.p;
        J.cf = 1;
        J.sbg = F.sbg.p;
        J.sng = F.ng;
        Launch3 L = cell_launch(F.g.n);
        {
            KScope ks(&kt, K_CELL);
            joule_energy_kernel<<<L.grid, L.block, 0, stream>>>(J);
            finalize_max_kernel<<<1, 1024, 0, stream>>>(F.part_e.p, F.nblk, F.red.p + ce);
        }
    }

    void scale_level(int l, unsigned mask, bool unscale, int clamp)
    {
        const int nt = (int)tsp.size();
        PostArgs S{};
        S.g = lgrid(l);
        S.state = lstate(l);
        S.nt = nt;
        S.active = mask;
        S.clamp = clamp;
        for (int c = 0; c < nt; ++c)
        {
            S.tsp[c] = tsp[c];
            S.scale[c] = unscale ? comp_scale(tsp[c]) : 1.0 / comp_scale(tsp[c]);
            S.floor[c] = (tsp[c] == EEN_ID) ? -std::numeric_limits<double>::max()
                                            : ((tsp[c] == eidx) ? min_electron_density : min_species_density);
        }
        scale_comps_kernel<<<(unsigned)((S.g.N + 255) / 256), 256, 0, stream>>>(S);
    }

    // composite species solves (and the electron energy), all levels
    void species_solve_ml(Real t, Real dt)
    {
        const int nt = (int)tsp.size();
        const int ce = een_comp();
        const unsigned all = (nt >= 32) ? 0xffffffffu : ((1u << nt) - 1u);
        const unsigned emask = (ce >= 0) ? (1u << ce) : 0u, smask = all & ~emask;
        spec_prepare(dt);
        for (int l = 1; l < nlev(); ++l)
        {
            FineLevel& F = *fine[l - 1];
            F.spec.alpha_a = 1.0 * ((1.0 / dt) * timescale);
            // components with a spatially uniform coefficient on this level
            // run the constant-coefficient kernels (as on level 0)
            auto h = read_bits(F.red2.p, 64);
            unsigned um = 0;
            std::vector<double> vals(nt, 0.0);
            for (int c = 0; c < nt; ++c)
            {
                const double mx = h[c], mn = -h[32 + c];
                if (mx == mn && !(mx != mx))
                {
                    um |= 1u << c;
                    vals[c] = 2.0 / (mx + mx);
                }
            }
            F.spec.set_uniform(um, vals);
        }
        if (ce >= 0)
            for (int l = 0; l < nlev(); ++l)
                hip_check(hipMemcpyAsync(l ? fine[l - 1]->sbsnap.p : sbsnap.p, lstate(l),
                                         sizeof(Real) * NUM_SPECIES * lgrid(l).N, hipMemcpyDeviceToDevice, stream),
                          "sborder snapshot");
        auto bnorm = [&]() {
            std::vector<double> b(nt, 0.0);
            for (int l = 0; l < nlev(); ++l)
            {
                auto m = read_bits(lred(l), nt);
                for (int c = 0; c < nt; ++c) b[c] = vmax(b[c], m[c]);
            }
            return b;
        };
        auto unit = [&](unsigned mask) {
            for (int c = 0; c < nt; ++c)
                if (((mask >> c) & 1u) && comp_scale(tsp[c]) != 1.0) return false;
            return true;
        };
        auto bn = bnorm();
        if (!unit(smask))
            for (int l = 0; l < nlev(); ++l) scale_level(l, smask, false, 0);
        int it = ml_solve(true, smask, bn, &last_spec_r0, &last_spec_r);
        if (it < 0) throw std::runtime_error("MLMG failed to converge (species)");
        last_spec_iters = std::max(last_spec_iters, it);
        if (ce >= 0)
        {
            launch_joule(t, dt, state.p);
            for (int l = 1; l < nlev(); ++l) launch_joule_fine(l, t, dt);
        }
        if (!unit(smask) || bound_specden)
            for (int l = 0; l < nlev(); ++l) scale_level(l, smask, true, bound_specden);
        if (ce < 0) return;
        bn = bnorm();
        if (!unit(emask))
            for (int l = 0; l < nlev(); ++l) scale_level(l, emask, false, 0);
        it = ml_solve(true, emask, bn, nullptr, nullptr);
        if (it < 0) throw std::runtime_error("MLMG failed to converge (electron energy)");
        last_spec_iters = std::max(last_spec_iters, it);
        if (!unit(emask))
            for (int l = 0; l < nlev(); ++l) scale_level(l, emask, true, 0);
        for (int l = 0; l < nlev(); ++l)
        {
            const long long N = lgrid(l).N;
            const Real* sbne = (l ? fine[l - 1]->sbsnap.p : sbsnap.p) + (long long)eidx * N;
            KScope ks(&kt, K_OTHER);
            energy_te_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(lgrid(l), lstate(l), sbne,
                                                                            min_electron_temp, eidx);
        }
    }

    void step_ml()
    {
        time_scales_ml(last_ts);
        const Real dt = fixed_dt;
        const Real t = cur_time;
        last_spec_iters = 0;
        solve_potential_ml(t);
        fill_sborder_ml();
        species_inputs_ml(t, dt);
        const int nt = (int)tsp.size();
        const unsigned all = (nt >= 32) ? 0xffffffffu : ((1u << nt) - 1u);
        if (elecenergy_solve)
        {
            species_solve_ml(t, dt);
            for (int l = 0; l < nlev(); ++l)
            {
                const long long N = lgrid(l).N;
                LfaArgs La{lgrid(l), lstate(l), gas_temperature, min_electron_temp, eidx,
                           l ? fine[l - 1]->sbsnap.p : sbsnap.p};
                const unsigned blocks = (unsigned)((N + 255) / 256);
                if (using_LFA) lfa_te_kernel<<<blocks, 256, 0, stream>>>(La);
                if (using_LFA) lfa_een_kernel<<<blocks, 256, 0, stream>>>(La, lstate(l) + (long long)CMASK_ID * N);
            }
        } else
        {
            for (int l = 0; l < nlev(); ++l)
            {
                const long long N = lgrid(l).N;
                LfaArgs La{lgrid(l), lstate(l), gas_temperature, min_electron_temp, eidx, lstate(l)};
                if (using_LFA) lfa_te_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(La);
            }
            (void)all;
            species_solve_ml(t, dt);
            for (int l = 0; l < nlev(); ++l)
            {
                const long long N = lgrid(l).N;
                LfaArgs La{lgrid(l), lstate(l), gas_temperature, min_electron_temp, eidx, lstate(l)};
                if (using_LFA)
                    lfa_een_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(La, lstate(l) + (long long)CMASK_ID * N);
            }
        }
        if (do_bg_reactions)
            for (int l = 0; l < nlev(); ++l)
                for (int sp : bg_species)
                {
                    const long long N = lgrid(l).N;
                    bg_update_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(
                        N, lstate(l), l ? fine[l - 1]->rxn.p : rxn.p, sp, dt, bound_specden, min_species_density);
                }
        // AverageDown (Source/GridWork.cpp:112-121): every component
        for (int l = nlev() - 1; l >= 1; --l)
            for (int c0 = 0; c0 < NVAR; c0 += kMaxComp)
            {
                CompSet cs{};
                cs.nc = std::min(kMaxComp, NVAR - c0);
                for (int c = 0; c < cs.nc; ++c) cs.map[c] = c0 + c;
                KScope ks(&kt, K_OTHER);
                average(l, lstate(l), cs, lstate(l - 1), cs);
            }
        hip_check(hipGetLastError(), "step (levels)");
        cur_time += dt;
        istep++;
    }

    void step()
    {
        if (!fine.empty())
        {
            step_ml();
            return;
        }
        time_scales(last_ts);
        const Real dt = fixed_dt;
        const long long N = g.N;
        const unsigned blocks = (unsigned)((N + 255) / 256);
        if (num_correctors > 1)
            hip_check(hipMemcpyAsync(state_old.p, state.p, sizeof(Real) * N * NVAR, hipMemcpyDeviceToDevice, stream),
                      "copy old");
        const Real* sb_old = (num_correctors > 1) ? state_old.p : state.p;
        if (track_current_den)
            hip_check(hipMemcpyAsync(eold.p, state.p + (long long)EFX_ID * N, sizeof(Real) * N * SD,
                                     hipMemcpyDeviceToDevice, stream),
                      "copy E");
        last_spec_iters = 0;
        const int nt = (int)tsp.size();
        const unsigned all = (nt >= 32) ? 0xffffffffu : ((1u << nt) - 1u);
        for (int it = 0; it < num_correctors; ++it)
        {
            const Real toff = (it > 0) ? 0.5 * dt : 0.0;
            const Real t = cur_time + toff;
            solve_potential(t, sb_old, nullptr);
            if (do_photo) photo_solve(t);
            photo_in_step = do_photo != 0;
            species_inputs(t, dt, sb_old, nullptr, nullptr);
            photo_in_step = false;
            if (elecenergy_solve)
            {
                species_energy_solve(t, dt, sb_old);
                // compute_electemp_lfa after the energy solve, from the
                // Sborder densities (Source/Evolve.cpp:374-381)
                LfaArgs La{g, state.p, gas_temperature, min_electron_temp, eidx, sbsnap.p};
                if (using_LFA) lfa_te_kernel<<<blocks, 256, 0, stream>>>(La);
                if (using_LFA) lfa_een_kernel<<<blocks, 256, 0, stream>>>(La, state.p + (long long)CMASK_ID * N);
            } else
            {
                LfaArgs La{g, state.p, gas_temperature, min_electron_temp, eidx, state.p};
                if (using_LFA) lfa_te_kernel<<<blocks, 256, 0, stream>>>(La);
                species_solve(dt, all, nullptr);
                if (using_LFA) lfa_een_kernel<<<blocks, 256, 0, stream>>>(La, state.p + (long long)CMASK_ID * N);
            }
            if (do_bg_reactions)
                for (int sp : bg_species)
                    bg_update_kernel<<<blocks, 256, 0, stream>>>(N, state.p, rxn.p, sp, dt, bound_specden,
                                                                min_species_density);
            if (track_current_den) track_currents(dt);
            if (it < num_correctors - 1)
                average_kernel<<<(unsigned)((N * NVAR + 255) / 256), 256, 0, stream>>>(N * NVAR, state.p, state_old.p);
        }
        if (track_current_den && track_integrated_currents)
            for (int q = 0; q < ncurrent_locs; ++q) int_time[q] += dt * (int_cond[q] + int_disp[q]);
        hip_check(hipGetLastError(), "step");
        cur_time += dt;
        istep++;
    }
};

// Case parameters are process-wide (one constant-memory copy per case
// library), so calls of contexts of a parameterised case hold this lock for
// their duration: one thread's parameters cannot change under another's
// launches.  Cases without run-time parameters take no lock.
#if defined(VD_NUM_CASE_PARAMS)
static std::recursive_mutex g_param_mutex;
#define VD_PARAM_LOCK std::lock_guard<std::recursive_mutex> vd_param_lock_(g_param_mutex)
#else
#define VD_PARAM_LOCK (void)0
#endif

#define VD_TRY(ctx, ...)                                                                                           \
    do                                                                                                             \
    {                                                                                                              \
        if (!(ctx)) return VD_E_BADARG;                                                                            \
        VD_PARAM_LOCK;                                                                                             \
        try                                                                                                        \
        {                                                                                                          \
            (ctx)->activate_params();                                                                              \
            __VA_ARGS__;                                                                                           \
            return VD_OK;                                                                                          \
        } catch (BadArg & e)                                                                                       \
        {                                                                                                          \
            (ctx)->err = e.what();                                                                                 \
            return VD_E_BADARG;                                                                                    \
        } catch (io::IoError & e)                                                                                  \
        {                                                                                                          \
            (ctx)->err = e.what();                                                                                 \
            return VD_E_IO;                                                                                        \
        } catch (std::exception & e)                                                                               \
        {                                                                                                          \
            (ctx)->err = e.what();                                                                                 \
            std::string m = e.what();                                                                              \
            if (m.find("non-finite") != std::string::npos) return VD_E_NAN;                                        \
            if (m.find("converge") != std::string::npos) return VD_E_NOCONV;                                       \
            if (m.find("HIP") != std::string::npos) return VD_E_HIP;                                               \
            return VD_E_INTERNAL;                                                                                  \
        }                                                                                                          \
    } while (0)

extern "C" {

int vd_create(const char* inputs_text, vd_ctx** out)
{
    if (!out) return VD_E_BADARG;
    *out = nullptr;
    auto* c = new vd_ctx;
    try
    {
        load_inputs(inputs_text);
        c->read_params();
        c->setup();
    } catch (std::exception& e)
    {
        g_global_error = e.what();
        delete c;
        return (std::string(e.what()).find("HIP") != std::string::npos) ? VD_E_HIP : VD_E_BADARG;
    }
    *out = c;
    return VD_OK;
}

void vd_destroy(vd_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    hipStream_t s = ctx->stream;
    if (g_param_owner == ctx) g_param_owner = nullptr;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
}

const char* vd_last_error(const vd_ctx* ctx) { return ctx ? ctx->err.c_str() : g_global_error.c_str(); }
const char* vd_global_error(void) { return g_global_error.c_str(); }

int vd_get_dims(const vd_ctx* ctx, int dims[8])
{
    if (!ctx || !dims) return VD_E_BADARG;
    dims[0] = ctx->g.n[0];
    dims[1] = ctx->g.n[1];
    dims[2] = ctx->g.n[2];
    dims[3] = NVAR;
    dims[4] = NUM_SPECIES;
    dims[5] = SD;
    dims[6] = ctx->eidx;
    dims[7] = (int)ctx->tsp.size();
    return VD_OK;
}

int vd_init_state(vd_ctx* ctx)
{
    VD_TRY(ctx, ctx->invalidate_adjoint(); ctx->init_data(); hip_check(hipStreamSynchronize(ctx->stream), "init sync"));
}

int vd_set_state(vd_ctx* ctx, const double* host)
{
    if (!host) return VD_E_BADARG;
    VD_TRY(ctx, ctx->invalidate_adjoint();
           hip_check(hipMemcpy(ctx->state.p, host, sizeof(double) * ctx->g.N * NVAR, hipMemcpyHostToDevice),
                     "set_state"));
}

int vd_get_state(vd_ctx* ctx, double* host)
{
    if (!host) return VD_E_BADARG;
    VD_TRY(ctx, hip_check(hipStreamSynchronize(ctx->stream), "sync");
           hip_check(hipMemcpy(host, ctx->state.p, sizeof(double) * ctx->g.N * NVAR, hipMemcpyDeviceToHost),
                     "get_state"));
}

int vd_state_device_ptr(vd_ctx* ctx, double** p)
{
    if (!ctx || !p) return VD_E_BADARG;
    *p = ctx->state.p;
    return VD_OK;
}

int vd_set_time(vd_ctx* ctx, double t)
{
    if (!ctx) return VD_E_BADARG;
    ctx->invalidate_adjoint();
    ctx->cur_time = t;
    return VD_OK;
}

int vd_time_scales(vd_ctx* ctx, double out[3])
{
    if (!out) return VD_E_BADARG;
    if (ctx && !ctx->fine.empty()) VD_TRY(ctx, ctx->time_scales_ml(out));
    VD_TRY(ctx, ctx->time_scales(out));
}

// operator entry points below act on one level; a context with refinement
// levels is advanced through vd_step
static void single_level_only(const vd_ctx* ctx)
{
    if (!ctx->fine.empty()) throw BadArg("this entry point needs a single-level context (use vd_step)");
}

int vd_time_scales_state(vd_ctx* ctx, const double* dev_state, double out[3])
{
    if (!out || !dev_state) return VD_E_BADARG;
    VD_TRY(ctx, single_level_only(ctx); ctx->time_scales(out, dev_state));
}

int vd_fill_ghosts(vd_ctx* ctx, double* dev_fab, int comp0, int ncomp, int ngrow)
{
    if (!dev_fab) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        const Grid& g = ctx->g;
        if (comp0 < 0 || ncomp < 1 || comp0 + ncomp > NVAR)
            throw BadArg("components [comp0, comp0 + ncomp) must lie in [0, " + std::to_string(NVAR) + ")");
        if (ngrow < 1 || ngrow > 8) throw BadArg("ngrow must be in [1, 8]");
        for (int d = 0; d < SD; ++d)
            if (ngrow > g.n[d]) throw BadArg("ngrow exceeds the extent of direction " + std::to_string(d));
        GhostFillArgs A{};
        A.fab = dev_fab;
        A.g = ngrow;
        A.comp0 = comp0;
        for (int d = 0; d < 3; ++d)
        {
            A.n[d] = g.n[d];
            A.e[d] = g.n[d] + ((d < SD) ? 2 * ngrow : 0);
            A.gd[d] = (d < SD) ? ngrow : 0;
            A.per[d] = (d < SD) ? g.periodic[d] : 0;
        }
        A.ncell = (long long)A.e[0] * A.e[1] * A.e[2];
        dim3 grd((unsigned)((A.e[0] + 63) / 64), (unsigned)A.e[1], (unsigned)(A.e[2] * ncomp));
        fill_ghosts_kernel<<<grd, 64, 0, ctx->stream>>>(A);
        hip_check(hipGetLastError(), "fill ghosts");
    });
}

int vd_solve_potential(vd_ctx* ctx, double time, vd_solve_info* info)
{
    VD_TRY(ctx, single_level_only(ctx); ctx->invalidate_adjoint(); ctx->solve_potential(time, ctx->state.p, info);
           hip_check(hipStreamSynchronize(ctx->stream), "sync"));
}

int vd_reaction_source(vd_ctx* ctx, double time, double* host_out)
{
    if (!host_out) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        const long long n = ctx->g.N * (NUM_SPECIES + 1);
        DevBuf<Real> r;
        r.alloc(n);
        fill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(n, r.p, 0.0);
        ctx->species_inputs(time, ctx->fixed_dt, ctx->state.p, nullptr, r.p);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        hip_check(hipMemcpy(host_out, r.p, sizeof(double) * n, hipMemcpyDeviceToHost), "d2h");
    });
}

int vd_explicit_source(vd_ctx* ctx, double time, double dt, double* host_out)
{
    if (!host_out) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        const long long n = ctx->g.N * (NUM_SPECIES + 1);
        DevBuf<Real> e;
        e.alloc(n);
        fill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(n, e.p, 0.0);
        ctx->species_inputs(time, dt, ctx->state.p, e.p, nullptr);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        hip_check(hipMemcpy(host_out, e.p, sizeof(double) * n, hipMemcpyDeviceToHost), "d2h");
    });
}

int vd_poisson_solve(vd_ctx* ctx, const double* dev_rhs, double* dev_phi, const vd_bc* bc, double reltol,
                     double abstol, double* dev_efield_face[3], vd_solve_info* info)
{
    if (!bc) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        if (ctx->mixed) throw BadArg("the operator entry points run the FP64 solvers (vidyut.mixed_precision = 0)");
        ctx->op_solve(ctx->pois, false, 1, dev_rhs, dev_phi, 0.0, 1.0, nullptr, *bc, reltol, abstol, dev_efield_face,
                      -1.0, info);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
    });
}

int vd_helmholtz_solve(vd_ctx* ctx, int ncomp, const double* dev_rhs, double* dev_sol, double alpha, double beta,
                       const double* dev_bcell, const vd_bc* bc, double reltol, double abstol,
                       double* dev_grad_face[3], vd_solve_info* info)
{
    if (!bc) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        ctx->op_solve(ctx->spec, true, ncomp, dev_rhs, dev_sol, alpha, beta, dev_bcell, *bc, reltol, abstol,
                      dev_grad_face, 1.0, info);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
    });
}

int vd_attach_state(vd_ctx* ctx, double* dev_state)
{
    VD_TRY(ctx, {
        single_level_only(ctx);
        ctx->invalidate_adjoint();
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        ctx->state.p = dev_state ? dev_state : ctx->state_own.p;
    });
}

int vd_reaction_src(vd_ctx* ctx, const double* dev_state, double time, double* dev_rxn_src)
{
    if (!dev_state || !dev_rxn_src) return VD_E_BADARG;
    VD_TRY(ctx, single_level_only(ctx); ctx->reaction_src(dev_state, time, dev_rxn_src));
}

int vd_explicit_rhs(vd_ctx* ctx, const double* dev_state, const double* dev_rxn_src, int startspec, int numspec,
                    const int bc_lo[3], const int bc_hi[3], double time, double dt, double* dev_expl_src)
{
    if (!dev_state || !dev_expl_src) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        ctx->check_range(startspec, numspec, bc_lo, bc_hi);
        const int nt = (int)ctx->tsp.size();
        if (numspec > nt) throw BadArg("numspec exceeds the context's face buffers (transported species)");
        Real* Fd[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < SD; ++d)
        {
            // the mixed form keeps FP32 faces: FP64 ones on first use
            if (!ctx->flux[d].p) ctx->flux[d].alloc(ctx->faces_of(d) * nt);
            Fd[d] = ctx->flux[d].p;
        }
        ctx->explicit_rhs(dev_state, dev_rxn_src, startspec, numspec, bc_lo, bc_hi, time, dt, dev_expl_src, Fd);
    });
}

int vd_explicit_rhs_noalloc(vd_ctx* ctx, const double* dev_state, const double* dev_rxn_src, int startspec,
                            int numspec, const int bc_lo[3], const int bc_hi[3], double time, double dt,
                            double* dev_expl_src)
{
    if (!dev_state || !dev_expl_src) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        ctx->check_range(startspec, numspec, bc_lo, bc_hi);
        long long need = 0;
        for (int d = 0; d < SD; ++d) need += ctx->faces_of(d) * numspec;
        if (need > kNoAllocScratchDoubles)
            throw BadArg("the face fluxes of this call need " + std::to_string(need) +
                         " doubles, more than the scratch holds (" + std::to_string(kNoAllocScratchDoubles) + ")");
        if (!ctx->scratch_dev) throw std::runtime_error("no-alloc scratch has no device address");
        // event timing creates events on demand: off for this call
        struct TimingOff
        {
            KTimer& k;
            bool was;
            explicit TimingOff(KTimer& t) : k(t), was(t.on) { k.on = false; }
            ~TimingOff() { k.on = was; }
        } toff(ctx->kt);
        Real* Fd[3] = {nullptr, nullptr, nullptr};
        long long off = 0;
        for (int d = 0; d < SD; ++d)
        {
            Fd[d] = ctx->scratch_dev + off;
            off += ctx->faces_of(d) * numspec;
        }
        ctx->explicit_rhs(dev_state, dev_rxn_src, startspec, numspec, bc_lo, bc_hi, time, dt, dev_expl_src, Fd);
    });
}

int vd_implicit_solve(vd_ctx* ctx, int sp, double time, double dt, const double* host_expl, vd_solve_info* info)
{
    if (!host_expl) return VD_E_BADARG;
    VD_TRY(ctx, {
        single_level_only(ctx);
        ctx->invalidate_adjoint();
        int comp = -1;
        for (size_t c = 0; c < ctx->tsp.size(); ++c)
            if (ctx->tsp[c] == sp) comp = int(c);
        if (comp < 0) throw BadArg("species is not transported");
        const long long N = ctx->g.N;
        // inputs computed as in the step, then the rhs rebuilt from the
        // given explicit source
        ctx->species_inputs(time, dt, ctx->state.p, nullptr, nullptr);
        std::vector<double> rhs(N);
        const double rs = ctx->timescale / ctx->comp_scale(sp);
        std::vector<double> st(N);
        hip_check(hipMemcpy(st.data(), ctx->state.p + (long long)sp * N, sizeof(double) * N, hipMemcpyDeviceToHost),
                  "d2h");
        double bn = 0.0;
        for (long long q = 0; q < N; ++q)
        {
            rhs[q] = ((1.0 / dt) * st[q] + 1.0 * host_expl[(long long)sp * N + q]) * rs;
            bn = std::max(bn, std::fabs(rhs[q]));
        }
        // fold the boundary data again on the fresh rhs: re-run side kernels
        // with a zeroed bnorm slot by uploading the rhs first
        hip_check(hipMemcpy(ctx->spec.f0.p + (long long)comp * N, rhs.data(), sizeof(double) * N, hipMemcpyHostToDevice),
                  "h2d");
        // fold boundary constants of this component
        for (int d = 0; d < SD; ++d)
        {
            if (ctx->g.periodic[d]) continue;
            for (int s = 0; s < 2; ++s)
            {
                auto& sl = ctx->sslab[d][s];
                const long long slab = (long long)sl.n1 * sl.n2;
                std::vector<double> gc(slab), iD(N);
                hip_check(hipMemcpy(gc.data(), sl.c.p + comp * slab, sizeof(double) * slab, hipMemcpyDeviceToHost),
                          "d2h");
                hip_check(hipMemcpy(iD.data(), ctx->spec.iD0.p + (long long)comp * N, sizeof(double) * N,
                                    hipMemcpyDeviceToHost),
                          "d2h");
                const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                const int kd = ctx->spec.side(comp, d, s).kind;
                if (!(kd == BK_DIR || kd == BK_INHNEU || kd == BK_ROBIN)) continue;
                for (int b = 0; b < sl.n2; ++b)
                    for (int a = 0; a < sl.n1; ++a)
                    {
                        int C[3] = {0, 0, 0};
                        C[t1] = a;
                        C[t2] = b;
                        C[d] = s ? ctx->g.n[d] - 1 : 0;
                        long long cell = ctx->g.idx(C[0], C[1], C[2]);
                        double h = ctx->g.dx[d];
                        rhs[cell] -= -ctx->spec.beta * (1.0 / iD[cell]) * gc[a + (long long)sl.n1 * b] / (h * h);
                    }
            }
        }
        hip_check(hipMemcpy(ctx->spec.f0.p + (long long)comp * N, rhs.data(), sizeof(double) * N, hipMemcpyHostToDevice),
                  "h2d");
        std::vector<unsigned long long> bits(32, 0ULL);
        for (size_t c = 0; c < ctx->tsp.size(); ++c) bits[c] = 0;
        double tmp = bn;
        std::memcpy(&bits[comp], &tmp, sizeof(double));
        hip_check(hipMemcpy(ctx->red.p, bits.data(), sizeof(unsigned long long) * 32, hipMemcpyHostToDevice), "h2d");
        if (sp == EEN_ID)
        {
            // the Sborder electron density for the temperature update
            DevBuf<Real> ne;
            ne.alloc(N);
            hip_check(hipMemcpyAsync(ne.p, ctx->state.p + (long long)ctx->eidx * N, sizeof(Real) * N,
                                     hipMemcpyDeviceToDevice, ctx->stream),
                      "ne copy");
            ctx->species_solve(dt, 1u << comp, info);
            energy_te_kernel<<<(unsigned)((N + 255) / 256), 256, 0, ctx->stream>>>(ctx->g, ctx->state.p, ne.p,
                                                                                ctx->min_electron_temp, ctx->eidx);
            hip_check(hipStreamSynchronize(ctx->stream), "sync");
        } else
        {
            ctx->species_solve(dt, 1u << comp, info);
            hip_check(hipStreamSynchronize(ctx->stream), "sync");
        }
    });
}

int vd_step(vd_ctx* ctx, int nsteps, vd_step_info* info)
{
    VD_TRY(ctx, {
        ctx->invalidate_adjoint();
        for (int q = 0; q < nsteps; ++q)
        {
            ctx->step();
            ctx->output_after_step(ctx->fixed_dt);
        }
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        if (info)
        {
            info->time = ctx->cur_time;
            info->dt = ctx->fixed_dt;
            info->poisson_iters = ctx->last_pot_iters;
            info->species_iters = ctx->last_spec_iters;
            info->dt_edrift = ctx->last_ts[0];
            info->dt_ediff = ctx->last_ts[1];
            info->dt_diel_relax = ctx->last_ts[2];
        }
    });
}

int vd_write_plotfile(vd_ctx* ctx, const char* dir)
{
    if (!dir) return VD_E_BADARG;
    VD_TRY(ctx, ctx->write_plotfile(dir));
}

int vd_write_checkpoint(vd_ctx* ctx, const char* dir)
{
    if (!dir) return VD_E_BADARG;
    VD_TRY(ctx, ctx->write_checkpoint(dir));
}

int vd_read_checkpoint(vd_ctx* ctx, const char* dir)
{
    if (!dir) return VD_E_BADARG;
    VD_TRY(ctx, ctx->invalidate_adjoint(); ctx->read_checkpoint(dir); ctx->reset_output_counters());
}

int vd_get_step(const vd_ctx* ctx, int* istep, double* time)
{
    if (!ctx) return VD_E_BADARG;
    if (istep) *istep = ctx->istep;
    if (time) *time = ctx->cur_time;
    return VD_OK;
}

int vd_synchronize(vd_ctx* ctx) { VD_TRY(ctx, hip_check(hipStreamSynchronize(ctx->stream), "sync")); }

void* vd_stream(vd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int vd_kernel_timing(vd_ctx* ctx, long long class_mask)
{
    VD_TRY(ctx, ctx->kt.reset(); ctx->kt.on = (class_mask != 0); ctx->kt.mask = (unsigned long long)class_mask);
}

int vd_kernel_times(vd_ctx* ctx, double* ms_total, long long* counts, int n)
{
    if (!ms_total || !counts) return VD_E_BADARG;
    VD_TRY(ctx, {
        ctx->kt.collect();
        for (int c = 0; c < n && c < K_NCLASS; ++c)
        {
            ms_total[c] = ctx->kt.ms[c];
            counts[c] = ctx->kt.count[c];
        }
    });
}

int vd_kernel_class_count(void) { return K_NCLASS; }
const char* vd_kernel_class_name(int c) { return kclass_name(c); }

int vd_currents(vd_ctx* ctx, double* out, int reset)
{
    if (!ctx || !out) return -VD_E_BADARG;
    for (int l = 0; l < ctx->ncurrent_locs; ++l)
    {
        out[4 * l + 0] = ctx->int_cond[l];
        out[4 * l + 1] = ctx->int_disp[l];
        out[4 * l + 2] = ctx->int_area[l];
        out[4 * l + 3] = ctx->int_time[l];
        if (reset) ctx->int_time[l] = 0.0;
    }
    return ctx->ncurrent_locs;
}

int vd_num_params(const vd_ctx* ctx)
{
    if (!ctx) return -VD_E_BADARG;
#if defined(VD_NUM_CASE_PARAMS)
    return VD_NUM_CASE_PARAMS + 1;
#else
    return 0;
#endif
}

int vd_get_params(vd_ctx* ctx, double* p, int n)
{
    VD_TRY(ctx, ctx->get_params(p, n));
}

int vd_set_params(vd_ctx* ctx, const double* p, int n)
{
    VD_TRY(ctx, ctx->invalidate_adjoint(); ctx->set_params(p, n));
}

int vd_objective_forward(vd_ctx* ctx, int nsteps, int checkpoint_every, double* J)
{
    VD_TRY(ctx, ctx->objective_forward(nsteps, checkpoint_every, J));
}

int vd_adjoint_gradient(vd_ctx* ctx, double* grad, int n)
{
    VD_TRY(ctx, ctx->adjoint_gradient(grad, n));
}

int vd_num_levels(const vd_ctx* ctx)
{
    if (!ctx) return -VD_E_BADARG;
    return 1 + (int)ctx->fine.size();
}

int vd_get_level_box(const vd_ctx* ctx, int lev, int box[6])
{
    if (!ctx || !box || lev < 0 || lev > (int)ctx->fine.size()) return VD_E_BADARG;
    for (int d = 0; d < 3; ++d)
    {
        box[d] = ctx->flo[lev][d];
        box[3 + d] = ctx->fnn[lev][d];
    }
    return VD_OK;
}

int vd_get_level_state(vd_ctx* ctx, int lev, double* host)
{
    if (!ctx || !host || lev < 0 || lev > (int)ctx->fine.size()) return VD_E_BADARG;
    VD_TRY(ctx, {
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        hip_check(hipMemcpy(host, ctx->lstate(lev), sizeof(double) * ctx->lgrid(lev).N * NVAR, hipMemcpyDeviceToHost),
                  "d2h");
    });
}

int vd_set_level_state(vd_ctx* ctx, int lev, const double* host)
{
    if (!ctx || !host || lev < 0 || lev > (int)ctx->fine.size()) return VD_E_BADARG;
    VD_TRY(ctx, {
        ctx->invalidate_adjoint();
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        hip_check(hipMemcpy(ctx->lstate(lev), host, sizeof(double) * ctx->lgrid(lev).N * NVAR, hipMemcpyHostToDevice),
                  "h2d");
    });
}

} // extern "C"
