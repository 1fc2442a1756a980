/*
 * vidyut3d_amd: C ABI of the MI355X-native vidyut3d per-step plasma-fluid
 * path (drift-diffusion explicit fluxes, reaction sources, implicit species
 * diffusion, multigrid Poisson).
 *
 * One shared library is built per case directory (libvd_<case>.so), exactly
 * as the reference builds one executable per case: the case plugin headers
 * (Chemistry.H, UserFunctions.H, UserSources.H, Prob.H, ProbParm.H) are
 * compile-time.  Every entry point returns 0 on success or a VD_E_* code;
 * vd_last_error() gives the message.  No C++ exceptions cross the ABI.
 *
 * State layout: the reference's MultiFab layout of one box without ghost
 * cells, double[NVAR][nz][ny][nx] (x fastest, component outermost), with the
 * component map of VarDefines.H (species, EEN, ETEMP, POT, EFX..EFZ, ...).
 *
 * A context is not thread-safe; use one per thread.  Work is ordered on the
 * context's HIP stream; entry points that return host values synchronise.
 */
#ifndef VIDYUT3D_AMD_H
#define VIDYUT3D_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum
{
    VD_OK = 0,
    VD_E_BADARG = 1,
    VD_E_NOCONV = 2,
    VD_E_HIP = 3,
    VD_E_NAN = 4,
    VD_E_INTERNAL = 5,
    VD_E_IO = 6
};

typedef struct vd_ctx vd_ctx;

typedef struct
{
    int iters;       /* V-cycles used (max over components) */
    double bnorm;    /* |rhs|_inf (max over components) */
    double resid0;   /* initial |r|_inf (max over components) */
    double resid;    /* final |r|_inf (max over components) */
} vd_solve_info;

typedef struct
{
    double time;            /* simulation time after the last step */
    double dt;              /* step size used */
    int poisson_iters;      /* V-cycles of the last potential solve */
    int species_iters;      /* max V-cycles of the last species solves */
    double dt_edrift;       /* electron drift, diffusion and dielectric */
    double dt_ediff;        /* relaxation time scales of the last step */
    double dt_diel_relax;
} vd_step_info;

/* Boundary conditions of one operator solve on the context's grid (the
 * reference's setDomainBC + setLevelBC with its boundary MultiFab,
 * Source/PotentialSolve.cpp:194-207,445-448; Source/ScalarSolve.cpp:573-585,
 * 820-822).  kind[d][s] for direction d and side s (0 lo, 1 hi) takes the
 * inputs' codes: 0 periodic (exactly on periodic directions), 1 Dirichlet,
 * 2 homogeneous Neumann, 3 inhomogeneous Neumann, 4 Robin.  value[d][s] is
 * a device pointer to ncomp * n1 * n2 doubles (the two other directions in
 * increasing order, the first fastest; NULL: zeros): the boundary value for
 * Dirichlet, du/dx_d in the coordinate direction for inhomogeneous Neumann
 * (the reference's convention, Source/PlasmaChem.H:75-200), f for Robin
 * a u + b du/dn = f with a = robin_a[d][s], b = robin_b[d][s] uniform over
 * the side.  Entries of directions beyond the dimension are ignored. */
typedef struct
{
    int kind[3][2];
    const double* value[3][2];
    double robin_a[3][2];
    double robin_b[3][2];
} vd_bc;

/* Create a solver from inputs text in the reference's ParmParse format
 * (keys geometry.*, amr.*, vidyut.*, prob.*), the Vidyut constructor and
 * ReadParameters (Source/Vidyut.cpp:27-218, 358-536) plus amrex_probinit.
 * On failure *out is NULL and vd_global_error() holds the message. */
int vd_create(const char* inputs_text, vd_ctx** out);
void vd_destroy(vd_ctx* ctx);
const char* vd_last_error(const vd_ctx* ctx);
const char* vd_global_error(void);

/* dims[0..7] = nx, ny, nz, NVAR, NUM_SPECIES, AMREX_SPACEDIM, electron
 * index, number of transported species. */
int vd_get_dims(const vd_ctx* ctx, int dims[8]);

/* InitData (Source/Vidyut.cpp:226-256): initdomaindata on the device
 * (Source/GridWork.cpp:97-106) at time 0, or the checkpoint named by
 * amr.restart; checkpoint 0 and the first plotfile when the inputs ask for
 * output (amr.chk_int, amr.plot_int, ...). */
int vd_init_state(vd_ctx* ctx);
/* Host <-> device copies of the whole state (NVAR * ncells doubles). */
int vd_set_state(vd_ctx* ctx, const double* host_state);
int vd_get_state(vd_ctx* ctx, double* host_state);
/* Device pointer of the state (owned by ctx). */
int vd_state_device_ptr(vd_ctx* ctx, double** dev_state);
int vd_set_time(vd_ctx* ctx, double time);

/* find_time_scales (Source/TimeStep.cpp:46-134): out = electron drift,
 * diffusion and dielectric-relaxation time scales. */
int vd_time_scales(vd_ctx* ctx, double out[3]);

/* solve_potential (Source/PotentialSolve.cpp:19-554) with Sborder and
 * Sborder_old taken as the current state; writes POT, EFX.., REF. */
int vd_solve_potential(vd_ctx* ctx, double time, vd_solve_info* info);

/* update_rxnsrc_at_all_levels (Source/ScalarSolve.cpp:156-216): host_out
 * receives (NUM_SPECIES+1)*ncells doubles. */
int vd_reaction_source(vd_ctx* ctx, double time, double* host_out);

/* update_explsrc_at_all_levels (Source/ScalarSolve.cpp:81-154) for every
 * transported species at once: host_out receives (NUM_SPECIES+1)*ncells
 * doubles, -div F + rxn for transported species and 0 elsewhere. */
int vd_explicit_source(vd_ctx* ctx, double time, double dt, double* host_out);

/* implicit_solve_scalar (Source/ScalarSolve.cpp:380-1000) for species sp
 * with a given explicit source ((NUM_SPECIES+1)*ncells doubles, host);
 * updates the state's component sp. */
int vd_implicit_solve(vd_ctx* ctx, int sp, double time, double dt, const double* host_expl, vd_solve_info* info);

/* The MLMG operators of the path on caller device buffers (they replace the
 * reference's MLABecLaplacian + MLMG objects, built per call in
 * Source/PotentialSolve.cpp:194-207,445-510 and Source/ScalarSolve.cpp:
 * 573-585,820-879), on the context's grid and multigrid hierarchy; they
 * synchronise and leave the context's own setup unchanged.
 *
 * vd_poisson_solve: laplacian(phi) = rhs (the potential operator: a = 0,
 * b = -1, B = 1), N doubles each; dev_phi is the initial guess on entry (the
 * reference starts from zero, Source/PotentialSolve.cpp:481-483) and the
 * solution on return.  dev_efield_face[d] (each nullable): E_d = -dphi/dx_d
 * on the faces of direction d, (n_d + 1) faces along d times the other
 * extents, x fastest (getGradSolution times -1, :500-510).  Stopping test:
 * |r| <= max(abstol, reltol * max(|rhs|, |r0|)).  FP64 contexts only.
 *
 * vd_helmholtz_solve: alpha u - beta div(b grad u) = rhs for ncomp
 * components (at most the transported-species count), b the harmonic face
 * mean of the cell coefficients dev_bcell (ncomp * N, Source/ScalarSolve.cpp:
 * 658-690); dev_sol holds the initial guess, then the solution (component c
 * at c * N).  dev_grad_face[d] (nullable): du/dx_d on faces, ncomp blocks. */
int vd_poisson_solve(vd_ctx* ctx, const double* dev_rhs, double* dev_phi, const vd_bc* bc, double reltol,
                     double abstol, double* dev_efield_face[3], vd_solve_info* info);
int vd_helmholtz_solve(vd_ctx* ctx, int ncomp, const double* dev_rhs, double* dev_sol, double alpha, double beta,
                       const double* dev_bcell, const vd_bc* bc, double reltol, double abstol,
                       double* dev_grad_face[3], vd_solve_info* info);

/* Caller-owned state (the reference's phi_new MultiFab, owned by the Vidyut
 * object in Source/Vidyut.H; here the caller may own it): the context
 * advances the device buffer dev_state (NVAR * ncells doubles, the state
 * layout above) in place from now on, without copying it; NULL returns to
 * the context's own buffer.  The caller keeps ownership and keeps the buffer
 * alive while attached.  Single-level contexts. */
int vd_attach_state(vd_ctx* ctx, double* dev_state);

/* Device-resident operators on caller buffers (device pointers, state
 * layout; stream-ordered on vd_stream(ctx), no synchronisation).
 *
 * vd_reaction_src: update_rxnsrc_at_all_levels (Source/ScalarSolve.cpp:
 * 156-216) of dev_state into dev_rxn_src, (NUM_SPECIES+1)*ncells doubles:
 * the case's CKWC rates times N_A, the energy exchange in the last
 * component, plus add_user_react_sources.
 *
 * vd_explicit_rhs: update_explsrc_at_all_levels (Source/ScalarSolve.cpp:
 * 81-154) for species [startspec, startspec + numspec) with the side codes
 * bc_lo/bc_hi of that species class (0 periodic, 1 Dirichlet, 2 homogeneous
 * Neumann, 3 inhomogeneous Neumann, 4 Robin; the flux is zero at code-3
 * sides, Source/compute_explicit_flux.H:401-414): face fluxes
 * (compute_scalar_transport_flux, :218-329) then compute_dsdt (:17-79),
 * dev_expl_src = -div F + rxn for the range and 0 for the other
 * NUM_SPECIES+1 components.  dev_rxn_src is the reference's rxn_src input;
 * NULL evaluates the reactions in the same pass.  The face fluxes live in
 * the context's face buffers; numspec is at most its transported-species
 * count. */
int vd_reaction_src(vd_ctx* ctx, const double* dev_state, double time, double* dev_rxn_src);
int vd_explicit_rhs(vd_ctx* ctx, const double* dev_state, const double* dev_rxn_src, int startspec, int numspec,
                    const int bc_lo[3], const int bc_hi[3], double time, double dt, double* dev_expl_src);

/* find_time_scales (Source/TimeStep.cpp:46-134) of a caller's device state
 * (state layout, ngrow 0): out = electron drift, diffusion and
 * dielectric-relaxation time scales.  Synchronises (host scalars). */
int vd_time_scales_state(vd_ctx* ctx, const double* dev_state, double out[3]);

/* Ghost fill of a caller's ghosted FAB (the Sborder MultiFab filled by
 * FillPatch, Source/GridWork.cpp:135-175, with the reference's bcspec,
 * Source/Vidyut.cpp:118-130): dev_fab holds NVAR components of
 * (n_d + 2 ngrow) cells in each of the case's AMREX_SPACEDIM directions
 * (1 cell in the others), x fastest, component outermost, valid cells at
 * offset ngrow.  The ghosts of components [comp0, comp0 + ncomp) take the
 * periodic image in periodic directions and the nearest valid cell
 * (foextrap) in the others, per direction, so edges and corners take the
 * valid cell with every index wrapped or clamped.  1 <= ngrow <= 8 and at
 * most each extent.  Stream-ordered, no synchronisation.  The device
 * operators above read the unghosted state (ghosts are index maps); this
 * serves callers that keep ghosted FABs for kernels of their own. */
int vd_fill_ghosts(vd_ctx* ctx, double* dev_fab, int comp0, int ncomp, int ngrow);

/* vd_explicit_rhs without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same arguments, kernels and results as
 * vd_explicit_rhs; the face fluxes go to a file-scope device scratch array
 * of 64 MiB per GPU (8M doubles: sum over directions of faces x numspec must
 * fit, else VD_E_BADARG), held until the process ends.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (other streams or
 * contexts) race.  Kernel timing (vd_kernel_timing) is off during the call. */
int vd_explicit_rhs_noalloc(vd_ctx* ctx, const double* dev_state, const double* dev_rxn_src, int startspec,
                            int numspec, const int bc_lo[3], const int bc_hi[3], double time, double dt,
                            double* dev_expl_src);

/* nsteps coarse steps of Vidyut::Evolve (Source/Evolve.cpp:69-714).  With
 * amr.plot_int / amr.plot_time (amr.chk_int / amr.chk_time) in the inputs a
 * plotfile (checkpoint) named amr.plot_file (amr.chk_file) plus a five-digit
 * counter is written after the steps where Evolve writes one
 * (Source/Evolve.cpp:635-676); the final plotfile Evolve writes when its
 * whole run ends (:716) is the caller's. */
int vd_step(vd_ctx* ctx, int nsteps, vd_step_info* info);

/* Plotfile of every level (Vidyut::WritePlotFile, Source/Plot.cpp:32-44:
 * amrex::WriteMultiLevelPlotfile with the variable names of
 * Source/Vidyut.cpp:34-74, time t_new[0], level steps istep, refinement
 * ratio 2): AMReX's plotfile layout (Header, Level_<l>/Cell_H,
 * Level_<l>/Cell_D_00000; IEEE doubles, little-endian).  An existing
 * directory of that name is renamed to <dir>.old.<pid>.<n> first.
 * VD_E_IO on file errors.  Synchronises. */
int vd_write_plotfile(vd_ctx* ctx, const char* dir);

/* Checkpoint (Vidyut::WriteCheckpointFile, Source/Plot.cpp:46-135): the
 * reference's Header (title, finest level, istep, dt and t_new per level,
 * BoxArrays) and one VisMF MultiFab Level_<l>/phi per level. */
int vd_write_checkpoint(vd_ctx* ctx, const char* dir);

/* Restart (Vidyut::ReadCheckpointFile, Source/Plot.cpp:137-225): the state
 * of every level, istep and t_new[0] from a checkpoint written by this
 * library or by the reference (any number of boxes and data files per
 * level, either byte order; the boxes must cover the context's box of each
 * level and the level count must be the context's).  vd_init_state does the
 * same when the inputs name a checkpoint in amr.restart, as InitData does. */
int vd_read_checkpoint(vd_ctx* ctx, const char* dir);

/* Steps taken (istep[0]) and the simulation time (t_new[0]). */
int vd_get_step(const vd_ctx* ctx, int* istep, double* time);

/* Wait for all queued work of ctx. */
int vd_synchronize(vd_ctx* ctx);

/* HIP stream of ctx (hipStream_t), for event timing by callers. */
void* vd_stream(vd_ctx* ctx);

/* Per-kernel-class timing with HIP events on the context's stream (no
 * counterpart in the reference, whose profiling is BL_PROFILE scopes,
 * Source/Evolve.cpp:19,71).  vd_kernel_timing(ctx, mask) resets and times
 * the classes whose bit is set in mask (0 disables, -1 times all);
 * vd_kernel_times fills total milliseconds and launch counts of the first n
 * classes (synchronises); class names from vd_kernel_class_name. */
int vd_kernel_timing(vd_ctx* ctx, long long class_mask);
int vd_kernel_times(vd_ctx* ctx, double* ms_total, long long* counts, int n);
int vd_kernel_class_count(void);
const char* vd_kernel_class_name(int cls);

/* Electrode currents (vidyut.track_current_den and
 * vidyut.track_integrated_currents with vidyut.current_loc_surfaces):
 * compute_current_den, compute_disp_current_den and
 * compute_integrated_currents (Source/CurrentDens.cpp:17-456) run after
 * every step; the state's ECUR/ICUR/DCUR components hold the current
 * densities.  out[4*l + 0..3] = conduction current, displacement current and
 * collector area of location l after the last step, and the time integral
 * (sum over steps of dt times conduction + displacement current) since the
 * last reset (reset != 0 zeroes it after reading).  Returns the number of
 * locations (at most 10) or a negative VD_E_* code. */
int vd_currents(vd_ctx* ctx, double* out, int reset);

/* Run-time parameters (no counterpart in the reference, whose case constants
 * are compile-time): the case's own (case_param(i) in its plugin headers,
 * VD_NUM_CASE_PARAMS of them) followed by the voltage amplitude of the
 * electrode named by current_loc_surfaces[0].  vd_num_params returns their
 * number (0 when the case has none). */
int vd_num_params(const vd_ctx* ctx);
int vd_get_params(vd_ctx* ctx, double* p, int n);
int vd_set_params(vd_ctx* ctx, const double* p, int n);

/* Objective and discrete adjoint (no counterpart in the reference).
 * vd_objective_forward runs nsteps steps from the current state and time,
 * keeping a state checkpoint every checkpoint_every steps, and returns
 * *J = sum over the steps of dt * (conduction + displacement current) through
 * current location 0 (the currents of Source/CurrentDens.cpp:343-456).
 * vd_adjoint_gradient then returns dJ/dp for the vd_num_params parameters
 * (case parameters, then the electrode's voltage amplitude): it recomputes
 * each segment from its checkpoint and runs the hand-written adjoint step
 * kernels backwards; afterwards the state is again the forward's final state.
 * Supported: upwind fluxes, one corrector, FP64, table chemistry with
 * state-independent transport, homogeneous Neumann or periodic species
 * sides, Dirichlet/Neumann/periodic potential sides; VD_E_BADARG otherwise. */
int vd_objective_forward(vd_ctx* ctx, int nsteps, int checkpoint_every, double* J);
int vd_adjoint_gradient(vd_ctx* ctx, double* grad, int n);

/* Static refinement levels (amr.max_level > 0, ratio 2, each level the
 * central half of its parent per direction, no regridding): the state of
 * level lev is the reference's phi_new[lev] (Source/Vidyut.H, one box per
 * level here).  vd_step advances all levels with the coarse dt
 * (Source/Evolve.cpp:69-714 with finest_level > 0).  vd_num_levels returns
 * finest_level + 1; vd_get_level_box gives the box origin in global level
 * indices (box[0..2]) and its extents (box[3..5]); the level state is
 * NVAR * (box cells) doubles in the same layout as vd_get_state.  Level 0 is
 * also reachable through vd_get_state / vd_set_state. */
int vd_num_levels(const vd_ctx* ctx);
int vd_get_level_box(const vd_ctx* ctx, int lev, int box[6]);
int vd_get_level_state(vd_ctx* ctx, int lev, double* host_state);
int vd_set_level_state(vd_ctx* ctx, int lev, const double* host_state);

#ifdef __cplusplus
}
#endif
#endif
