/*
 * mahyco_gpu.h -- C-ABI of the MI355X (gfx950) MaHyCo Lagrange + Euler-ADI
 * remap time step.
 *
 * Everything here is plain C: pointers, sizes, int status codes.  No HIP or
 * torch type crosses the boundary.  A context owns all device memory; the
 * caller owns every host array it passes and the library never keeps a host
 * pointer after a call returns.  A context is not thread-safe.  All calls are
 * host-synchronous unless their name ends in _async.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   - the MahycoModule entry points of the compute loop
 *       mahyco/src/Mahyco.config:30-41, declared in mahyco/src/Mahyco.axl:150-168
 *       and mahyco/src/MahycoModule.h:145-381;
 *   - the IRemap service (mahyco/src/Remap/IRemap.h:32-65) with its RemapADI
 *     implementation (mahyco/src/Remap/RemapADIService.cc:19-60);
 *   - the IEquationOfState services (mahyco/src/eos/IEquationOfState.h:32-65):
 *     PerfectGas (mahyco/src/eos/perfectgas/PerfectGasEOSService.cc:20-137) and
 *     StiffenedGas (mahyco/src/eos/stiffenedgas/StiffenedGasEOSService.cc:10-59);
 *   - the option surface of Mahyco.axl:170-340, RemapADI.axl:136-178,
 *     PerfectGasEOS.axl:99-105.
 *
 * Error behaviour: the reference calls exit(1) when dt falls below deltat-min
 * (mahyco/src/MahycoModule.cc:2152-2159) and stopComputeLoop(true) at the end
 * of the run (MahycoModule.cc:2177-2183).  Here these become the return codes
 * MHG_E_DT_MIN and MHG_STOP.
 */
#ifndef MAHYCO_GPU_H
#define MAHYCO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHG_ABI_VERSION 1
#define MHG_MAX_ENV 8
#define MHG_MAX_BC 16

/* Status codes. */
enum {
  MHG_OK = 0,
  MHG_STOP = 1,           /* end of run reached (stopComputeLoop)           */
  MHG_E_ARG = -1,         /* bad argument / inconsistent sizes              */
  MHG_E_DT_MIN = -2,      /* dt < deltat-min (reference exits with code 1)  */
  MHG_E_HIP = -3,         /* HIP runtime error                              */
  MHG_E_STATE = -4,       /* call out of order (e.g. step before init)      */
  MHG_E_UNSUPPORTED = -5, /* option value outside the GPU path              */
  MHG_E_ALLOC = -6        /* device allocation failed                       */
};

/* Cartesian mesh as built by Arcane's cartesian generator
 * (mahyco/data/Data.1.arc:13-24): n[d] cells of uniform size length[d]/n[d]
 * from origin[d].  n[2] and the z entries are ignored when dim == 2. */
typedef struct {
  int32_t dim;
  int32_t n[3];
  double origin[3];
  double length[3];
} mhg_mesh_desc;

/* Boundary conditions (Mahyco.axl:312-329, TypesMahyco.h:11-17).  The face
 * groups are the ones PrepareFaceGroup builds (MahycoAnnexe.cc:183-251). */
enum { MHG_BC_VX = 0, MHG_BC_VY = 1, MHG_BC_VZ = 2 };
enum {
  MHG_SURF_XMIN = 0, MHG_SURF_XMAX = 1, MHG_SURF_YMIN = 2,
  MHG_SURF_YMAX = 3, MHG_SURF_ZMIN = 4, MHG_SURF_ZMAX = 5
};
typedef struct {
  int32_t surface;
  int32_t type;
  double value;
} mhg_bc;

/* Equation of state of one environment (<eos-model name=...>). */
enum { MHG_EOS_PERFECT_GAS = 0, MHG_EOS_STIFFENED_GAS = 1 };
typedef struct {
  int32_t eos;
  double adiabatic_cst;  /* PerfectGasEOS.axl:101 default 1.4 */
  double limit_tension;  /* PerfectGasEOS.axl:104 default 0.0 */
} mhg_env_desc;

/* Characteristic length (Mahyco.axl "longueur-caracteristique"). */
enum { MHG_LC_RACINE_CUBIQUE_VOLUME = 0, MHG_LC_FACES_OPPOSEES = 1 };

/* Limiters (TypesMahyco.h:65-74).  The generalised limiters MHG_LIM_MINMODG
 * .. MHG_LIM_ULTRABEEG (fluxLimiterG, UtilesRemap.cc:127-173) take the cell
 * widths along the sweep into account (h_cell, RemapADIService.cc:223-255, and
 * the dual widths of UtilesRemap.cc:46-57); they run on 2D (single and
 * batched) and 3D meshes, where the reference runs them on the host only.
 * MHG_LIM_LEARNED is an extension with no
 * counterpart in the reference: the cell sweeps take the slope
 * alpha_dir(c) * (g_back + g_front) / 2, with alpha_dir(c) a per-cell,
 * per-direction coefficient, and fall back to the minmod-limited slope
 * wherever that slope puts a face value of the cell outside the min/max of
 * the cell and its neighbours along the sweep, and wherever alpha_dir(c) < 0;
 * the dual sweeps use minmod. */
enum { MHG_LIM_MINMOD = 0, MHG_LIM_SUPERBEE = 1, MHG_LIM_VANLEER = 2, MHG_LIM_MINMODG = 3,
       MHG_LIM_SUPERBEEG = 4, MHG_LIM_VANLEERG = 5, MHG_LIM_ARITHMETICG = 6, MHG_LIM_ULTRABEEG = 7,
       MHG_LIM_LEARNED = 64, MHG_LIM_LEARNED_ALL = 65 };

/* Arithmetic of the bulk fields. */
enum { MHG_PREC_FP64 = 0, MHG_PREC_FP32_BULK = 1 };

typedef struct {
  /* Mahyco.axl */
  double cfl;               /* default 0.3   */
  double threshold;         /* default 1e-10 */
  double deltat_init;
  double deltat_min;
  double deltat_max;
  double final_time;
  int32_t with_projection;  /* default 1; 0: Lagrange only (no remap, the mesh
                               moves; FP64 single contexts, run through the
                               entry-point kernels) */
  int32_t pseudo_centree;   /* default 0 */
  int32_t schema_csts;      /* must be 0 (host-only path in the reference) */
  int32_t with_newton;      /* must be 0 (host-only path in the reference) */
  int32_t sans_lagrange;    /* must be 0 */
  int32_t longueur_caracteristique;
  int32_t nb_bc;
  mhg_bc bc[MHG_MAX_BC];
  int32_t nb_env;
  mhg_env_desc env[MHG_MAX_ENV];
  /* RemapADI.axl */
  double remap_threshold;             /* default 1e-10 */
  int32_t is_euler_scheme;            /* must be 1     */
  int32_t ordre_projection;           /* 1, 2 or 3 (3: not learned, not in batches) */
  int32_t projection_limiteur_id;     /* MHG_LIM_MINMOD .. MHG_LIM_ULTRABEEG, or MHG_LIM_LEARNED[_ALL] */
  int32_t projection_pente_borne;     /* default 0; 1: pente-borne fluxes
                                         (order 2, classical or generalised
                                         limiter; FP64 single contexts) */
  int32_t conservation_energie_totale;/* total-energy fix-up of the remap
                                         (RemapADIFinal.cc:751-776); not in
                                         batches */
  /* Extensions (no counterpart in the reference). */
  double visc_quad_coeff;   /* < 0: reference value (gamma+1)/2 */
  int32_t precision;        /* MHG_PREC_FP64 (the drop-in) or MHG_PREC_FP32_BULK */
  int32_t land_on_final_time; /* 1: the last step is shortened to end at final_time
                                 (0, the reference: the run stops before the step
                                 that would pass it, MahycoModule.cc:2177-2183) */
  /* RemapADI.axl, pente-borne options (RemapADIService.cc:319-431) */
  int32_t projection_pente_borne_mixte;     /* default 0; 1: cells with a pure
                                               neighbourhood along the sweep use
                                               projection_limiteur_pure_id's
                                               pente-borne reconstruction */
  int32_t projection_limiteur_pure_id;      /* default 0 */
  int32_t projection_pente_borne_debar_fix; /* default 0; 0, 1 or 2 */
  int32_t reserved[3];
} mhg_options;

/* Full per-step state, structure-of-arrays, caller-owned host memory.
 * Any pointer may be NULL (that field is skipped).  Sizes: N_c cells,
 * N_n nodes (see mhg_sizes).  "env" arrays are [nb_env][N_c] and only
 * meaningful where the cell holds two or more environments; for a cell with
 * one environment the environment value is the cell value (Arcane's pure-cell
 * aliasing, mahyco/src/accenv/MultiEnvUtils.h:303-440). */
typedef struct {
  double *density, *internal_energy, *pressure, *sound_speed;
  double *pseudo_viscosity, *cell_mass;
  double *fracvol, *mass_fraction;
  int32_t *env_mask;                 /* bit k <=> cell holds environment k */
  double *density_env, *internal_energy_env, *pressure_env, *sound_speed_env;
  double *pseudo_viscosity_env, *cell_mass_env, *fracvol_env, *mass_fraction_env;
  double *velocity;                  /* [dim][N_n] */
  double *node_mass;                 /* [N_n]      */
  double *node_coord;                /* [dim][N_n] coordinates at the end of the last step */
  double *dual_grad_phi;             /* [dim][N_n] persistent limited nodal gradient      */
  double *caracteristic_length;      /* [N_c] L of the last Lagrange phase (output only)  */
  double global_deltat, old_deltat, global_time;
  int32_t sens_projection, global_iteration;
} mhg_state;

typedef struct mhg_ctx mhg_ctx;

/* ---- life cycle ------------------------------------------------------- */
int mhg_abi_version(void);
void mhg_default_options(mhg_options* opt);
int mhg_create(const mhg_mesh_desc* mesh, const mhg_options* opt, mhg_ctx** out);
int mhg_destroy(mhg_ctx* ctx);
const char* mhg_last_error(const mhg_ctx* ctx);
int mhg_sizes(const mhg_ctx* ctx, int64_t* n_cells, int64_t* n_nodes);

/* ---- initialisation (hydroStartInit + ComputeCellMass + ComputeNodeMass,
 *      mahyco/src/MahycoModule.cc:48-266) ------------------------------- */
/* SODService (mahyco/src/casTest/SOD/SODService.cc:3-135): cas_test 3/4/5
 * mono-material tube along x/y/z, 13/14/15 two-material tube. */
int mhg_init_sod(mhg_ctx* ctx, int32_t cas_test);
/* Three-environment triple point (TypesMahyco.h:26 names it; the reference
 * has no service for it): x<1: env 0; x>1,y<1.5: env 1; x>1,y>1.5: env 2. */
int mhg_init_triple_point(mhg_ctx* ctx, const double rho[3], const double p[3]);
/* Generic pure-cell initial state: env index, density and pressure per cell,
 * velocity [dim][N_n] (NULL = at rest). */
int mhg_init_cells(mhg_ctx* ctx, const int32_t* env_of_cell, const double* density,
                   const double* pressure, const double* velocity);

int mhg_upload_state(mhg_ctx* ctx, const mhg_state* st);
int mhg_download_state(mhg_ctx* ctx, mhg_state* st);

/* ---- compute loop, one call per Mahyco.config entry point ------------ */
/* 2D and 3D contexts in FP64 (MHG_E_UNSUPPORTED in the precision form). */
/* Arcane's time-loop manager increments time by dt and the iteration count
 * before the compute-loop entry points of an iteration; a standalone caller
 * does it with this call (an Arcane shim would not call it). */
int mhg_begin_iteration(mhg_ctx* ctx);
int mhg_save_values_at_n(mhg_ctx* ctx);
int mhg_update_velocity(mhg_ctx* ctx);
int mhg_apply_boundary_condition(mhg_ctx* ctx);
int mhg_update_position(mhg_ctx* ctx);
int mhg_compute_geometric_values(mhg_ctx* ctx);
int mhg_update_density(mhg_ctx* ctx);
int mhg_compute_artificial_viscosity(mhg_ctx* ctx);
/* updateEnergyAndPressure (MahycoModule.cc:1439-1478); with
 * with_projection = 0 it also applies every EOS and the pressure average. */
int mhg_update_energy_and_pressure(mhg_ctx* ctx);
/* remap (PrepareRemap.cc:448-563) fused: mhg_remap_prepare, mhg_remap_appli,
 * mhg_eos_apply of every environment and mhg_compute_pression_moyenne in one
 * call.  With with_projection = 0 it does nothing (the reference returns at
 * once) but close the iteration's Lagrange phase. */
int mhg_remap(mhg_ctx* ctx);
/* computePressionMoyenne (MahycoModule.cc:1836-1915): cells holding several
 * environments take p = sum f_k p_k and c = max(1e-20, max c_k). */
int mhg_compute_pression_moyenne(mhg_ctx* ctx);
/* computeDeltaT (MahycoModule.cc:2027-2191) from the current state. */
int mhg_compute_delta_t(mhg_ctx* ctx, double* new_dt, int32_t* stop);

/* ---- IRemap (mahyco/src/Remap/IRemap.h:32-65), RemapADI ---------------- */
/* computeVariablesForRemap + computeFaceQuantitesForRemap
 * (PrepareRemap.cc:15-437), the MahycoModule side of the remap: the remap
 * vector of every cell from the Lagrangian state; the face quantities are
 * formed inside the sweeps. */
int mhg_remap_prepare(mhg_ctx* ctx);
/* appliRemap (IRemap.h:46, RemapADIService.cc:19-60): the cell and dual
 * sweeps in SensProjection order, then remapVariables.  Pressures and sound
 * speeds are the EOS's: call mhg_eos_apply per environment and
 * mhg_compute_pression_moyenne next.  dimension, with_dual_projection,
 * nb_vars_to_project (3 nb_env + 5) and nb_env must match the context
 * (MHG_E_ARG otherwise). */
int mhg_remap_appli(mhg_ctx* ctx, int32_t dimension, int32_t with_dual_projection,
                    int32_t nb_vars_to_project, int32_t nb_env);
/* remapVariables (IRemap.h:64, RemapADIFinal.cc:19-778) from the remap
 * vectors of this iteration's mhg_remap_appli, which already ends with it. */
int mhg_remap_variables(mhg_ctx* ctx, int32_t dimension, int32_t with_dual_projection,
                        int32_t nb_vars_to_project, int32_t nb_env);
/* resizeRemapVariables (IRemap.h:50): the context sizes its remap arrays at
 * creation; the call checks the counts (MHG_E_ARG if they differ). */
int mhg_remap_resize_variables(mhg_ctx* ctx, int32_t nb_vars_to_project, int32_t nb_env);
/* synchronizeUremap (IRemap.h:54): ghost-cell exchange; a context holds the
 * whole mesh on one device, so it has none to exchange. */
int mhg_remap_synchronize_uremap(mhg_ctx* ctx);
int mhg_remap_get_ordre_projection(const mhg_ctx* ctx);
int mhg_remap_has_projection_pente_borne(const mhg_ctx* ctx);
int mhg_remap_has_conservation_energie_totale(const mhg_ctx* ctx);
int mhg_remap_is_euler(const mhg_ctx* ctx);

/* ---- IEquationOfState (mahyco/src/eos/IEquationOfState.h:32-65) ------- */
/* Per environment env (its <eos-model>: PerfectGas,
 * PerfectGasEOSService.cc:20-137, or StiffenedGas,
 * StiffenedGasEOSService.cc:10-59), on the cells holding it (the cell values
 * of a cell holding env alone, the environment values of a mixed cell).
 * FP64 contexts. */
/* initEOS: internal energy and sound speed from density and pressure. */
int mhg_eos_init(mhg_ctx* ctx, int32_t env);
/* applyEOS: pressure and sound speed from density and internal energy. */
int mhg_eos_apply(mhg_ctx* ctx, int32_t env);
/* applyOneCellEOS: the same for one cell. */
int mhg_eos_apply_one_cell(mhg_ctx* ctx, int32_t env, int64_t cell);
int mhg_eos_get_adiabatic_cst(const mhg_ctx* ctx, int32_t env, double* value);
int mhg_eos_get_tension_limit_cst(const mhg_ctx* ctx, int32_t env, double* value);

/* ---- fused fast path ---------------------------------------------------- */
/* One full iteration (all entry points of Mahyco.config:30-41), fused. */
int mhg_step(mhg_ctx* ctx, int32_t* stop);
/* Up to max_steps fused iterations with dt kept on the device; stops early
 * at the end of the run.  *steps_done receives the number executed. */
int mhg_run(mhg_ctx* ctx, int32_t max_steps, int32_t* steps_done);
/* Enqueue n fused iterations on the context's stream without host sync. */
int mhg_run_async(mhg_ctx* ctx, int32_t n_steps);
int mhg_synchronize(mhg_ctx* ctx);

/* ---- tuning (extension, no counterpart in the reference) --------------- */
/* Times every launch variant of the tunable kernels on this context's mesh
 * (reps launches each) and keeps the fastest; the simulation state is not
 * changed.  report (may be NULL) receives the per-variant timings. */
int mhg_autotune(mhg_ctx* ctx, int32_t reps, char* report, int32_t report_len);
/* Number of launch variants of a tunable kernel ("cell_sweep_y",
 * "node_sweep_x", "node_sweep_y"; "step_graph": direct launches or graph
 * replay; "final_fusion": the fused step's final cell phase as a kernel of
 * its own or inside the last cell sweep; "lagrange_fusion": its Lagrange
 * cell phase as a kernel of its own or inside the first cell sweep;
 * "store_policy": which of the step kernels' outputs are written with
 * streaming stores, one bit per kind of array (1 the cell state, 2 the remap
 * vector between the cell sweeps, 4 the dual fluxes and dual state, 8 node
 * velocities and coordinates; 0 all ordinary stores, which stay in the
 * caches for the next kernel, 15 all streaming; the default follows the
 * working set: above 320 MB, 1 in FP32-bulk and 3 in FP64, else 0);
 * "ytile_width": which y sweeps run on LDS tiles of 8 columns
 * instead of 16, one bit per kind (1 the first cell sweep with the Lagrange
 * cell phase, 2 the last with the final cell phase, 4 the dual sweeps; the
 * default 6 in FP64, 7 in FP32-bulk)), 0 for an unknown name. */
int32_t mhg_variant_count(const char* kernel);
/* Selects launch variant `variant` of a tunable kernel for later steps
 * (MHG_E_ARG for an unknown kernel or an index out of range). */
int mhg_select_variant(mhg_ctx* ctx, const char* kernel, int32_t variant);

/* ---- batching form (extension, no counterpart in the reference) ------- */
/* n independent 2D problems of uneven sizes advanced together.  Each problem
 * runs the fused iteration of mhg_step on its own mesh with its own options
 * (final time, cfl, dt bounds, EOS constants, boundary conditions), dt, time,
 * iteration count and stop test, bit for bit as a context of its own would.
 * One launch of each persistent kernel covers every running problem; the
 * iterations are replayed from a captured HIP graph; problems that stop drop
 * out of the kernels' work lists.  All problems share dim = 2, the precision,
 * nb_env (at most 3), projection_limiteur_id, ordre_projection and
 * pseudo_centree. */
typedef struct mhg_batch mhg_batch;
int mhg_batch_create(int32_t n, const mhg_mesh_desc* meshes, const mhg_options* opts, mhg_batch** out);
int mhg_batch_destroy(mhg_batch* b);
const char* mhg_batch_last_error(const mhg_batch* b);
/* The context of problem i, owned by the batch.  mhg_init_*, mhg_upload_state
 * (before the first mhg_batch_run), mhg_download_state, mhg_sizes and
 * mhg_conservation work on it; the stepping and tuning calls return
 * MHG_E_STATE. */
mhg_ctx* mhg_batch_problem(mhg_batch* b, int32_t i);
/* Advances every running problem until it stops, or for at most max_steps
 * iterations of the batch.  *n_running receives the problems still running,
 * *zone_cycles the cells x iterations done by all problems in this call.
 * Returns MHG_STOP when no problem is left running, MHG_E_DT_MIN if a problem
 * stopped on deltat-min (mhg_batch_status names it). */
int mhg_batch_run(mhg_batch* b, int32_t max_steps, int32_t* n_running, int64_t* zone_cycles);
/* Problem i: iterations done, time, dt, and 0 running / 1 stopped at the end
 * of its run / 2 stopped on deltat-min. */
int mhg_batch_status(mhg_batch* b, int32_t i, int32_t* iterations, double* time, double* dt, int32_t* state);
/* Times the launch shapes of the batch's sweep kernels (persistent grid size;
 * for the y sweeps also the column-segment or 64-row tile form) on the
 * batch's meshes and keeps the fastest; the simulation state is unchanged. */
int mhg_batch_autotune(mhg_batch* b, int32_t reps, char* report, int32_t report_len);

/* ---- learned limiter (extension, no counterpart in the reference) ----- */
/* Fixed coefficients alpha [2][N_c] (x then y) of a context created with
 * projection_limiteur_id = MHG_LIM_LEARNED; until set they are 0. */
int mhg_set_learned_alpha(mhg_ctx* ctx, const double* alpha);
/* The coefficients the next step uses: with a network attached, computed
 * from the current state; otherwise the fixed ones.  alpha [2][N_c]. */
int mhg_learned_alpha(mhg_ctx* ctx, double* alpha);

/* The network: a 128-64-64-2 perceptron with ReLU, BF16 weights and FP32
 * accumulation on the matrix cores.  Input: a 5 x 5 stencil of density,
 * specific internal energy, cell-mean x and y velocity and largest
 * environment volume fraction, each channel centred on the cell's value and
 * divided by its largest deviation over the stencil (cells beyond the domain
 * repeat the boundary cell), channel-major, zero-padded to 128.  Output:
 * alpha_x, alpha_y, clamped to [0, 2]; with MHG_LIM_LEARNED a cell whose
 * 3 x 3 neighbourhood is not one and the same single environment gets -1
 * (minmod at material interfaces), with MHG_LIM_LEARNED_ALL every cell takes
 * the network's value.  Parameters, FP32:
 * W1 [64][128], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2]. */
typedef struct mhg_mlp mhg_mlp;
int mhg_mlp_create(mhg_mlp** out);
int mhg_mlp_destroy(mhg_mlp* net);
const char* mhg_mlp_last_error(const mhg_mlp* net);
int32_t mhg_mlp_param_count(void);
int mhg_mlp_set_params(mhg_mlp* net, const float* params);  /* also resets the optimiser */
int mhg_mlp_get_params(mhg_mlp* net, float* params);
/* Y [m][2] of X [m][128] (rounded to BF16). */
int mhg_mlp_forward(mhg_mlp* net, const float* X, int64_t m, float* Y);
/* Training samples kept on the device: X [m][128], targets T [m][16] =
 * for direction d and channel k < 4 (density, energy, x and y velocity), the
 * normalised central slope and the normalised reference slope; the loss is
 * sum over d, k of (alpha_d sc - st)^2, averaged over the samples. */
int mhg_mlp_set_data(mhg_mlp* net, const float* X, const float* T, int64_t m);
/* Loss and gradient over samples [row0, row0 + m) of the data. */
int mhg_mlp_gradient(mhg_mlp* net, int64_t row0, int64_t m, float* loss, float* grad);
/* The same loss and gradient without any allocation, for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid it); no counterpart in the reference.  d_out: the caller's
 * device memory of mhg_mlp_param_count() + 1 floats (gradient, then loss).
 * At most 4096 rows (MHG_E_ARG beyond).  The per-block partial gradients
 * live in one fixed 6.4 MB device array per GPU, shared by every call on
 * that GPU and overwritten by each: calls that overlap on one GPU race.
 * Host-synchronous. */
int mhg_mlp_gradient_noalloc(mhg_mlp* net, int64_t row0, int64_t m, float* d_out);
/* Adam (beta 0.9, 0.999, eps 1e-8) over shuffled minibatches; the mean
 * minibatch loss of each epoch in epoch_loss [epochs]. */
int mhg_mlp_train(mhg_mlp* net, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* epoch_loss);
/* Attaches a copy of the network to a learned-limiter context: every later
 * step first computes alpha from the state at its start. */
int mhg_set_limiter_mlp(mhg_ctx* ctx, const mhg_mlp* net);
/* Coefficient given, under MHG_LIM_LEARNED with a network attached, to the
 * cells at material interfaces (whose 3 x 3 neighbourhood does not hold one
 * and the same single environment): alpha < 0 (the default, -1) keeps the
 * classical limiter there, 0 <= alpha <= 2 applies the slope alpha (g_b +
 * g_f) / 2 with the same bounds check as the network's.  No counterpart in
 * the reference. */
int mhg_set_learned_interface_alpha(mhg_ctx* ctx, double alpha);
/* Network inputs of every cell of a 2D context's current state, X [N_c][128],
 * and the channel scales (largest deviation plus 1e-6 |centre|), [N_c][5]
 * (may be NULL); the coefficients are not touched. */
int mhg_learned_features(mhg_ctx* ctx, float* X, float* scale);

/* ---- diagnostics ---------------------------------------------------- */
/* Compensated global sums: total mass, total internal energy
 * (sum m*e), total kinetic energy (sum 1/2 m_n |u|^2). */
int mhg_conservation(mhg_ctx* ctx, double out[3]);
/* Kernel timing: average duration (ms) of each fused kernel over the
 * launches recorded since the last reset, measured with HIP events on the
 * context's stream.  names/ms arrays of length cap; returns count. */
int mhg_kernel_times(mhg_ctx* ctx, int32_t enable, const char** names, double* ms,
                     int32_t* launches, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* MAHYCO_GPU_H */
