/*
 * mahyco_oracle.h -- CPU FP64 oracle of the MaHyCo time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  It is never part of the product path.
 *
 * It exports the same interface as include/mahyco_gpu.h with the prefix mho_
 * instead of mhg_, using the same option/state/mesh structures, so a harness
 * can swap the oracle and the GPU library.  In addition it exports a few
 * single-function probes (mho_probe_*) used by the closed-form unit tests.
 */
#ifndef MAHYCO_ORACLE_H
#define MAHYCO_ORACLE_H

#include "../include/mahyco_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mho_ctx mho_ctx;

int mho_abi_version(void);
void mho_default_options(mhg_options* opt);
int mho_create(const mhg_mesh_desc* mesh, const mhg_options* opt, mho_ctx** out);
int mho_destroy(mho_ctx* ctx);
const char* mho_last_error(const mho_ctx* ctx);
int mho_sizes(const mho_ctx* ctx, int64_t* n_cells, int64_t* n_nodes);
int mho_init_sod(mho_ctx* ctx, int32_t cas_test);
int mho_init_triple_point(mho_ctx* ctx, const double rho[3], const double p[3]);
int mho_init_cells(mho_ctx* ctx, const int32_t* env_of_cell, const double* density,
                   const double* pressure, const double* velocity);
int mho_upload_state(mho_ctx* ctx, const mhg_state* st);
/* Arcane's time-loop manager: time += dt, iteration += 1 before the
 * compute-loop entry points of an iteration. */
int mho_begin_iteration(mho_ctx* ctx);
int mho_download_state(mho_ctx* ctx, mhg_state* st);
int mho_save_values_at_n(mho_ctx* ctx);
int mho_update_velocity(mho_ctx* ctx);
int mho_apply_boundary_condition(mho_ctx* ctx);
int mho_update_position(mho_ctx* ctx);
int mho_compute_geometric_values(mho_ctx* ctx);
int mho_update_density(mho_ctx* ctx);
int mho_compute_artificial_viscosity(mho_ctx* ctx);
int mho_update_energy_and_pressure(mho_ctx* ctx);
int mho_remap(mho_ctx* ctx);
int mho_compute_delta_t(mho_ctx* ctx, double* new_dt, int32_t* stop);
int mho_remap_prepare(mho_ctx* ctx);
int mho_remap_appli(mho_ctx* ctx, int32_t dimension, int32_t with_dual_projection, int32_t nb_vars_to_project,
                    int32_t nb_env);
int mho_remap_variables(mho_ctx* ctx, int32_t dimension, int32_t with_dual_projection, int32_t nb_vars_to_project,
                        int32_t nb_env);
int mho_remap_resize_variables(mho_ctx* ctx, int32_t nb_vars_to_project, int32_t nb_env);
int mho_remap_synchronize_uremap(mho_ctx* ctx);
int mho_eos_init(mho_ctx* ctx, int32_t env);
int mho_eos_apply(mho_ctx* ctx, int32_t env);
int mho_eos_apply_one_cell(mho_ctx* ctx, int32_t env, int64_t cell);
int mho_eos_get_adiabatic_cst(const mho_ctx* ctx, int32_t env, double* value);
int mho_eos_get_tension_limit_cst(const mho_ctx* ctx, int32_t env, double* value);
int mho_compute_pression_moyenne(mho_ctx* ctx);
int mho_remap_get_ordre_projection(const mho_ctx* ctx);
int mho_remap_has_projection_pente_borne(const mho_ctx* ctx);
int mho_remap_has_conservation_energie_totale(const mho_ctx* ctx);
int mho_remap_is_euler(const mho_ctx* ctx);
int mho_step(mho_ctx* ctx, int32_t* stop);
int mho_run(mho_ctx* ctx, int32_t max_steps, int32_t* steps_done);
int mho_conservation(mho_ctx* ctx, double out[3]);
/* Number of OpenMP threads the oracle uses (0 = leave the runtime default). */
int mho_set_threads(int32_t n);

/* ---- probes for closed-form tests -------------------------------------- */
/* Corner normals of one cell (MahycoModule.cc:1093-1112 in 2D,
 * computeCQs MahycoModule.cc:2195-2258 in 3D).  coord: [nb_node][3],
 * out_cqs: [nb_node][3]; returns the volume (1/dim) sum x_i . C_i. */
double mho_probe_cqs(int32_t dim, const double* coord, double* out_cqs);
/* Limiters of RemapADIService.h:260-286 (id 0 minmod, 1 superbee, 2 vanLeer,
 * other: order 1). */
double mho_probe_limiter(int32_t id, double r);
/* Generalised limiters (fluxLimiterG, UtilesRemap.cc:127-173; id
 * MHG_LIM_MINMODG .. MHG_LIM_ULTRABEEG): slope from the front / back
 * gradients, front / back values and the widths h0 (cell), hp (front
 * neighbour), hm (back neighbour). */
double mho_probe_limiter_g(int32_t id, double gp, double gm, double yp, double ym, double h0, double hp,
                           double hm);
/* Order-3 face value (ComputeFluxOrdre3, Remap-ordre-3.cc:69-131): y6 and h6
 * the values and widths of the cells bbb, bb, b | f, ff, fff around the face,
 * vdt the face velocity times dt. */
double mho_probe_flux3(const double* y6, const double* h6, double vdt);
/* Pente-borne (UtilesRemap.cc:73-603): INTY, and one variable's integrated
 * face flux (computeY0 with limiter lim, computexgxd / computeygyd with the
 * cell's limited gradient grady, the type-0 (back face) or type-1 (front
 * face) integral over ppv, formula 16); thr the remap threshold. */
double mho_probe_pb_inty(double X, double x0, double y0, double x1, double y1, double thr);
double mho_probe_pb_flux(int32_t type, int32_t lim, double ppv, double h0, double hp, double hm, double y0,
                         double yp, double ym, double grady, double thr);
/* Cell widths along idir (RemapADIService.cc:223-255) of the present
 * coordinates, out [N_c]. */
int mho_probe_cell_widths(mho_ctx* ctx, int32_t idir, double* out);
/* Learned-limiter slope of one cell and variable (MHG_LIM_LEARNED). */
double mho_probe_learned_grad(double alpha, double phim, double phi, double phip, int32_t back_in,
                              int32_t front_in, double gb, double gf, double db, double df);
/* Consistent environment slopes of the learned mode in a mixed cell: phi and
 * g [3 ne] (volumes, masses, energies), gcl the classical slopes, used[k]
 * whether the learned volume slope of environment k was kept; g is updated
 * in place (remap threshold 1e-10). */
void mho_probe_consistent_slopes(int32_t ne, const double* phi, double* g, const double* gcl, const int32_t* used);
/* Coefficients of the learned limiter, [dim][N_c] (kept until replaced). */
int mho_set_learned_alpha(mho_ctx* ctx, const double* alpha);
/* compute_eint (MahycoModule.cc:1630-1654). */
double mho_probe_eint(int32_t pseudo_centree, double gamma, double q_n, double q,
                      double rho_n, double rho, double p, double e_n);
/* Pseudo-viscosity formula (MahycoModule.cc:517-528). */
double mho_probe_viscosity(double gamma, double tau, double L, double c, double divu);
/* Perfect-gas EOS (PerfectGasEOSService.cc:38-45): out[0]=p, out[1]=c, out[2]=dpde. */
void mho_probe_eos_pg(double gamma, double rho, double e, double* out3);
/* EOS of one cell, perfect gas (eos 0) or stiffened gas (eos 1,
 * StiffenedGasEOSService.cc:27-44): out[0]=p, out[1]=c, out[2]=dpde. */
void mho_probe_eos(int32_t eos, double gamma, double limit_tension, double rho, double e, double* out3);
/* initEOS of one cell (PerfectGasEOSService.cc:20-31, StiffenedGasEOSService.cc:10-22):
 * out[0]=e, out[1]=c from rho and p. */
void mho_probe_eos_init(int32_t eos, double gamma, double limit_tension, double rho, double p, double* out2);

#ifdef __cplusplus
}
#endif
#endif
