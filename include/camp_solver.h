/* C ABI of the MI355X-native CAMP chemistry solver (libcamp_mi355x.so).
 *
 * The first block of entry points is the drop-in boundary: the symbols the
 * reference's Fortran layer binds with bind(c) (src/camp_solver_data.F90:
 * 43-363), with the reference signatures, argument meaning and error
 * convention (setup errors print and exit(EXIT_FAILURE); solver_run
 * returns 0 = CAMP_SOLVER_SUCCESS or 1 = CAMP_SOLVER_FAIL).  The state array
 * is n_state_var * n_cells doubles, cell-major; env is 2 * n_cells doubles
 * [temperature K, pressure Pa] per cell; species ids inside the packed int
 * parameters are 1-based.
 *
 * The second block has no counterpart in the reference: mode selection,
 * device-resident runs, bulk rate updates and the test/inspection hooks.
 */
#ifndef CAMP_MI355X_SOLVER_H
#define CAMP_MI355X_SOLVER_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define CAMP_SOLVER_SUCCESS 0 /* camp_solver.c:49 */
#define CAMP_SOLVER_FAIL 1    /* camp_solver.c:50 */

/* ---- drop-in boundary ---------------------------------------------------- */

/* replaces solver_new, src/camp_solver.c:83-362 (camp_solver.h:16-24) */
void *solver_new(int n_state_var, int n_cells, int *var_type, int n_rxn,
                 int n_rxn_int_param, int n_rxn_float_param,
                 int n_rxn_env_param, int n_aero_phase,
                 int n_aero_phase_int_param, int n_aero_phase_float_param,
                 int n_aero_rep, int n_aero_rep_int_param,
                 int n_aero_rep_float_param, int n_aero_rep_env_param,
                 int n_sub_model, int n_sub_model_int_param,
                 int n_sub_model_float_param, int n_sub_model_env_param);

/* replaces solver_initialize, src/camp_solver.c:374-494 */
void solver_initialize(void *solver_data, double *abs_tol, double rel_tol,
                       int max_steps, int max_conv_fails);

/* replaces solver_run, src/camp_solver.c:544-693 */
int solver_run(void *solver_data, double *state, double *env, double t_initial,
               double t_final);

/* replaces solver_get_statistics, src/camp_solver.c:725-791 */
void solver_get_statistics(void *solver_data, int *solver_flag, int *num_steps,
                           int *RHS_evals, int *LS_setups,
                           int *error_test_fails, int *NLS_iters,
                           int *NLS_convergence_fails, int *DLS_Jac_evals,
                           int *DLS_RHS_evals, double *last_time_step__s,
                           double *next_time_step__s, int *Jac_eval_fails,
                           int *RHS_evals_total, int *Jac_evals_total,
                           double *RHS_time__s, double *Jac_time__s,
                           double *max_loss_precision);

/* replaces solver_reset_timers, src/camp_solver.c:1764-1775 */
void solver_reset_timers(void *solver_data);

/* replaces solver_free, src/camp_solver.c:1832-1873 */
void solver_free(void *solver_data);

/* replaces rxn_add_condensed_data, src/rxn_solver.c:658-693 */
void rxn_add_condensed_data(int rxn_type, int n_int_param, int n_float_param,
                            int n_env_param, int *int_param,
                            double *float_param, void *solver_data);

/* replaces rxn_update_data, src/rxn_solver.c:708-763; update_data is the
 * {int id; double rate} payload of the *_create_rate_update_data helpers */
void rxn_update_data(int cell_id, int *rxn_id, int update_rxn_type,
                     void *update_data, void *solver_data);

/* replaces rxn_print_data, src/rxn_solver.c:769-845 */
void rxn_print_data(void *solver_data);

/* replaces rxn_free_update_data, src/rxn_solver.c:851 */
void rxn_free_update_data(void *update_data);

/* replace the bind(c) helpers of rxn_photolysis.F90:135-155,
 * rxn_emission.F90:118-138 and rxn_first_order_loss.F90:121-141, whose C
 * bodies are rxn_photolysis.c:262-283, rxn_emission.c:192-213 and
 * rxn_first_order_loss.c:208-229 */
void *rxn_photolysis_create_rate_update_data(void);
void rxn_photolysis_set_rate_update_data(void *update_data, int photo_id,
                                         double base_rate);
void *rxn_emission_create_rate_update_data(void);
void rxn_emission_set_rate_update_data(void *update_data, int rxn_id,
                                       double base_rate);
void *rxn_first_order_loss_create_rate_update_data(void);
void rxn_first_order_loss_set_rate_update_data(void *update_data, int rxn_id,
                                               double base_rate);

/* replace aero_phase_add_condensed_data (aero_phase_solver.c:175),
 * aero_rep_add_condensed_data (aero_rep_solver.c:443) and
 * sub_model_add_condensed_data (sub_model_solver.c:413), bound at
 * camp_solver_data.F90:201-354: packed data of aerosol phases, aerosol
 * representations (MODAL_BINNED_MASS) and sub models (ZSR_AEROSOL_WATER),
 * in the reference's layouts.  Unsupported types print an error and exit
 * at solver_initialize, the reference's setup-error convention. */
void aero_phase_add_condensed_data(int n_int_param, int n_float_param,
                                   int *int_param, double *float_param,
                                   void *solver_data);
void aero_rep_add_condensed_data(int aero_rep_type, int n_int_param,
                                 int n_float_param, int n_env_param,
                                 int *int_param, double *float_param,
                                 void *solver_data);
void sub_model_add_condensed_data(int sub_model_type, int n_int_param,
                                  int n_float_param, int n_env_param,
                                  int *int_param, double *float_param,
                                  void *solver_data);

/* replaces sub_model_update_data, src/sub_model_solver.c:466-504 (bound at
 * camp_solver_data.F90:267-280): scans the sub models from *sub_model_id,
 * advancing it; no supported sub-model type takes update data */
void sub_model_update_data(int cell_id, int *sub_model_id,
                           int update_sub_model_type, void *update_data,
                           void *solver_data);

/* replaces sub_model_print_data, src/sub_model_solver.c:509-541
 * (camp_solver_data.F90:283) */
void sub_model_print_data(void *solver_data);

/* replaces aero_phase_print_data, src/aero_phase_solver.c:199-218
 * (camp_solver_data.F90:306) */
void aero_phase_print_data(void *solver_data);

/* replaces aero_rep_update_data, src/aero_rep_solver.c:492-541 with
 * aero_rep_modal_binned_mass.c:594-642 (camp_solver_data.F90:334-347):
 * GMD/GSD updates address MODAL sections; a matching update of this
 * solver's BINNED-only representations is the reference's error exit */
void aero_rep_update_data(int cell_id, int *aero_rep_id,
                          int update_aero_rep_type, void *update_data,
                          void *solver_data);

/* replaces aero_rep_print_data, src/aero_rep_solver.c:546-582
 * (camp_solver_data.F90:350) */
void aero_rep_print_data(void *solver_data);

/* replace the bind(c) helpers of aero_rep_modal_binned_mass.F90:258-304,
 * whose C bodies are aero_rep_modal_binned_mass.c:677-738 and
 * aero_rep_solver.c:589: payload {int aero_rep_id, int update_type, int
 * section_id, double value} */
void *aero_rep_modal_binned_mass_create_gmd_update_data(void);
void aero_rep_modal_binned_mass_set_gmd_update_data(void *update_data,
                                                    int aero_rep_id,
                                                    int section_id, double gmd);
void *aero_rep_modal_binned_mass_create_gsd_update_data(void);
void aero_rep_modal_binned_mass_set_gsd_update_data(void *update_data,
                                                    int aero_rep_id,
                                                    int section_id, double gsd);
void aero_rep_free_update_data(void *update_data);

/* replace the payload helpers of the reaction and representation types this
 * solver does not integrate, so that every Fortran module links
 * (rxn_wet_deposition.F90:124-144 / rxn_wet_deposition.c:218-239,
 * rxn_condensed_phase_photolysis.F90:140-160 / .c:464-485,
 * aero_rep_single_particle.F90:169-189 / .c:405-433); a
 * mechanism with those types stops at solver_initialize (setup error) */
void *rxn_wet_deposition_create_rate_update_data(void);
void rxn_wet_deposition_set_rate_update_data(void *update_data, int rxn_id,
                                             double base_rate);
void *rxn_condensed_phase_photolysis_create_rate_update_data(void);
void rxn_condensed_phase_photolysis_set_rate_update_data(void *update_data,
                                                         int photo_id,
                                                         double base_rate);
void *aero_rep_single_particle_create_number_update_data(void);
void aero_rep_single_particle_set_number_update_data__n_m3(void *update_data,
                                                           int aero_rep_id,
                                                           int particle_id,
                                                           double number_conc);

/* replace solver_set_debug_out and solver_set_eval_jac,
 * src/camp_solver.c:502-533 (CAMP_DEBUG builds; camp_solver_data.F90:
 * 100-118): debug output is accepted and ignored; eval_jac turns on a
 * finite-difference check (camp_solver_check_jac's rule) of every cell's
 * last Jacobian after each solver_run, and solver_get_statistics reports the
 * cells that fail it as Jac_eval_fails (the reference checks inside every
 * Jac() call, camp_solver.c:1116-1122) */
int solver_set_debug_out(void *solver_data, bool do_output);
int solver_set_eval_jac(void *solver_data, bool eval_jac);

/* ---- extensions (no counterpart in the reference) ------------------------ */

#define CAMP_MODE_PER_CELL 0 /* each cell its own BDF control (default) */
#define CAMP_MODE_BLOCK 1    /* one BDF over all cells, launch per f/J/LU/solve */

/* Select the integration mode; returns the previous mode. */
int camp_solver_set_mode(void *solver_data, int mode);

/* Block mode: replay each Newton iteration's device work as HIP graphs (1,
 * default) or launch it kernel by kernel (0).  Results are identical.
 * Returns the previous setting. */
int camp_solver_set_graphs(void *solver_data, int on);

/* Cells per workgroup of the persistent kernel (1, 2, 4, 8, 16, 32; one
 * wavefront per cell; a value whose workgroup does not fit a compute unit
 * is ignored); 0 selects the built-in default.  Returns the value in
 * effect. */
int camp_solver_set_cells_per_block(void *solver_data, int cells_per_block);

/* Cell placement of the per-cell kernel (gas-phase programs; multi-phase
 * programs always use 1): 0 = each cell's Nordsieck history, work vectors
 * and J history in LDS (fewer cells per compute unit), 1 = those in global
 * memory (more cells per compute unit, two wavefronts per SIMD; the
 * default), 2 = those and the LU factors in global memory.  Returns the
 * placement in effect.  Results do not depend on it. */
int camp_solver_set_placement(void *solver_data, int place);

/* LDS bytes one cell of the per-cell kernel uses with the given placement
 * (-1 for an unknown placement or an uninitialised solver; makes the
 * solver's device buffers). */
long camp_solver_cell_lds_bytes(void *solver_data, int place);

/* Autotuner for the per-cell kernel: solves a sample of n_sample cells
 * (every (n_cells / n_sample)-th; 0 = up to 8192) over [t_initial, t_final]
 * from the device-resident state/env once per candidate -- placement {0, 1, 2}
 * times cells per workgroup {1, 2, 4, 8, 16, 32} -- that fits a compute
 * unit, on copies (the solver's own cells, rate data and J history are
 * untouched), and keeps the fastest.  ms_out[18] receives each candidate's
 * kernel time, placement-major (-1: does not fit, or not available for the
 * program).  Returns the chosen cells per workgroup plus 100 times the chosen
 * placement. */
int camp_solver_autotune(void *solver_data, const double *d_state,
                         const double *d_env, double t_initial, double t_final,
                         int n_sample, double *ms_out);

/* Measurement hook: the per-cell kernel's cycles for f(y), J(y), the LU
 * factorisation and one triangular solve, each timed reps times at the
 * state of every cell of a sample of n_sample cells (every
 * (n_cells/n_sample)-th, copies; 0 = up to 4096) with the solver's placement
 * and cells per workgroup, as many cells resident as in a solve.  out4
 * receives the mean cycles over the sample.  Returns 0, or 1 for gas-phase
 * programs at placement 2. */
int camp_solver_phase_cycles(void *solver_data, const double *d_state,
                             const double *d_env, int n_sample, int reps,
                             double *out4);

/* Enable (1, default) or disable (0) the negative-prediction guess helper
 * (camp_solver.c:1321-1408). */
void camp_solver_set_guess_helper(void *solver_data, int on);

/* Queue order of the per-cell kernel: order[n] is a permutation of the
 * cells, which the launch's wavefronts then take first to last.  n = 0 or
 * order = NULL restores the default: index order on the first call, then
 * the cells by the previous call's BDF step counts, most first (so that a
 * launch ends on cheap cells).  Every cell's result is independent of the
 * order.  Returns 0, or 1 if order is not a permutation of n_cells cells. */
int camp_solver_set_cell_order(void *solver_data, const int *order, int n);

/* Independence check of the per-cell kernel (on = 1): before it stages each
 * cell, a wavefront fills all of its working memory -- the cell's LDS region
 * and, for multi-phase mechanisms, its global scratch -- with NaNs, so that a
 * value read before it is written in the same cell would show in the result.
 * Results with it on equal those with it off. */
void camp_solver_set_poison(void *solver_data, int on);

/* solver_run on device-resident arrays (HIP device pointers, same layout
 * as solver_run's host arrays); ordered on the solver's stream and
 * synchronised before returning. */
int camp_solver_run_device(void *solver_data, double *d_state, double *d_env,
                           double t_initial, double t_final);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  Same as camp_solver_run_device, but the per-cell kernel's
 * queue counter and wavefront scratch come from a fixed 64 MiB device array
 * in the library (one per GPU, kept until the process ends) and nothing is
 * allocated, on the device or on the host.  Per-cell mode only, without
 * phase timing or the Jacobian check; the solver's own buffers must exist
 * (any earlier call that uses the device makes them, e.g. solver_run,
 * camp_solver_run_device or camp_solver_get_rxn_env).  Launches as many
 * wavefronts as the array holds (fewer than camp_solver_run_device for
 * multi-phase mechanisms, so slower; results are the same).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU race.
 * Returns 0 / 1 like solver_run, or CAMP_NOALLOC_NOT_READY (solver not
 * initialised or its device buffers missing), CAMP_NOALLOC_UNSUPPORTED
 * (block mode, timing or Jacobian check on), CAMP_NOALLOC_TOO_LARGE (one
 * wavefront's scratch exceeds the array), with a message on stderr. */
#define CAMP_NOALLOC_NOT_READY -1
#define CAMP_NOALLOC_UNSUPPORTED -2
#define CAMP_NOALLOC_TOO_LARGE -3
int camp_solver_run_device_noalloc(void *solver_data, double *d_state, double *d_env,
                                   double t_initial, double t_final);

/* Set the base rate of updatable reaction `unique_id` (photolysis,
 * emission, first-order loss) for every cell from rates[n_cells]; the bulk
 * equivalent of one rxn_update_data call per cell. */
int camp_solver_set_base_rates(void *solver_data, int rxn_type, int unique_id,
                               const double *rates);

/* Number of solver variables, solver-Jacobian nonzeros, LU nonzeros and
 * reaction-Jacobian nonzeros per cell (after solver_initialize). */
int camp_solver_dims(void *solver_data, int *n_dep, int *nnz, int *lu_nnz,
                     int *nnz_rxn);

/* Copies of the CAMP sparsity pattern and ids built by solver_initialize:
 * reaction Jacobian (CSC over state variables), per-cell solver Jacobian
 * (CSC over solver variables), deriv_ids[n_state], and the packed int
 * parameters after the id update (n_rxn_int_param + n_rxn entries). */
void camp_solver_get_pattern(void *solver_data, int *rxn_colptr,
                             int *rxn_rowval, int *colptr, int *rowval,
                             int *deriv_ids, int *rxn_int_data);

/* Test hooks: f(y) and J(y) of all cells on the GPU for the given host state
 * (n_cells x n_state_var) and env; rate constants are computed first.
 * deriv is n_cells x n_dep, jac n_cells x nnz (CSC element order).
 * use_deriv_est selects the Jacobian-extrapolated blend against the stored
 * J history; camp_eval_jac also stores the history, like Jac().  Return 0,
 * or 1 when a cell has y < -1e-30. */
int camp_eval_deriv(void *solver_data, const double *state, const double *env,
                    double time_step, int use_deriv_est, double *deriv);
int camp_eval_jac(void *solver_data, const double *state, const double *env,
                  double time_step, double *jac);
/* Finite-difference Jacobian check on the GPU, check_Jac's role
 * (camp_solver.c:1146-1257) with its pass rule: of the difference
 * quotients of f_i at the steps h = y_j * s, s = 1e-12 .. 1 by decades
 * (forward differences at h and h/2, Richardson-extrapolated, with a
 * truncation plus rounding error estimate err), the one with the smallest
 * err must agree with J_ij to within max(1.2 err, 1e-9) absolute or 1e-4
 * relative; columns of species below 1e-30 are skipped.  jac (n_cells x nnz, CSC element order) is checked as
 * given, or, when NULL, J(y) is evaluated first as camp_eval_jac does.
 * fails[n_cells] (may be NULL) receives each cell's failing elements.
 * Returns the total, or -1 when f() fails at a cell's state. */
int camp_solver_check_jac(void *solver_data, const double *state,
                          const double *env, double time_step,
                          const double *jac, int *fails);

/* Kernel family of camp_eval_deriv / camp_eval_jac for gas-phase
 * mechanisms: 0 (default) the one-wavefront-per-cell f/J code the per-cell
 * kernel runs, 1 the lane-group kernels of the block path (8 lanes per
 * cell).  Returns the previous setting. */
int camp_solver_set_eval_kernels(void *solver_data, int group);

/* Batched LU solve test hook: factor I - gamma*J per cell (jac n_cells x
 * nnz) and solve for b (n_cells x n_dep) in place. Returns 0 or 1. */
int camp_eval_lu_solve(void *solver_data, const double *jac, double gamma,
                       double *b);

/* Per-cell statistics of the last per-cell run: n_cells x 10 doubles
 * [num_steps, RHS_evals, LS_setups, error_test_fails, NLS_iters,
 *  NLS_convergence_fails, Jac_evals, flag, last_h, next_h]. */
void camp_solver_get_cell_stats(void *solver_data, double *stats);

/* Duration (ms) of the last persistent-kernel launch, from HIP events on
 * the launch stream. */
double camp_solver_last_kernel_ms(void *solver_data);

/* Per-cell phase timing in per-cell mode (the device-side counterpart of the
 * RHS/Jac timers that solver_get_statistics reports, camp_solver.c:725-791):
 * when on, each solver_run records per cell the shader-clock cycles spent in
 * f(y), J(y), LU factorisation, triangular solves and the whole integration,
 * and solver_get_statistics reports the f and J sums as RHS_time__s and
 * Jac_time__s.  get_cell_profile copies [n_cells][7] cycles: f, J, LU,
 * solves, whole integration, then f's two parts (state update + rates,
 * species gather + estimate blend); returns 1 if no profiled run has
 * happened. */
void camp_solver_set_profiling(void *solver_data, int on);
int camp_solver_get_cell_profile(void *solver_data, double *out);

/* Algorithmic work per cell and evaluation, for roofline accounting:
 * out8 = [flops of f(y) (with the Jacobian-estimate blend), of J(y), of one
 * LU factorisation, of one triangular solve, then the HBM bytes the same
 * four operations must move when run as separate kernels]. */
void camp_solver_work_model(void *solver_data, double *out8);

/* Measurement hook: average duration (ms, HIP events on the solver's
 * stream) of the plain-path f(y) and J(y) kernels over all cells for the
 * device-resident state/env, each launched `reps` times. */
int camp_solver_time_fj(void *solver_data, const double *d_state,
                        const double *d_env, int reps, double *ms_deriv,
                        double *ms_jac);

/* Autotuner of the block path's f(y) and J(y) kernels (gas-phase
 * mechanisms): times the lane-group kernels at 8, 16 and 32 lanes per cell
 * on the device-resident state/env (reps launches each) and keeps the
 * fastest width for f and for J.  ms_out[6] receives [f at 8, 16, 32, J at
 * 8, 16, 32] (ms per launch).  Returns the chosen f width plus 100 times
 * the chosen J width. */
int camp_solver_tune_fj(void *solver_data, const double *d_state,
                        const double *d_env, int reps, double *ms_out);
/* Set the lanes per cell (8, 16 or 32; other values leave it) of the block
 * path's f and J kernels; returns f lanes plus 100 times J lanes. */
int camp_solver_set_fj_lanes(void *solver_data, int f_lanes, int j_lanes);

/* Copy of the device-side per-cell reaction environment data (n_cells x
 * n_rxn_env_param: rate constants and base rates) as last computed. */
void camp_solver_get_rxn_env(void *solver_data, double *rxn_env);

/* Test hook: the host build of the correctly rounded elementary functions
 * that the device code uses in place of the C library's in the restated
 * rate-constant, SIMPOL, aqueous-equilibrium and ZSR formulas.
 * fn 0 exp(x), 1 log(x), 2 log10(x), 3 pow(x, y); n values. */
void camp_cr_math(int fn, int n, const double *x, const double *y, double *out);

#ifdef __cplusplus
}
#endif
#endif
