/* CPU oracle for the CAMP chemistry hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference algorithm, used as the parity checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg.  Nothing in
 * the product library (camp_amd/) links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   pattern + ids     camp_solver.c:1417-1712, Jacobian.c:105-214,
 *                     rxn_{arrhenius,troe,photolysis,emission,first_order_loss}.c
 *                     get_used_jac_elem / update_ids
 *   rate constants    rxn_solver.c:224-313 and the per-type update_env_state
 *   rate updates      rxn_solver.c:708-763 and the per-type update_data
 *   f(y)              camp_solver.c:843-958, time_derivative.c:46-97,
 *                     per-type calc_deriv_contrib (long double accumulation)
 *   J(y)              camp_solver.c:972-1126, Jacobian.c:223-241,
 *                     per-type calc_jac_contrib
 *   solver_run        camp_solver.c:544-693, is_anything_going_on_here
 *                     camp_solver.c:1875-1898, guess_helper :1321-1408
 *   BDF integrator    CVODE (absent from the reference tree): restated
 *                     from the published algorithm, Hindmarsh et al.,
 *                     ACM TOMS 31(3):363-396, 2005, and the CVODE 3.x
 *                     user guide (variable-order Nordsieck BDF, modified
 *                     Newton, WRMS error control).  PARITY UNPINNED at the
 *                     trajectory level: pinned only by analytic solutions,
 *                     high-accuracy reference solutions and invariants.
 *   linear solver     KLU (absent): a sparse LU with a static
 *                     minimum-degree ordering, cross-checked in tests
 *                     against dense partial-pivoting LU.  PARITY UNPINNED
 *                     against KLU itself (rounding-level differences).
 */
#ifndef CAMP_ORACLE_H
#define CAMP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

void *oracle_new(int n_state_var, int n_cells, const int *var_type, int n_rxn,
                 int n_rxn_int_param, int n_rxn_float_param,
                 int n_rxn_env_param);
void oracle_add_rxn(void *om, int rxn_type, int n_int, int n_float, int n_env,
                    const int *int_param, const double *float_param);
void oracle_initialize(void *om, const double *abs_tol, double rel_tol,
                       int max_steps, int max_conv_fails);
void oracle_free(void *om);

/* rxn_update_data for a {int id; double rate} payload. Returns 1 if found. */
int oracle_update_data(void *om, int cell_id, int *rxn_id, int rxn_type,
                       int unique_id, double rate);

/* Pattern / id getters (copies). */
int oracle_n_dep(void *om);
int oracle_nnz(void *om);
int oracle_nnz_rxn(void *om);
void oracle_get_pattern(void *om, int *rxn_colptr, int *rxn_rowval,
                        int *colptr, int *rowval, int *deriv_ids);
void oracle_get_jac_map(void *om, int *solver_id, int *rxn_id, int *param_id);
int oracle_n_mapped(void *om);
void oracle_get_rxn_int_data(void *om, int *dst, int n);
int oracle_rxn_int_size(void *om);
int oracle_n_env(void *om);
void oracle_get_rxn_env(void *om, double *dst);

/* Rate constants for all cells from env (what solver_run does first). */
void oracle_update_env_state(void *om, const double *env);

/* f and J of one cell.  y_dep has n_dep entries.  use_deriv_est selects the
 * Jacobian-extrapolated blend of time_derivative_output.  Returns 0 or 1
 * (recoverable failure on y < -SMALL). oracle_eval_jac also updates the
 * cell's J history (J_solver/J_state/J_deriv) like Jac() does. */
int oracle_eval_f(void *om, int cell, const double *y_dep, double *deriv,
                  double time_step, int use_deriv_est);
int oracle_eval_jac(void *om, int cell, const double *y_dep, double *jac,
                    double time_step);
void oracle_reset_jac_history(void *om);
/* J_deriv + J_solver (y - J_state) of one cell from its J history (the
 * estimate the blend of time_derivative.c:58-64 uses). */
void oracle_eval_est(void *om, int cell, const double *state_cell, double *est);

/* solver_run: mode 0 = reference block semantics (one BDF over all
 * cells), mode 1 = per-cell (each cell its own BDF, n_cells = 1
 * semantics).  state is n_cells*n_state_var AoS, env 2*n_cells. */
int oracle_solver_run(void *om, double *state, const double *env,
                      double t_initial, double t_final, int mode);
/* Same as mode 1 but only for cells [c0, c1); returns # failed cells. */
int oracle_solver_run_cells(void *om, double *state, const double *env,
                            double t_initial, double t_final, int c0, int c1);

/* Statistics of the last run: [nst, nfe, nsetups, netf, nni, ncfn, nje,
 * flag] summed over problems (mode 1) or of the block (mode 0), plus last
 * and next step size of the last problem solved. */
void oracle_get_stats(void *om, long *counters8, double *h2);
void oracle_set_guess_helper(void *om, int on);
/* 1 (default): reject a converged Newton iterate with y < -SMALL as a
 * recoverable f() failure; 0: accept it (plain CVODE behaviour). */
void oracle_set_neg_mode(void *om, int mode);

/* Linear-algebra checks used by tests: factor M = I - gamma*J (solver
 * pattern, one cell) with the oracle's sparse LU and with dense partial
 * pivoting, then solve M x = b.  Returns 0 on success. */
int oracle_lu_solve_sparse(void *om, const double *jac, double gamma,
                           const double *b, double *x);
int oracle_lu_solve_dense(void *om, const double *jac, double gamma,
                          const double *b, double *x);
int oracle_lu_nnz(void *om);

/* Known-answer hooks: the pattern builder on explicit (dep, ind)
 * registrations (returns nnz), the long-double Jacobian accumulators, and
 * the time-derivative accumulators + output. */
int oracle_kat_pattern(int n, int n_pairs, const int *dep, const int *ind,
                       int *colptr, int *rowval);
void oracle_kat_jac_accum(int nnz, int n_adds, const int *elem, const int *flag,
                          const double *val, double *out);
void oracle_kat_time_derivative(int n_spec, int n_contrib, const int *spec,
                                const double *contrib, const double *est,
                                double *out);

#ifdef __cplusplus
}
#endif
#endif
