/* Driver for the reference's own C sources -- TEST INFRASTRUCTURE ONLY.
 *
 * oracle/Makefile compiles, from /root/reference/src as they lie, the
 * reference files that build without SUNDIALS (Jacobian.c,
 * time_derivative.c, rxn_solver.c, rxns/*.c, aero_*_solver.c,
 * sub_model_solver.c, aero_reps/*.c, sub_models/*.c) and links them with
 * this driver into oracle/_ref/libcampref.so.  The driver sets up the
 * reference ModelData the way solver_new does (camp_solver.c:83-362, which
 * itself needs SUNDIALS and is not built) and exposes the reference's own
 * pattern registration, id update, rate-constant, rate-update and
 * accumulator functions to the tests, which use them to pin the oracle.
 */
#include <stdlib.h>
#include <string.h>

#include "Jacobian.h"
#include "camp_common.h"
#include "rxn_solver.h"
#include "time_derivative.h"

typedef struct {
  SolverData sd;
  Jacobian jac;
  int built;
} RefModel;

void *ref_new(int n_state, int n_cells, const int *var_type, int n_rxn,
              int n_int, int n_float, int n_env) {
  RefModel *rm = calloc(1, sizeof(RefModel));
  ModelData *md = &rm->sd.model_data;
  md->n_per_cell_state_var = n_state;
  md->n_cells = n_cells;
  md->var_type = malloc(n_state * sizeof(int));
  memcpy(md->var_type, var_type, n_state * sizeof(int));
  md->rxn_int_data = malloc((n_int + n_rxn) * sizeof(int));
  md->rxn_float_data = malloc((n_float ? n_float : 1) * sizeof(double));
  md->rxn_env_data = calloc((size_t)n_cells * n_env + 1, sizeof(double));
  md->rxn_int_indices = malloc((n_rxn + 1) * sizeof(int));
  md->rxn_float_indices = malloc((n_rxn + 1) * sizeof(int));
  md->rxn_env_idx = malloc((n_rxn + 1) * sizeof(int));
  md->n_rxn = n_rxn;
  md->n_added_rxns = 0;
  md->n_rxn_env_data = 0;
  md->rxn_int_indices[0] = 0;
  md->rxn_float_indices[0] = 0;
  md->rxn_env_idx[0] = 0;
  return rm;
}

void ref_add_rxn(void *p, int type, int n_int, int n_float, int n_env,
                 int *int_param, double *float_param) {
  RefModel *rm = p;
  rxn_add_condensed_data(type, n_int, n_float, n_env, int_param, float_param,
                         &rm->sd);
}

/* Diagonal + rxn_get_used_jac_elem + jacobian_build_matrix, exactly the
 * first part of get_jac_init (camp_solver.c:1445-1465). */
int ref_build_pattern(void *p, int *colptr, int *rowval, int max_nnz) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  int ns = md->n_per_cell_state_var;
  jacobian_initialize_empty(&rm->jac, (unsigned int)ns);
  for (unsigned int i = 0; i < (unsigned int)ns; ++i)
    jacobian_register_element(&rm->jac, i, i);
  rxn_get_used_jac_elem(md, &rm->jac);
  jacobian_build_matrix(&rm->jac);
  rm->built = 1;
  int nnz = (int)jacobian_number_of_elements(rm->jac);
  for (int i = 0; i <= ns; ++i)
    colptr[i] = (int)jacobian_column_pointer_value(rm->jac, i);
  for (int e = 0; e < nnz && e < max_nnz; ++e)
    rowval[e] = (int)jacobian_row_index(rm->jac, e);
  return nnz;
}

void ref_update_ids(void *p, int *deriv_ids) {
  RefModel *rm = p;
  rxn_update_ids(&rm->sd.model_data, deriv_ids, rm->jac);
}

void ref_get_int_data(void *p, int *dst, int n) {
  RefModel *rm = p;
  memcpy(dst, rm->sd.model_data.rxn_int_data, n * sizeof(int));
}

/* rxn_update_env_state for one cell (camp_solver.c:587-603) */
void ref_update_env(void *p, int cell, double *env2) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  md->grid_cell_id = cell;
  md->grid_cell_env = env2;
  md->grid_cell_rxn_env_data = &md->rxn_env_data[cell * md->n_rxn_env_data];
  rxn_update_env_state(md);
}

/* rxn_update_data with the {int id; double rate} payload of
 * rxn_photolysis_create_rate_update_data and friends */
void ref_update_data(void *p, int cell, int *rxn_id, int type, int uid,
                     double rate) {
  RefModel *rm = p;
  char payload[sizeof(int) + sizeof(double)];
  memcpy(payload, &uid, sizeof(int));
  memcpy(payload + sizeof(int), &rate, sizeof(double));
  rxn_update_data(cell, rxn_id, type, payload, &rm->sd);
}

void ref_get_env_data(void *p, double *dst, int n) {
  RefModel *rm = p;
  memcpy(dst, rm->sd.model_data.rxn_env_data, n * sizeof(double));
}

/* time_derivative accumulation + output on given contributions:
 * contrib[k] is added to species spec[k] (time_derivative.c:90-97), then
 * time_derivative_output (:46-88) with or without an estimate. */
void ref_time_derivative(int n_spec, int n_contrib, const int *spec,
                         const double *contrib, const double *est,
                         double *out) {
  TimeDerivative td;
  time_derivative_initialize(&td, n_spec);
  time_derivative_reset(td);
  for (int k = 0; k < n_contrib; ++k)
    time_derivative_add_value(td, spec[k], (long double)contrib[k]);
  time_derivative_output(td, out, (double *)est, 0);
  time_derivative_free(td);
}

/* ---------------------------------------------------------------------
 * Aerosol phases, representations and sub models: the ModelData arrays
 * solver_new allocates for them (camp_solver.c:200-349), then the
 * reference's own add/update/pattern functions.
 * --------------------------------------------------------------------- */
#include "aero_phase_solver.h"
#include "aero_rep_solver.h"
#include "sub_model_solver.h"

void ref_set_aero(void *p, int n_phase, int n_phase_int, int n_phase_float, int n_rep,
                  int n_rep_int, int n_rep_float, int n_rep_env, int n_sub, int n_sub_int,
                  int n_sub_float, int n_sub_env, const double *abs_tol) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  int nc = md->n_cells;
  md->aero_phase_int_data = calloc(n_phase_int + 1, sizeof(int));
  md->aero_phase_float_data = calloc(n_phase_float + 1, sizeof(double));
  md->aero_phase_int_indices = calloc(n_phase + 1, sizeof(int));
  md->aero_phase_float_indices = calloc(n_phase + 1, sizeof(int));
  md->n_aero_phase = n_phase;
  md->n_added_aero_phases = 0;
  md->aero_rep_int_data = calloc(n_rep_int + n_rep + 1, sizeof(int));
  md->aero_rep_float_data = calloc(n_rep_float + 1, sizeof(double));
  md->aero_rep_env_data = calloc((size_t)nc * n_rep_env + 1, sizeof(double));
  md->aero_rep_int_indices = calloc(n_rep + 1, sizeof(int));
  md->aero_rep_float_indices = calloc(n_rep + 1, sizeof(int));
  md->aero_rep_env_idx = calloc(n_rep + 1, sizeof(int));
  md->n_aero_rep = n_rep;
  md->n_added_aero_reps = 0;
  md->n_aero_rep_env_data = 0;
  md->sub_model_int_data = calloc(n_sub_int + n_sub + 1, sizeof(int));
  md->sub_model_float_data = calloc(n_sub_float + 1, sizeof(double));
  md->sub_model_env_data = calloc((size_t)nc * n_sub_env + 1, sizeof(double));
  md->sub_model_int_indices = calloc(n_sub + 1, sizeof(int));
  md->sub_model_float_indices = calloc(n_sub + 1, sizeof(int));
  md->sub_model_env_idx = calloc(n_sub + 1, sizeof(int));
  md->n_sub_model = n_sub;
  md->n_added_sub_models = 0;
  md->n_sub_model_env_data = 0;
  md->abs_tol = malloc(md->n_per_cell_state_var * sizeof(double));
  memcpy(md->abs_tol, abs_tol, md->n_per_cell_state_var * sizeof(double));
}

void ref_add_phase(void *p, int n_int, int n_float, int *ip, double *fp) {
  RefModel *rm = p;
  aero_phase_add_condensed_data(n_int, n_float, ip, fp, &rm->sd);
}

void ref_add_rep(void *p, int type, int n_int, int n_float, int n_env, int *ip, double *fp) {
  RefModel *rm = p;
  aero_rep_add_condensed_data(type, n_int, n_float, n_env, ip, fp, &rm->sd);
}

void ref_add_sub(void *p, int type, int n_int, int n_float, int n_env, int *ip, double *fp) {
  RefModel *rm = p;
  sub_model_add_condensed_data(type, n_int, n_float, n_env, ip, fp, &rm->sd);
}

/* the sub-model parameter pattern: dummy (0,0) + sub_model_get_used_jac_elem
 * (camp_solver.c:1509-1530), then sub_model_update_ids (:1552) */
static Jacobian ref_param_jac;
int ref_param_pattern(void *p, int *deriv_ids, int *colptr, int *rowval, int max_nnz) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  int ns = md->n_per_cell_state_var;
  jacobian_initialize_empty(&ref_param_jac, (unsigned int)ns);
  jacobian_register_element(&ref_param_jac, 0, 0);
  sub_model_get_used_jac_elem(md, &ref_param_jac);
  jacobian_build_matrix(&ref_param_jac);
  int nnz = (int)jacobian_number_of_elements(ref_param_jac);
  for (int i = 0; i <= ns; ++i) colptr[i] = (int)jacobian_column_pointer_value(ref_param_jac, i);
  for (int e = 0; e < nnz && e < max_nnz; ++e) rowval[e] = (int)jacobian_row_index(ref_param_jac, e);
  sub_model_update_ids(md, deriv_ids, ref_param_jac);
  return nnz;
}

/* per-cell environment update of every component (camp_solver.c:587-603) */
void ref_update_env_all(void *p, int cell, double *env2) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  md->grid_cell_id = cell;
  md->grid_cell_env = env2;
  md->grid_cell_rxn_env_data = &md->rxn_env_data[cell * md->n_rxn_env_data];
  md->grid_cell_aero_rep_env_data = &md->aero_rep_env_data[cell * md->n_aero_rep_env_data];
  md->grid_cell_sub_model_env_data = &md->sub_model_env_data[cell * md->n_sub_model_env_data];
  aero_rep_update_env_state(md);
  sub_model_update_env_state(md);
  rxn_update_env_state(md);
}

/* aero_rep_update_state + sub_model_calculate on one cell's state (the
 * first two steps of every f() and Jac() cell loop, camp_solver.c:913-917) */
void ref_cell_update(void *p, int cell, double *state) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  md->grid_cell_id = cell;
  md->grid_cell_state = state;
  md->grid_cell_aero_rep_env_data = &md->aero_rep_env_data[cell * md->n_aero_rep_env_data];
  md->grid_cell_sub_model_env_data = &md->sub_model_env_data[cell * md->n_sub_model_env_data];
  aero_rep_update_state(md);
  sub_model_calculate(md);
}

/* aerosol-representation getters with their partial derivatives
 * (aero_rep_solver.c:199-441): which = 0 radius, 1 number concentration,
 * 2 phase mass, 3 phase average MW. */
void ref_rep_get(void *p, int cell, double *state, int rep, int phase, int which,
                 double *value, double *partial) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  md->grid_cell_id = cell;
  md->grid_cell_state = state;
  md->grid_cell_aero_rep_env_data = &md->aero_rep_env_data[cell * md->n_aero_rep_env_data];
  switch (which) {
    case 0: aero_rep_get_effective_radius__m(md, rep, phase, value, partial); break;
    case 1: aero_rep_get_number_conc__n_m3(md, rep, phase, value, partial); break;
    case 2: aero_rep_get_aero_phase_mass__kg_m3(md, rep, phase, value, partial); break;
    case 3: aero_rep_get_aero_phase_avg_MW__kg_mol(md, rep, phase, value, partial); break;
  }
}

void ref_get_arrays(void *p, int which, void *dst, int n) {
  RefModel *rm = p;
  ModelData *md = &rm->sd.model_data;
  switch (which) {
    case 0: memcpy(dst, md->rxn_float_data, n * sizeof(double)); break;
    case 1: memcpy(dst, md->aero_rep_int_data, n * sizeof(int)); break;
    case 2: memcpy(dst, md->aero_rep_float_data, n * sizeof(double)); break;
    case 3: memcpy(dst, md->sub_model_int_data, n * sizeof(int)); break;
    case 4: memcpy(dst, md->sub_model_env_data, n * sizeof(double)); break;
  }
}
