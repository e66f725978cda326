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
