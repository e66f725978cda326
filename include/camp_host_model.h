/* Host-model interface of libcamp_mi355x.so: a grid model's tracer arrays in,
 * chemistry on the GPU, tracer arrays out.
 *
 * Replaces the gather / solve / scatter of the reference's MONARCH interface,
 * monarch_interface_t%integrate (test/monarch/camp_monarch_interface.F90:
 * 343-509), over a solver made with the C ABI of camp_solver.h.  The arrays
 * are the host model's own, in Fortran (column-major) order and single
 * precision, as that subroutine takes them:
 *
 *   NMMB style (W->E, S->N, top->bottom, ...):
 *     temperature[ni][nj][nk]                 (K)
 *     conc[ni][nj][nk][n_tracers]             (ppm or ug/m3; updated in place)
 *     water_conc[ni][nj][nk][n_water]         (kg H2O / kg air)
 *   WRF style (W->E, bottom->top, N->S):
 *     air_density[ni][nk][nj]                 (kg air / m3)
 *     pressure[ni][nk][nj]                    (Pa)
 *
 * (first index fastest).  For each grid cell (i, j, k) of the W->E range
 * i_start..i_end, the S->N range j_start..j_end and every level k (1-based),
 * with k_flip = nk - k + 1, the cell's CAMP state gets
 *   state[map_camp_id[m]]  = conc(i, j, k_flip, map_tracer_id[m])  (each m)
 *   state[water_camp_id]   = water_conc(i, j, k_flip, water_vapor_index)
 *                            * air_density(i, k, j) * 1e9
 *     (the product of the two single-precision values is rounded to single
 *     precision before the double-precision scaling, as the reference's
 *     expression evaluates)
 *   T = temperature(i, j, k_flip), P = pressure(i, k, j)
 * (camp_monarch_interface.F90:423-435, 455-471), the cells are integrated
 * over [0, time_step] (the value is passed to the solver as given, as the
 * reference passes it to camp_core%solve, :440-441, 476-477) and
 *   conc(i, j, k_flip, map_tracer_id[m]) = state[map_camp_id[m]]
 * rounded to single precision (:447-449, 480-489).  Cell z of the solver is
 * (k-1) * nI * nJ + (j - j_start) * nI + (i - i_start), nI and nJ the
 * range widths (:463-464 with i_start = j_start = 1).  All ids are 1-based.
 *
 * Multi-cell solvers (n_cells = nI * nJ * nk, checked as the reference does,
 * :451-455) keep each cell's CAMP state on the GPU between calls, so
 * species the map does not name carry over from the previous call, as the
 * reference's camp_state does in this mode.  A solver of one cell integrates
 * the cells one at a time with the whole state cleared before each gather
 * (the reference's single-cell branch, :375-449).
 */
#ifndef CAMP_HOST_MODEL_H
#define CAMP_HOST_MODEL_H

#ifdef __cplusplus
extern "C" {
#endif

/* New interface over an initialised solver (camp_solver.h).  map_tracer_id
 * and map_camp_id have n_map entries (1-based tracer index into the
 * n_tracers dimension, 1-based CAMP state id); water_camp_id is the
 * gas-phase water species' state id (1-based).  The map is the reference's
 * species map (create_map, camp_monarch_interface.F90:613-713).  Invalid
 * ids print a message and exit, like the reference's assertions. */
void *camp_hostmodel_new(void *solver_data, int ni, int nj, int nk,
                         int n_tracers, int n_water, int n_map,
                         const int *map_tracer_id, const int *map_camp_id,
                         int water_camp_id);

/* The gather, solver_run and scatter described above for the given range.
 * Returns solver_run's code (0 success, 1 failure).  A range whose cell
 * count differs from a multi-cell solver's n_cells prints a message and
 * exits (the reference's assertion 559245176). */
int camp_hostmodel_integrate(void *hm, double time_step, int i_start,
                             int i_end, int j_start, int j_end,
                             const float *temperature, float *conc,
                             const float *water_conc, int water_vapor_index,
                             const float *air_density, const float *pressure);

/* Set the GPU-resident CAMP state of every cell (n_cells x n_state_var,
 * cell-major) that a multi-cell interface carries between calls (zero when
 * the interface is made). */
void camp_hostmodel_set_state(void *hm, const double *state);

/* Test hook: the gather alone, for the range, into host arrays: state
 * (cells x n_state_var, from the resident state as integrate would see it)
 * and env (cells x 2: T, P).  Nothing is integrated or stored. */
void camp_hostmodel_gather(void *hm, int i_start, int i_end, int j_start,
                           int j_end, const float *temperature,
                           const float *conc, const float *water_conc,
                           int water_vapor_index, const float *air_density,
                           const float *pressure, double *state_out,
                           double *env_out);

void camp_hostmodel_free(void *hm);

#ifdef __cplusplus
}
#endif

#endif /* CAMP_HOST_MODEL_H */
