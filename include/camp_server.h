/* Request server over the per-cell CAMP solver (libcamp_mi355x.so).
 *
 * No counterpart in the reference: CAMP integrates one host model's cells
 * per solver_run call (camp_solver.c:544-693).  The server admits chemistry
 * solve requests of any cell count for any initialised mechanism while
 * others run: a request is admitted when its cells fit in the mechanism's
 * HBM cell arena (sized from free device memory), its state stays in HBM
 * across its timesteps, and admitted requests form one first-come queue of
 * cells.  Whenever a stream is free the scheduler launches the persistent
 * kernel over the next cells of the oldest request's mechanism (from that
 * request and later ones of the same mechanism, up to twice the cells the
 * GPU holds at once; a large request spans several launches; one mechanism
 * never holds every stream while another has cells waiting), each cell
 * running its request's timesteps back to back (continuous batching:
 * requests that arrive while others run join the queue and start as soon
 * as a stream frees up; a request completes when its last cells do).
 * Cells never share integrator control, so a request's results are
 * bit-identical alone and in any mix.
 *
 * Threading: one scheduler thread per server owns every HIP call; submit,
 * wait and the counters may be called from any thread.  Host buffers passed
 * to submit must stay valid until the request completes.
 */
#ifndef CAMP_SERVER_H
#define CAMP_SERVER_H

#ifdef __cplusplus
extern "C" {
#endif

/* n_streams concurrent launches (1..8; the GPU runs at most as many
 * kernels at once as the process has hardware queues, 4 by default);
 * max_batch_cells caps one launch below the default chunk. */
void *camp_server_new(int n_streams, long max_batch_cells);

/* Registers an initialised per-cell solver (solver_new + solver_initialize)
 * as a mechanism; its program and controls (rtol, atol, max_steps) are
 * used, its own cell buffers are not.  max_cells = 0 sizes the arena from
 * half the free HBM at the time of the call.  Up to 16 mechanisms per
 * server.  Returns the mechanism id. */
int camp_server_add_mechanism(void *server, void *solver_data, long max_cells);

/* Arena capacity (cells) of a mechanism. */
long camp_server_capacity(void *server, int mech);

/* Base rates a request supplies per cell, in this order: the photolysis,
 * emission and first-order-loss reactions in the order they were added,
 * each identified by (rxn_type, unique id) as rxn_update_data expects.
 * Returns the count; types/uids may be NULL to query it. */
int camp_solver_rate_layout(void *solver_data, int *types, int *uids);

/* Queues a request: n_cells cells, state [n_cells][n_state], env
 * [n_cells][2], base_rates [n_cells][n_rates] (NULL if n_rates is 0),
 * integrated n_steps times over dt seconds each.  out_state receives the
 * final state; out_stats (may be NULL) the last step's per-cell statistics
 * as camp_solver_get_cell_stats lays them out.  Returns the request id
 * (>= 0), or -1 if the request can never fit the arena. */
long camp_server_submit(void *server, int mech, int n_cells, const double *state,
                        const double *env, const double *base_rates, double dt, int n_steps,
                        double *out_state, double *out_stats);

/* Blocks until the request completes; latency_s (may be NULL) receives
 * completion minus submission time.  Returns 0 if every cell of every step
 * succeeded, 1 otherwise. */
int camp_server_wait(void *server, long request, double *latency_s);

/* Host-resident pool of cells that camp_server_replay draws requests from
 * (copied into pinned memory): state [n][n_state], env [n][2], rates
 * [n][n_rates]. */
int camp_server_set_pool(void *server, int mech, long n_cells, const double *state,
                         const double *env, const double *base_rates);

/* Open-loop load generator: request i arrives t_arrive[i] seconds after the
 * call (ascending), asks for n_cells[i] cells of mechanism mech[i] starting
 * at pool row offset[i], n_steps[i] steps of dt[i] seconds; outputs go to a
 * pinned scratch area.  Fills latency_s[i]; *makespan_s is the time from
 * the call to the last completion.  Returns the number of failed requests. */
int camp_server_replay(void *server, int n_requests, const double *t_arrive, const int *mech,
                       const int *n_cells, const long *offset, const double *dt,
                       const int *n_steps, double *latency_s, double *makespan_s);

/* Counters since creation: [requests done, requests failed, cell-timesteps
 * done, batches launched, mean cells per batch, peak cells resident]. */
void camp_server_counters(void *server, double *out6);

/* Waits for every queued request, stops the scheduler, frees everything. */
void camp_server_free(void *server);

#ifdef __cplusplus
}
#endif

#endif
