/* CPU oracle for the CAMP chemistry hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (see camp_oracle.h for
 * the list of reference functions each part follows).  Used only as the
 * parity checker (tests/, __graft_entry__.smoke) and as the CPU baseline
 * in bench.py.  The product library never links or calls this file.
 *
 * Build: oracle/Makefile, with -ffp-contract=off so that every a*b+c is
 * two roundings, as in the reference's C99 build on x86-64.
 */
#include "camp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* camp_common.h:34-45, time_derivative.h:18 */
#define VARIABLE 1
#define SMALL 1.0e-30
#define TINY 1.0e-60
#define MAX_PRECISION_LOSS 1.0e-14
/* rxn_solver.c:19-35 */
#define RXN_ARRHENIUS 1
#define RXN_TROE 2
#define RXN_PHOTOLYSIS 5
#define RXN_FIRST_ORDER_LOSS 12
#define RXN_EMISSION 13
/* camp_solver.c:47 */
#define GUESS_MAX_ITER 5

/* --------------------------------------------------------------------- */
/* Sparse pattern builder: restates Jacobian.c:105-214                    */
/* (register unique (dep, ind) pairs per column, then sort each column).  */
/* --------------------------------------------------------------------- */
typedef struct {
  int n;
  int *cnt, *cap, **rows;
} ColElems;

typedef struct {
  int n, nnz;
  int *colptr, *rowval;
} Csc;

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    exit(EXIT_FAILURE);
  }
  return p;
}
static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    exit(EXIT_FAILURE);
  }
  return p;
}

static void cols_init(ColElems *c, int n) {
  c->n = n;
  c->cnt = xcalloc(n, sizeof(int));
  c->cap = xcalloc(n, sizeof(int));
  c->rows = xcalloc(n, sizeof(int *));
}

static void cols_register(ColElems *c, int dep, int ind) {
  for (int i = 0; i < c->cnt[ind]; ++i)
    if (c->rows[ind][i] == dep) return;
  if (c->cnt[ind] == c->cap[ind]) {
    c->cap[ind] += 10; /* Jacobian.c BUFFER_SIZE */
    c->rows[ind] = realloc(c->rows[ind], c->cap[ind] * sizeof(int));
  }
  c->rows[ind][c->cnt[ind]++] = dep;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static Csc cols_build(ColElems *c) {
  Csc m;
  m.n = c->n;
  m.nnz = 0;
  for (int j = 0; j < c->n; ++j) {
    qsort(c->rows[j], c->cnt[j], sizeof(int), cmp_int);
    m.nnz += c->cnt[j];
  }
  m.colptr = xmalloc((c->n + 1) * sizeof(int));
  m.rowval = xmalloc(m.nnz * sizeof(int));
  int e = 0;
  for (int j = 0; j < c->n; ++j) {
    m.colptr[j] = e;
    for (int i = 0; i < c->cnt[j]; ++i) m.rowval[e++] = c->rows[j][i];
    free(c->rows[j]);
  }
  m.colptr[c->n] = e;
  free(c->rows);
  free(c->cnt);
  free(c->cap);
  return m;
}

static int csc_find(const Csc *m, int dep, int ind) {
  for (int e = m->colptr[ind]; e < m->colptr[ind + 1]; ++e)
    if (m->rowval[e] == dep) return e;
  return -1;
}

/* --------------------------------------------------------------------- */
/* Model                                                                  */
/* --------------------------------------------------------------------- */
typedef struct {
  int ns, nc, nv;
  int *var_type, *deriv_ids, *dep_state; /* dep_state[d] = state index */
  int n_rxn, n_added, n_int_total, n_float_total;
  int *rxn_int, *rxn_int_idx, *rxn_float_idx, *rxn_env_idx, n_env;
  double *rxn_float, *rxn_env;
  double *abs_tol, rtol;
  int max_steps, max_conv_fails, initialized, guess_helper_on, neg_mode;
  Csc rjac, pjac, sjac;          /* rxn, param, solver Jacobians (state ids) */
  int nnz, *colptr, *rowval;     /* per-cell solver pattern (dep ids)   */
  int n_mapped, *map_solver, *map_rxn, *map_param;
  /* LU (static minimum-degree ordering) */
  int lu_nnz, *perm, *iperm, *lu_rowptr, *lu_col, *lu_diag, *m_to_lu;
  /* CAMP J history (camp_solver.c:877-879, 1111-1114) */
  double *J_solver, *J_state, *J_deriv;
  /* work */
  long double *td_p, *td_l, *jp, *jl;
  double *J_rxn, *est, *tmp1, *tmp2, *lu_w;
  double *state;                 /* borrowed total state during a run */
  const double *env;
  /* stats */
  long st[8];
  double h_last, h_next;
} Oracle;

void *oracle_new(int n_state_var, int n_cells, const int *var_type, int n_rxn,
                 int n_rxn_int_param, int n_rxn_float_param,
                 int n_rxn_env_param) {
  Oracle *o = xcalloc(1, sizeof(Oracle));
  o->ns = n_state_var;
  o->nc = n_cells;
  o->var_type = xmalloc(n_state_var * sizeof(int));
  memcpy(o->var_type, var_type, n_state_var * sizeof(int));
  o->nv = 0;
  for (int i = 0; i < n_state_var; ++i)
    if (var_type[i] == VARIABLE) o->nv++;
  o->n_rxn = n_rxn;
  o->rxn_int = xmalloc((n_rxn_int_param + n_rxn) * sizeof(int));
  o->rxn_float = xmalloc(n_rxn_float_param * sizeof(double));
  o->rxn_env = xcalloc((size_t)n_cells * n_rxn_env_param, sizeof(double));
  o->rxn_int_idx = xcalloc(n_rxn + 1, sizeof(int));
  o->rxn_float_idx = xcalloc(n_rxn + 1, sizeof(int));
  o->rxn_env_idx = xcalloc(n_rxn + 1, sizeof(int));
  o->n_int_total = n_rxn_int_param + n_rxn;
  o->n_float_total = n_rxn_float_param;
  o->guess_helper_on = 1;
  o->neg_mode = 1;
  return o;
}

/* rxn_solver.c:658-693 */
void oracle_add_rxn(void *om, int rxn_type, int n_int, int n_float, int n_env,
                    const int *int_param, const double *float_param) {
  Oracle *o = om;
  int k = o->n_added;
  int *id = &o->rxn_int[o->rxn_int_idx[k]];
  double *fd = &o->rxn_float[o->rxn_float_idx[k]];
  o->rxn_int_idx[k + 1] = o->rxn_int_idx[k] + n_int + 1;
  o->rxn_float_idx[k + 1] = o->rxn_float_idx[k] + n_float;
  o->rxn_env_idx[k + 1] = o->rxn_env_idx[k] + n_env;
  o->n_added++;
  *(id++) = rxn_type;
  for (int i = 0; i < n_int; ++i) id[i] = int_param[i];
  for (int i = 0; i < n_float; ++i) fd[i] = float_param[i];
  o->n_env += n_env;
}

#define RX_INT(o, r) (&(o)->rxn_int[(o)->rxn_int_idx[r]])
#define RX_FLT(o, r) (&(o)->rxn_float[(o)->rxn_float_idx[r]])

/* Per-type accessors for the packed layouts of the gas-phase types with
 * reactant/product lists (rxn_arrhenius.c:20-36, rxn_troe.c:20-39,
 * rxn_photolysis.c:20-33). */
static int nprop_int(int type) { return type == RXN_PHOTOLYSIS ? 3 : 2; }
static int nprop_flt(int type) {
  return type == RXN_ARRHENIUS ? 6 : type == RXN_TROE ? 10 : 1;
}
#define NR(d) ((d)[0])
#define NP(d) ((d)[1])
#define REACT(d, t, x) ((d)[nprop_int(t) + (x)] - 1)
#define PROD(d, t, x) ((d)[nprop_int(t) + NR(d) + (x)] - 1)
#define DERIV_ID(d, t, x) ((d)[nprop_int(t) + NR(d) + NP(d) + (x)])
#define JAC_ID(d, t, x) ((d)[nprop_int(t) + 2 * (NR(d) + NP(d)) + (x)])
#define YIELD(f, t, x) ((f)[nprop_flt(t) + (x)])

static int is_mass_action(int type) {
  return type == RXN_ARRHENIUS || type == RXN_TROE || type == RXN_PHOTOLYSIS;
}

/* rxn_get_used_jac_elem, rxn_solver.c:42-129 */
static void used_jac_elem(Oracle *o, ColElems *jac) {
  for (int r = 0; r < o->n_rxn; ++r) {
    int *id = RX_INT(o, r);
    int t = *(id++);
    if (is_mass_action(t)) {
      for (int i = 0; i < NR(id); ++i) {
        for (int d = 0; d < NR(id); ++d)
          cols_register(jac, REACT(id, t, d), REACT(id, t, i));
        for (int d = 0; d < NP(id); ++d)
          cols_register(jac, PROD(id, t, d), REACT(id, t, i));
      }
    } else if (t == RXN_FIRST_ORDER_LOSS) {
      cols_register(jac, id[1] - 1, id[1] - 1);
    }
  }
}

/* rxn_update_ids, rxn_solver.c:131-222 */
static void update_ids(Oracle *o) {
  for (int r = 0; r < o->n_rxn; ++r) {
    int *id = RX_INT(o, r);
    int t = *(id++);
    if (is_mass_action(t)) {
      for (int i = 0; i < NR(id); ++i)
        DERIV_ID(id, t, i) = o->deriv_ids[REACT(id, t, i)];
      for (int i = 0; i < NP(id); ++i)
        DERIV_ID(id, t, i + NR(id)) = o->deriv_ids[PROD(id, t, i)];
      int e = 0;
      for (int i = 0; i < NR(id); ++i) {
        for (int d = 0; d < NR(id); ++d)
          JAC_ID(id, t, e++) =
              csc_find(&o->rjac, REACT(id, t, d), REACT(id, t, i));
        for (int d = 0; d < NP(id); ++d)
          JAC_ID(id, t, e++) =
              csc_find(&o->rjac, PROD(id, t, d), REACT(id, t, i));
      }
    } else if (t == RXN_EMISSION) {
      id[2] = o->deriv_ids[id[1] - 1];
    } else if (t == RXN_FIRST_ORDER_LOSS) {
      id[2] = o->deriv_ids[id[1] - 1];
      id[3] = csc_find(&o->rjac, id[1] - 1, id[1] - 1);
    }
  }
}

static void build_lu(Oracle *o);

/* solver_initialize pattern part: get_jac_init, camp_solver.c:1417-1712 */
void oracle_initialize(void *om, const double *abs_tol, double rel_tol,
                       int max_steps, int max_conv_fails) {
  Oracle *o = om;
  int ns = o->ns;
  o->abs_tol = xmalloc(ns * sizeof(double));
  memcpy(o->abs_tol, abs_tol, ns * sizeof(double));
  o->rtol = rel_tol;
  o->max_steps = max_steps;
  o->max_conv_fails = max_conv_fails;

  ColElems rj;
  cols_init(&rj, ns);
  for (int i = 0; i < ns; ++i) cols_register(&rj, i, i); /* diagonal */
  used_jac_elem(o, &rj);
  o->rjac = cols_build(&rj);

  o->deriv_ids = xmalloc(ns * sizeof(int));
  o->dep_state = xmalloc((o->nv ? o->nv : 1) * sizeof(int));
  int d = 0;
  for (int i = 0; i < ns; ++i) {
    if (o->var_type[i] == VARIABLE) {
      o->dep_state[d] = i;
      o->deriv_ids[i] = d++;
    } else {
      o->deriv_ids[i] = -1;
    }
  }
  update_ids(o);

  /* sub-model parameter Jacobian: only the dummy (0,0) element */
  ColElems pj;
  cols_init(&pj, ns);
  cols_register(&pj, 0, 0);
  o->pjac = cols_build(&pj);

  /* solver Jacobian structure and number of mapped values */
  ColElems sj;
  cols_init(&sj, ns);
  int n_mapped = 0;
  for (int i_ind = 0; i_ind < ns; ++i_ind) {
    for (int i_dep = 0; i_dep < ns; ++i_dep) {
      if (o->var_type[i_dep] != VARIABLE || csc_find(&o->rjac, i_dep, i_ind) == -1)
        continue;
      if (o->var_type[i_ind] == VARIABLE) {
        cols_register(&sj, i_dep, i_ind);
        ++n_mapped;
        continue;
      }
      for (int j_ind = 0; j_ind < ns; ++j_ind)
        if (csc_find(&o->pjac, i_ind, j_ind) != -1 && o->var_type[j_ind] == VARIABLE) {
          cols_register(&sj, i_dep, j_ind);
          ++n_mapped;
        }
    }
  }
  o->sjac = cols_build(&sj);
  o->nnz = o->sjac.nnz;

  /* per-cell CSC over dependent variables (block structure, :1605-1633) */
  o->colptr = xmalloc((o->nv + 1) * sizeof(int));
  o->rowval = xmalloc((o->nnz ? o->nnz : 1) * sizeof(int));
  for (int c = 0; c < ns; ++c)
    if (o->deriv_ids[c] >= 0) o->colptr[o->deriv_ids[c]] = o->sjac.colptr[c];
  o->colptr[o->nv] = o->nnz;
  for (int e = 0; e < o->nnz; ++e) o->rowval[e] = o->deriv_ids[o->sjac.rowval[e]];

  /* jac_map, :1643-1683 */
  o->n_mapped = n_mapped;
  o->map_solver = xmalloc((n_mapped ? n_mapped : 1) * sizeof(int));
  o->map_rxn = xmalloc((n_mapped ? n_mapped : 1) * sizeof(int));
  o->map_param = xmalloc((n_mapped ? n_mapped : 1) * sizeof(int));
  int m = 0;
  for (int i_ind = 0; i_ind < ns; ++i_ind) {
    for (int e = o->rjac.colptr[i_ind]; e < o->rjac.colptr[i_ind + 1]; ++e) {
      int i_dep = o->rjac.rowval[e];
      if (o->var_type[i_dep] != VARIABLE || csc_find(&o->rjac, i_dep, i_ind) == -1)
        continue;
      if (o->var_type[i_ind] == VARIABLE) {
        o->map_solver[m] = csc_find(&o->sjac, i_dep, i_ind);
        o->map_rxn[m] = e;
        o->map_param[m] = 0;
        ++m;
        continue;
      }
      for (int j_ind = 0; j_ind < ns; ++j_ind)
        if (csc_find(&o->pjac, i_ind, j_ind) != -1 && o->var_type[j_ind] == VARIABLE) {
          o->map_solver[m] = csc_find(&o->sjac, i_dep, j_ind);
          o->map_rxn[m] = e;
          o->map_param[m] = csc_find(&o->pjac, i_ind, j_ind);
          ++m;
        }
    }
  }
  if (m != n_mapped) {
    fprintf(stderr, "oracle: internal jac_map error\n");
    exit(EXIT_FAILURE);
  }

  size_t nc = o->nc, nv = o->nv, nnz = o->nnz;
  o->J_solver = xcalloc(nc * nnz, sizeof(double));
  o->J_state = xcalloc(nc * nv, sizeof(double));
  o->J_deriv = xcalloc(nc * nv, sizeof(double));
  o->td_p = xcalloc(nv, sizeof(long double));
  o->td_l = xcalloc(nv, sizeof(long double));
  o->jp = xcalloc(o->rjac.nnz, sizeof(long double));
  o->jl = xcalloc(o->rjac.nnz, sizeof(long double));
  o->J_rxn = xcalloc(o->rjac.nnz, sizeof(double));
  o->est = xcalloc(nv, sizeof(double));
  o->tmp1 = xcalloc(nv, sizeof(double));
  o->tmp2 = xcalloc(nv, sizeof(double));
  o->lu_w = xcalloc(nv, sizeof(double));
  build_lu(o);
  o->initialized = 1;
}

void oracle_free(void *om) {
  Oracle *o = om;
  if (!o) return;
  free(o->var_type); free(o->deriv_ids); free(o->dep_state);
  free(o->rxn_int); free(o->rxn_float); free(o->rxn_env);
  free(o->rxn_int_idx); free(o->rxn_float_idx); free(o->rxn_env_idx);
  free(o->abs_tol);
  if (o->initialized) {
    free(o->rjac.colptr); free(o->rjac.rowval);
    free(o->pjac.colptr); free(o->pjac.rowval);
    free(o->sjac.colptr); free(o->sjac.rowval);
    free(o->colptr); free(o->rowval);
    free(o->map_solver); free(o->map_rxn); free(o->map_param);
    free(o->perm); free(o->iperm); free(o->lu_rowptr); free(o->lu_col);
    free(o->lu_diag); free(o->m_to_lu);
    free(o->J_solver); free(o->J_state); free(o->J_deriv);
    free(o->td_p); free(o->td_l); free(o->jp); free(o->jl); free(o->J_rxn);
    free(o->est); free(o->tmp1); free(o->tmp2); free(o->lu_w);
  }
  free(o);
}

/* --------------------------------------------------------------------- */
/* Rate constants and rate updates                                        */
/* --------------------------------------------------------------------- */
/* rxn_update_env_state, rxn_solver.c:224-313, and per type:
 * rxn_arrhenius.c:103-118, rxn_troe.c:106-130, rxn_photolysis.c:137-147,
 * rxn_emission.c:109-119, rxn_first_order_loss.c:119-129 */
static void update_env_cell(Oracle *o, int cell, const double *env_data) {
  double T = env_data[0], P = env_data[1];
  for (int r = 0; r < o->n_rxn; ++r) {
    int *id = RX_INT(o, r);
    int t = *(id++);
    double *f = RX_FLT(o, r);
    double *e = &o->rxn_env[(size_t)cell * o->n_env + o->rxn_env_idx[r]];
    switch (t) {
      case RXN_ARRHENIUS:
        e[0] = f[0] * exp(f[2] / T) * (f[1] == 0.0 ? 1.0 : pow(T / f[3], f[1])) *
               (f[4] == 0.0 ? 1.0 : (1.0 + f[4] * P)) *
               pow(f[5] * P / T, NR(id) - 1);
        break;
      case RXN_TROE: {
        double conv = f[9] * P / T;
        double k0 = f[0] * (f[2] == 0.0 ? 1.0 : exp(f[2] / T)) *
                    (f[1] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[1])) * conv;
        double kinf = k0 / (f[3] * (f[5] == 0.0 ? 1.0 : exp(f[5] / T)) *
                            (f[4] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[4])));
        e[0] = (k0 / (1.0 + kinf)) *
               pow(f[6], (1.0 / (1.0 + pow(log10(kinf) / f[7], 2)))) *
               pow(conv, NR(id) - 1) * f[8];
        break;
      }
      case RXN_PHOTOLYSIS:
      case RXN_EMISSION:
      case RXN_FIRST_ORDER_LOSS:
        e[0] = f[0] * e[1];
        break;
      default:
        fprintf(stderr, "oracle: unsupported reaction type %d\n", t);
        exit(EXIT_FAILURE);
    }
  }
}

void oracle_update_env_state(void *om, const double *env) {
  Oracle *o = om;
  for (int c = 0; c < o->nc; ++c) update_env_cell(o, c, &env[2 * c]);
}

/* rxn_update_data, rxn_solver.c:708-763, with rxn_photolysis.c:109-125,
 * rxn_emission.c:82-98, rxn_first_order_loss.c:89-106 */
int oracle_update_data(void *om, int cell_id, int *rxn_id, int rxn_type,
                       int unique_id, double rate) {
  Oracle *o = om;
  for (; *rxn_id < o->n_rxn; (*rxn_id)++) {
    int *id = RX_INT(o, *rxn_id);
    int t = *(id++);
    double *f = RX_FLT(o, *rxn_id);
    double *e = &o->rxn_env[(size_t)cell_id * o->n_env + o->rxn_env_idx[*rxn_id]];
    if (t != rxn_type) continue;
    int my_id = (t == RXN_PHOTOLYSIS) ? id[2] : id[0];
    if (t != RXN_PHOTOLYSIS && t != RXN_EMISSION && t != RXN_FIRST_ORDER_LOSS) continue;
    if (unique_id == my_id && my_id > 0) {
      e[1] = rate;
      e[0] = f[0] * e[1];
      return 1;
    }
  }
  return 0;
}

/* --------------------------------------------------------------------- */
/* f(y): camp_solver.c:843-958 with time_derivative.c and calc_deriv      */
/* --------------------------------------------------------------------- */
static void td_add(Oracle *o, int i, long double c) {
  if (c > 0.0) o->td_p[i] += c;
  else o->td_l[i] += -c;
}

/* one cell; state = the cell's (already updated) state vector */
static void cell_deriv(Oracle *o, int cell, const double *state, double *deriv,
                       const double *est, double time_step) {
  for (int i = 0; i < o->nv; ++i) o->td_p[i] = o->td_l[i] = 0.0;
  const double *env = &o->rxn_env[(size_t)cell * o->n_env];
  for (int r = 0; r < o->n_rxn; ++r) {
    int *id = RX_INT(o, r);
    int t = *(id++);
    double *f = RX_FLT(o, r);
    const double *e = env + o->rxn_env_idx[r];
    if (is_mass_action(t)) {
      long double rate = e[0];
      for (int i = 0; i < NR(id); ++i) rate *= state[REACT(id, t, i)];
      if (rate != 0.0) {
        int dv = 0;
        for (int i = 0; i < NR(id); ++i, ++dv) {
          if (DERIV_ID(id, t, dv) < 0) continue;
          td_add(o, DERIV_ID(id, t, dv), -rate);
        }
        for (int i = 0; i < NP(id); ++i, ++dv) {
          if (DERIV_ID(id, t, dv) < 0) continue;
          if (-rate * YIELD(f, t, i) * time_step <= state[PROD(id, t, i)])
            td_add(o, DERIV_ID(id, t, dv), rate * YIELD(f, t, i));
        }
      }
    } else if (t == RXN_EMISSION) {
      if (id[2] >= 0) td_add(o, id[2], (long double)e[0]);
    } else if (t == RXN_FIRST_ORDER_LOSS) {
      long double rate = e[0] * state[id[1] - 1];
      if (id[2] >= 0) td_add(o, id[2], -rate);
    }
  }
  /* time_derivative_output, time_derivative.c:46-88 */
  for (int i = 0; i < o->nv; ++i) {
    long double rp = o->td_p[i], rl = o->td_l[i];
    if (rp + rl != 0.0) {
      if (est) {
        long double sf = 1.0 / (rp + rl) /
                         (1.0 / (rp + rl) + MAX_PRECISION_LOSS / fabsl(rp - rl));
        deriv[i] = sf * (rp - rl) + (1.0 - sf) * est[i];
      } else {
        deriv[i] = rp - rl;
      }
    } else {
      deriv[i] = 0.0;
    }
  }
}

/* camp_solver_update_model_state (camp_solver.c:805-833) for one cell */
static int update_cell_state(Oracle *o, double *cell_state, const double *y) {
  for (int d = 0; d < o->nv; ++d)
    if (y[d] < -SMALL) return 1;
  for (int d = 0; d < o->nv; ++d)
    cell_state[o->dep_state[d]] = y[d] > -SMALL ? y[d] : TINY;
  return 0;
}

/* Jacobian-estimated derivative J_deriv + J_solver (y - J_state) */
static void cell_est(Oracle *o, int cell, const double *y, double *est) {
  const double *Js = &o->J_solver[(size_t)cell * o->nnz];
  const double *ys = &o->J_state[(size_t)cell * o->nv];
  const double *fs = &o->J_deriv[(size_t)cell * o->nv];
  for (int i = 0; i < o->nv; ++i) {
    o->tmp1[i] = y[i] - ys[i];
    o->tmp2[i] = 0.0;
  }
  for (int j = 0; j < o->nv; ++j)
    for (int k = o->colptr[j]; k < o->colptr[j + 1]; ++k)
      o->tmp2[o->rowval[k]] += Js[k] * o->tmp1[j];
  for (int i = 0; i < o->nv; ++i) est[i] = fs[i] + o->tmp2[i];
}

/* f over cells [c0, c0+nc) of the borrowed state; y, deriv: nc*nv */
static int range_f(Oracle *o, int c0, int nc, const double *y, double *deriv,
                   double time_step, int use_est) {
  int nv = o->nv;
  for (int c = 0; c < nc; ++c)
    for (int d = 0; d < nv; ++d)
      if (y[c * nv + d] < -SMALL) return 1;
  for (int c = 0; c < nc; ++c)
    update_cell_state(o, &o->state[(size_t)(c0 + c) * o->ns], &y[c * nv]);
  for (int c = 0; c < nc; ++c) {
    if (use_est) cell_est(o, c0 + c, &y[c * nv], o->est);
    cell_deriv(o, c0 + c, &o->state[(size_t)(c0 + c) * o->ns], &deriv[c * nv],
               use_est ? o->est : NULL, time_step);
  }
  return 0;
}

/* J(y): camp_solver.c:972-1126 with Jacobian.c:216-241 and calc_jac */
static void cell_jac(Oracle *o, int cell, const double *state, double *J,
                     double time_step) {
  for (int e = 0; e < o->rjac.nnz; ++e) o->jp[e] = o->jl[e] = 0.0;
  const double *env = &o->rxn_env[(size_t)cell * o->n_env];
  for (int r = 0; r < o->n_rxn; ++r) {
    int *id = RX_INT(o, r);
    int t = *(id++);
    double *f = RX_FLT(o, r);
    const double *e = env + o->rxn_env_idx[r];
    if (is_mass_action(t)) {
      int el = 0;
      for (int i_ind = 0; i_ind < NR(id); ++i_ind) {
        double rate = e[0];
        for (int i = 0; i < NR(id); ++i)
          if (i != i_ind) rate *= state[REACT(id, t, i)];
        for (int d = 0; d < NR(id); ++d, ++el) {
          if (JAC_ID(id, t, el) < 0) continue;
          o->jl[JAC_ID(id, t, el)] += rate;
        }
        for (int d = 0; d < NP(id); ++d, ++el) {
          if (JAC_ID(id, t, el) < 0) continue;
          if (-rate * state[REACT(id, t, i_ind)] * YIELD(f, t, d) * time_step <=
              state[PROD(id, t, d)])
            o->jp[JAC_ID(id, t, el)] += YIELD(f, t, d) * rate;
        }
      }
    } else if (t == RXN_FIRST_ORDER_LOSS) {
      if (id[3] >= 0) o->jl[id[3]] += e[0];
    }
  }
  for (int e = 0; e < o->rjac.nnz; ++e) o->J_rxn[e] = o->jp[e] - o->jl[e];
  for (int k = 0; k < o->nnz; ++k) J[k] = 0.0;
  /* J_params[0] = 1 is the only parameter value without sub models */
  for (int m = 0; m < o->n_mapped; ++m)
    J[o->map_solver[m]] += o->J_rxn[o->map_rxn[m]] * (o->map_param[m] == 0 ? 1.0 : 0.0);
}

static int range_jac(Oracle *o, int c0, int nc, const double *y, double *fy,
                     double *J, double time_step) {
  if (range_f(o, c0, nc, y, fy, time_step, 0)) return 1;
  for (int c = 0; c < nc; ++c)
    update_cell_state(o, &o->state[(size_t)(c0 + c) * o->ns], &y[c * o->nv]);
  for (int c = 0; c < nc; ++c)
    cell_jac(o, c0 + c, &o->state[(size_t)(c0 + c) * o->ns], &J[(size_t)c * o->nnz],
             time_step);
  memcpy(&o->J_solver[(size_t)c0 * o->nnz], J, (size_t)nc * o->nnz * sizeof(double));
  memcpy(&o->J_state[(size_t)c0 * o->nv], y, (size_t)nc * o->nv * sizeof(double));
  memcpy(&o->J_deriv[(size_t)c0 * o->nv], fy, (size_t)nc * o->nv * sizeof(double));
  return 0;
}

/* --------------------------------------------------------------------- */
/* Sparse LU with a static minimum-degree ordering (stands in for KLU)    */
/* --------------------------------------------------------------------- */
static void build_lu(Oracle *o) {
  int n = o->nv;
  o->perm = xmalloc((n ? n : 1) * sizeof(int));
  o->iperm = xmalloc((n ? n : 1) * sizeof(int));
  char *G = xcalloc((size_t)n * n, 1);
  for (int j = 0; j < n; ++j)
    for (int k = o->colptr[j]; k < o->colptr[j + 1]; ++k) {
      int i = o->rowval[k];
      if (i != j) G[i * n + j] = G[j * n + i] = 1;
    }
  char *alive = xmalloc(n ? n : 1);
  memset(alive, 1, n);
  for (int s = 0; s < n; ++s) { /* minimum degree on A + A^T */
    int best = -1, bdeg = 0;
    for (int v = 0; v < n; ++v) {
      if (!alive[v]) continue;
      int deg = 0;
      for (int u = 0; u < n; ++u) deg += alive[u] && G[v * n + u];
      if (best < 0 || deg < bdeg) best = v, bdeg = deg;
    }
    o->perm[s] = best;
    alive[best] = 0;
    for (int a = 0; a < n; ++a)
      if (alive[a] && G[best * n + a])
        for (int b = 0; b < n; ++b)
          if (b != a && alive[b] && G[best * n + b]) G[a * n + b] = 1;
  }
  free(alive);
  free(G);
  for (int s = 0; s < n; ++s) o->iperm[o->perm[s]] = s;
  /* symbolic factorization on the permuted pattern */
  char *F = xcalloc((size_t)n * n, 1);
  for (int j = 0; j < n; ++j)
    for (int k = o->colptr[j]; k < o->colptr[j + 1]; ++k)
      F[o->iperm[o->rowval[k]] * n + o->iperm[j]] = 1;
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i)
      if (F[i * n + k])
        for (int j = k + 1; j < n; ++j)
          if (F[k * n + j]) F[i * n + j] = 1;
  o->lu_rowptr = xmalloc((n + 1) * sizeof(int));
  int nnz = 0;
  for (int i = 0; i < n * n; ++i) nnz += F[i];
  o->lu_nnz = nnz;
  o->lu_col = xmalloc((nnz ? nnz : 1) * sizeof(int));
  o->lu_diag = xmalloc((n ? n : 1) * sizeof(int));
  int p = 0;
  for (int i = 0; i < n; ++i) {
    o->lu_rowptr[i] = p;
    for (int j = 0; j < n; ++j)
      if (F[i * n + j]) {
        if (i == j) o->lu_diag[i] = p;
        o->lu_col[p++] = j;
      }
  }
  o->lu_rowptr[n] = p;
  o->m_to_lu = xmalloc((o->nnz ? o->nnz : 1) * sizeof(int));
  for (int j = 0; j < n; ++j)
    for (int k = o->colptr[j]; k < o->colptr[j + 1]; ++k) {
      int pi = o->iperm[o->rowval[k]], pj = o->iperm[j];
      for (int q = o->lu_rowptr[pi]; q < o->lu_rowptr[pi + 1]; ++q)
        if (o->lu_col[q] == pj) o->m_to_lu[k] = q;
    }
  free(F);
}

/* M = I - gamma*J (SUNMatScaleAddI), then factor; returns 1 on a zero
 * pivot (a recoverable linear-setup failure). */
static int lu_factor(Oracle *o, const double *J, double gamma, double *lu) {
  int n = o->nv;
  for (int q = 0; q < o->lu_nnz; ++q) lu[q] = 0.0;
  for (int j = 0; j < n; ++j)
    for (int k = o->colptr[j]; k < o->colptr[j + 1]; ++k) {
      double v = -gamma * J[k];
      if (o->rowval[k] == j) v = v + 1.0;
      lu[o->m_to_lu[k]] = v;
    }
  double *w = o->lu_w;
  for (int i = 0; i < n; ++i) {
    for (int q = o->lu_rowptr[i]; q < o->lu_rowptr[i + 1]; ++q) w[o->lu_col[q]] = lu[q];
    for (int q = o->lu_rowptr[i]; q < o->lu_diag[i]; ++q) {
      int k = o->lu_col[q];
      double lik = w[k] / lu[o->lu_diag[k]];
      w[k] = lik;
      for (int s = o->lu_diag[k] + 1; s < o->lu_rowptr[k + 1]; ++s)
        w[o->lu_col[s]] -= lik * lu[s];
    }
    for (int q = o->lu_rowptr[i]; q < o->lu_rowptr[i + 1]; ++q) lu[q] = w[o->lu_col[q]];
    if (lu[o->lu_diag[i]] == 0.0 || !isfinite(lu[o->lu_diag[i]])) return 1;
  }
  return 0;
}

static void lu_solve(Oracle *o, const double *lu, double *b) {
  int n = o->nv;
  double *x = o->lu_w;
  for (int i = 0; i < n; ++i) x[i] = b[o->perm[i]];
  for (int i = 0; i < n; ++i)
    for (int q = o->lu_rowptr[i]; q < o->lu_diag[i]; ++q) x[i] -= lu[q] * x[o->lu_col[q]];
  for (int i = n - 1; i >= 0; --i) {
    for (int q = o->lu_diag[i] + 1; q < o->lu_rowptr[i + 1]; ++q)
      x[i] -= lu[q] * x[o->lu_col[q]];
    x[i] /= lu[o->lu_diag[i]];
  }
  for (int i = 0; i < n; ++i) b[o->perm[i]] = x[i];
}

/* --------------------------------------------------------------------- */
/* BDF integrator: CVODE algorithm (cvode.c, CVODE 3.x) restated          */
/* --------------------------------------------------------------------- */
#define Q_MAX 5
#define L_MAX (Q_MAX + 1)
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 0.000001
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define ONEPSM 1.000001
#define THRESH 1.5
#define SMALL_NST 10
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define CRDOWN 0.3
#define DGMAX 0.3
#define RDIV 2.0
#define MSBP 20
#define NLSCOEF 0.1
#define MAXCOR 3
#define MSBJ 50
#define DGMAX_JBAD 0.2

enum { CV_SUCCESS = 0, CV_TOO_MUCH_WORK = -1, CV_TOO_MUCH_ACC = -2,
       CV_ERR_FAILURE = -3, CV_CONV_FAILURE = -4, CV_LSETUP_FAIL = -6,
       CV_LSOLVE_FAIL = -7, CV_RHSFUNC_FAIL = -8, CV_FIRST_RHSFUNC_ERR = -9,
       CV_REPTD_RHSFUNC_ERR = -10, CV_UNREC_RHSFUNC_ERR = -11,
       CV_ILL_INPUT = -22 };
enum { FIRST_CALL = 1, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN, CONV_FAIL, TRY_AGAIN, RHSFUNC_RECVR };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

typedef struct {
  Oracle *o;
  int c0, nc, n;
  double *zn[L_MAX], *ewt, *acor, *y, *tempv, *ftemp, *last_yn, *gh_tmp,
      *gh_corr, *acor_init, *Jw, *lu;
  double tn, h, hscale, hprime, next_h, eta, etamax, etaq, etaqm1, etaqp1, hu;
  double l[L_MAX + 1], tq[6], tau[L_MAX + 2];
  double rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5, init_h;
  int q, qprime, next_q, L, qwait, qu, indx_acor, jcur, use_acor_init;
  long nst, nfe, nsetups, nni, ncfn, netf, nje, nstlp, nstlj, nscon;
} Cv;

static double wrms(const Cv *cv, const double *x) {
  double s = 0.0;
  for (int i = 0; i < cv->n; ++i) {
    double p = x[i] * cv->ewt[i];
    s += p * p;
  }
  return sqrt(s / cv->n);
}

static double rpowerr(double b, double e) { return b <= 0.0 ? 0.0 : pow(b, e); }
static double rpoweri(double b, int e) {
  double p = 1.0;
  for (int i = 1; i <= abs(e); ++i) p *= b;
  return e < 0 ? 1.0 / p : p;
}

static int ewt_set(Cv *cv, const double *ycur) {
  Oracle *o = cv->o;
  for (int c = 0; c < cv->nc; ++c)
    for (int d = 0; d < o->nv; ++d) {
      int i = c * o->nv + d;
      double t = o->rtol * fabs(ycur[i]) + o->abs_tol[o->dep_state[d]];
      if (t <= 0.0) return -1;
      cv->ewt[i] = 1.0 / t;
    }
  return 0;
}

/* f as CVODE calls it; the time step is CVodeGetCurrentStep (next_h) */
static int cv_f(Cv *cv, const double *y, double *ydot) {
  double ts = cv->next_h > 0.0 ? cv->next_h : cv->init_h;
  cv->nfe++;
  return range_f(cv->o, cv->c0, cv->nc, y, ydot, ts, 1);
}

/* guess_helper, camp_solver.c:1321-1408 */
static int guess_helper(Cv *cv, double t_n, double h_n, const double *y_n,
                        const double *y_n1, double *hf, double *tmp1, double *corr) {
  int n = cv->n;
  double mn = y_n[0];
  for (int i = 1; i < n; ++i) mn = y_n[i] < mn ? y_n[i] : mn;
  if (mn > -SMALL) return 0;
  for (int i = 0; i < n; ++i) tmp1[i] = y_n1[i];
  if (h_n > 0.0) {
    double c = 1.0 / h_n;
    for (int i = 0; i < n; ++i) corr[i] = c * hf[i];
  } else {
    for (int i = 0; i < n; ++i) corr[i] = hf[i];
  }
  double t_0 = h_n > 0.0 ? t_n - h_n : t_n - 1.0;
  double t_j = 0.0;
  int iter = 0;
  for (; iter < GUESS_MAX_ITER && t_0 + t_j < t_n; iter++) {
    double h_j = t_n - (t_0 + t_j);
    int i_fast = -1;
    for (int i = 0; i < n; ++i) {
      double t_star = -tmp1[i] / corr[i];
      if ((t_star > 0.0 || (t_star == 0.0 && corr[i] < 0.0)) && t_star < h_j) {
        h_j = t_star;
        i_fast = i;
      }
    }
    if (i_fast >= 0 && h_n > 0.0) h_j *= 0.95 + 0.1 * iter / (double)GUESS_MAX_ITER;
    h_j = t_n < t_0 + t_j + h_j ? t_n - (t_0 + t_j) : h_j;
    if (h_n == 0.0 && t_n - (h_j + t_j + t_0) > cv->o->rtol) return -1;
    for (int i = 0; i < n; ++i) tmp1[i] += h_j * corr[i];
    t_j += h_j;
    if (cv_f(cv, tmp1, corr) != 0) {
      for (int i = 0; i < n; ++i) corr[i] = 0.0;
      return -1;
    }
    if (iter == GUESS_MAX_ITER - 1 && t_0 + t_j < t_n)
      if (h_n == 0.0) return -1;
  }
  for (int i = 0; i < n; ++i) corr[i] = tmp1[i] - y_n[i];
  if (h_n > 0.0)
    for (int i = 0; i < n; ++i) corr[i] *= 0.999;
  for (int i = 0; i < n; ++i) hf[i] = tmp1[i] - y_n1[i];
  return 1;
}

static void cv_rescale(Cv *cv) {
  double factor = cv->eta;
  for (int j = 1; j <= cv->q; ++j) {
    for (int i = 0; i < cv->n; ++i) cv->zn[j][i] *= factor;
    factor *= cv->eta;
  }
  cv->h = cv->hscale * cv->eta;
  cv->next_h = cv->h;
  cv->hscale = cv->h;
  cv->nscon = 0;
}

static void cv_predict(Cv *cv) {
  cv->tn += cv->h;
  for (int k = 1; k <= cv->q; ++k)
    for (int j = cv->q; j >= k; --j)
      for (int i = 0; i < cv->n; ++i) cv->zn[j - 1][i] += cv->zn[j][i];
}

static void cv_restore(Cv *cv, double saved_t) {
  cv->tn = saved_t;
  for (int k = 1; k <= cv->q; ++k)
    for (int j = cv->q; j >= k; --j)
      for (int i = 0; i < cv->n; ++i) cv->zn[j - 1][i] -= cv->zn[j][i];
}

static void cv_set_tq(Cv *cv, double hsum, double alpha0, double alpha0_hat,
                      double xi_inv, double xistar_inv) {
  int q = cv->q;
  double A1 = 1.0 - alpha0_hat + alpha0;
  double A2 = 1.0 + q * A1;
  cv->tq[2] = fabs(A1 / (alpha0 * A2));
  cv->tq[5] = fabs(A2 * xistar_inv / (cv->l[q] * xi_inv));
  if (cv->qwait == 1) {
    if (q > 1) {
      double C = xistar_inv / cv->l[q];
      double A3 = alpha0 + 1.0 / q;
      double A4 = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      cv->tq[1] = fabs(C * Cpinv);
    } else {
      cv->tq[1] = 1.0;
    }
    hsum += cv->tau[q];
    xi_inv = cv->h / hsum;
    double A5 = alpha0 - (1.0 / (q + 1));
    double A6 = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    cv->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  cv->tq[4] = NLSCOEF / cv->tq[2];
}

static void cv_set(Cv *cv) {
  int q = cv->q;
  double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  cv->l[0] = cv->l[1] = xi_inv = xistar_inv = 1.0;
  for (int z = 2; z <= q; ++z) cv->l[z] = 0.0;
  alpha0 = alpha0_hat = -1.0;
  hsum = cv->h;
  if (q > 1) {
    for (int j = 2; j < q; ++j) {
      hsum += cv->tau[j - 1];
      xi_inv = cv->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; --i) cv->l[i] += cv->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -cv->l[1] - alpha0;
    hsum += cv->tau[q - 1];
    xi_inv = cv->h / hsum;
    alpha0_hat = -cv->l[1] - xi_inv;
    for (int i = q; i >= 1; --i) cv->l[i] += cv->l[i - 1] * xistar_inv;
  }
  cv_set_tq(cv, hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
  cv->rl1 = 1.0 / cv->l[1];
  cv->gamma = cv->h * cv->rl1;
  if (cv->nst == 0) cv->gammap = cv->gamma;
  cv->gamrat = (cv->nst > 0) ? cv->gamma / cv->gammap : 1.0;
}

static void cv_increase_bdf(Cv *cv) {
  double alpha0, alpha1, prod, xi, xiold, hsum, A1;
  for (int j = 0; j <= Q_MAX; ++j) cv->l[j] = 0.0;
  cv->l[2] = alpha1 = prod = xiold = 1.0;
  alpha0 = -1.0;
  hsum = cv->hscale;
  if (cv->q > 1) {
    for (int j = 1; j < cv->q; ++j) {
      hsum += cv->tau[j + 1];
      xi = hsum / cv->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; --i) cv->l[i] = cv->l[i] * xiold + cv->l[i - 1];
      xiold = xi;
    }
  }
  A1 = (-alpha0 - alpha1) / prod;
  for (int i = 0; i < cv->n; ++i) cv->zn[cv->L][i] = A1 * cv->zn[cv->indx_acor][i];
  for (int j = 2; j <= cv->q; ++j)
    for (int i = 0; i < cv->n; ++i) cv->zn[j][i] += cv->l[j] * cv->zn[cv->L][i];
}

static void cv_decrease_bdf(Cv *cv) {
  double hsum, xi;
  for (int j = 0; j <= Q_MAX; ++j) cv->l[j] = 0.0;
  cv->l[2] = 1.0;
  hsum = 0.0;
  for (int j = 1; j <= cv->q - 2; ++j) {
    hsum += cv->tau[j];
    xi = hsum / cv->hscale;
    for (int i = j + 2; i >= 2; --i) cv->l[i] = cv->l[i] * xi + cv->l[i - 1];
  }
  for (int j = 2; j < cv->q; ++j)
    for (int i = 0; i < cv->n; ++i) cv->zn[j][i] += (-cv->l[j]) * cv->zn[cv->q][i];
}

static void cv_adjust_params(Cv *cv) {
  if (cv->qprime != cv->q) {
    int deltaq = cv->qprime - cv->q;
    if (deltaq == 1) cv_increase_bdf(cv);
    else if (deltaq == -1) cv_decrease_bdf(cv);
    cv->q = cv->qprime;
    cv->L = cv->q + 1;
    cv->qwait = cv->L;
  }
  cv_rescale(cv);
}

/* CVDls lsetup (cvDlsSetup) with the KLU setup replaced by lu_factor */
static int cv_lsetup(Cv *cv, int convfail, double *ypred, double *fpred) {
  Oracle *o = cv->o;
  double dgamma = fabs(cv->gamma / cv->gammap - 1.0);
  int jbad = (cv->nst == 0) || (cv->nst > cv->nstlj + MSBJ) ||
             ((convfail == CV_FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) ||
             (convfail == CV_FAIL_OTHER);
  double *J = cv->Jw;
  if (!jbad) {
    cv->jcur = 0;
    memcpy(J, &o->J_solver[(size_t)cv->c0 * o->nnz], (size_t)cv->nc * o->nnz * sizeof(double));
  } else {
    cv->nje++;
    cv->nstlj = cv->nst;
    cv->jcur = 1;
    double ts = cv->next_h > 0.0 ? cv->next_h : cv->init_h;
    /* Jac() computes f(ypred) (without the estimate) into fpred first */
    cv->nfe += 0;
    if (range_jac(o, cv->c0, cv->nc, ypred, fpred, J, ts)) return 1;
  }
  for (int c = 0; c < cv->nc; ++c)
    if (lu_factor(o, &J[(size_t)c * o->nnz], cv->gamma, &cv->lu[(size_t)c * o->lu_nnz]))
      return 1;
  return 0;
}

static int cv_lsolve(Cv *cv, double *b) {
  Oracle *o = cv->o;
  for (int c = 0; c < cv->nc; ++c) lu_solve(o, &cv->lu[(size_t)c * o->lu_nnz], &b[c * o->nv]);
  if (cv->gamrat != 1.0) {
    double s = 2.0 / (1.0 + cv->gamrat);
    for (int i = 0; i < cv->n; ++i) b[i] *= s;
  }
  return 0;
}

static int cv_newton_iteration(Cv *cv) {
  int m = 0;
  double del = 0.0, delp = 0.0, dcon;
  for (;;) {
    double *b = cv->tempv;
    for (int i = 0; i < cv->n; ++i) b[i] = cv->rl1 * cv->zn[1][i] + cv->acor[i];
    for (int i = 0; i < cv->n; ++i) b[i] = cv->gamma * cv->ftemp[i] - b[i];
    int retval = cv_lsolve(cv, b);
    cv->nni++;
    if (retval < 0) return CV_LSOLVE_FAIL;
    if (retval > 0) return cv->jcur ? CONV_FAIL : TRY_AGAIN;
    del = wrms(cv, b);
    for (int i = 0; i < cv->n; ++i) cv->acor[i] += b[i];
    for (int i = 0; i < cv->n; ++i) cv->y[i] = cv->zn[0][i] + cv->acor[i];
    if (m > 0) cv->crate = fmax(CRDOWN * cv->crate, del / delp);
    dcon = del * fmin(1.0, cv->crate) / cv->tq[4];
    if (dcon <= 1.0) {
      /* Non-negativity safeguard (not in CVODE): CAMP's f() rejects
       * y < -SMALL (camp_solver.c:816-822), so a converged iterate below
       * that would make every later f() at the prediction fail; it is
       * treated as the recoverable f() failure it implies. */
      if (cv->o->neg_mode == 1) {
        double mn = cv->y[0];
        for (int i = 1; i < cv->n; ++i) mn = cv->y[i] < mn ? cv->y[i] : mn;
        if (mn < -SMALL) return cv->jcur ? RHSFUNC_RECVR : TRY_AGAIN;
      }
      cv->acnrm = (m == 0) ? del : wrms(cv, cv->acor);
      cv->jcur = 0;
      return CV_SUCCESS;
    }
    m++;
    if ((m == MAXCOR) || ((m >= 2) && (del > RDIV * delp)))
      return cv->jcur ? CONV_FAIL : TRY_AGAIN;
    delp = del;
    retval = cv_f(cv, cv->y, cv->ftemp);
    if (retval > 0) return cv->jcur ? RHSFUNC_RECVR : TRY_AGAIN;
  }
}

/* cvNlsNewton.  The guess helper (CAMP's CVODE fork hook) is applied once
 * per call, at the predictor: when the prediction has a component below
 * -SMALL, the helper's correction becomes the Newton starting correction
 * and f/J are evaluated at the corrected prediction. */
static int cv_nls(Cv *cv, int nflag) {
  int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES
                                                                     : CV_FAIL_OTHER;
  int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) ||
                  (cv->nst == 0) || (cv->nst >= cv->nstlp + MSBP) ||
                  (fabs(cv->gamrat - 1.0) > DGMAX);
  cv->use_acor_init = 0;
  if (cv->o->guess_helper_on) {
    for (int i = 0; i < cv->n; ++i) cv->acor_init[i] = cv->zn[0][i] - cv->last_yn[i];
    int gh = guess_helper(cv, cv->tn, cv->h, cv->zn[0], cv->last_yn, cv->acor_init,
                          cv->gh_tmp, cv->gh_corr);
    if (gh == 1) {
      memcpy(cv->acor_init, cv->gh_corr, cv->n * sizeof(double));
      cv->use_acor_init = 1;
    }
  }
  for (;;) {
    if (cv->use_acor_init) {
      for (int i = 0; i < cv->n; ++i) cv->y[i] = cv->zn[0][i] + cv->acor_init[i];
    } else {
      memcpy(cv->y, cv->zn[0], cv->n * sizeof(double));
    }
    int retval = cv_f(cv, cv->y, cv->ftemp);
    if (retval > 0) return RHSFUNC_RECVR;
    if (callSetup) {
      int ier = cv_lsetup(cv, convfail, cv->y, cv->ftemp);
      cv->nsetups++;
      callSetup = 0;
      cv->gamrat = cv->crate = 1.0;
      cv->gammap = cv->gamma;
      cv->nstlp = cv->nst;
      if (ier < 0) return CV_LSETUP_FAIL;
      if (ier > 0) return CONV_FAIL;
    }
    if (cv->use_acor_init) {
      memcpy(cv->acor, cv->acor_init, cv->n * sizeof(double));
    } else {
      for (int i = 0; i < cv->n; ++i) cv->acor[i] = 0.0;
    }
    int ier = cv_newton_iteration(cv);
    if (ier != TRY_AGAIN) return ier;
    callSetup = 1;
    convfail = CV_FAIL_BAD_J;
  }
}

static int cv_handle_nflag(Cv *cv, int *nflagPtr, double saved_t, int *ncfPtr) {
  int nflag = *nflagPtr;
  if (nflag == CV_SUCCESS) return DO_ERROR_TEST;
  cv->ncfn++;
  cv_restore(cv, saved_t);
  if (nflag == CV_LSETUP_FAIL) return CV_LSETUP_FAIL;
  if (nflag == CV_LSOLVE_FAIL) return CV_LSOLVE_FAIL;
  if (nflag == CV_RHSFUNC_FAIL) return CV_RHSFUNC_FAIL;
  (*ncfPtr)++;
  cv->etamax = 1.0;
  if (*ncfPtr == cv->o->max_conv_fails) {
    if (nflag == CONV_FAIL) return CV_CONV_FAILURE;
    if (nflag == RHSFUNC_RECVR) return CV_REPTD_RHSFUNC_ERR;
  }
  cv->eta = ETACF; /* max(ETACF, hmin/|h|) with hmin = 0 */
  *nflagPtr = PREV_CONV_FAIL;
  cv_rescale(cv);
  return PREDICT_AGAIN;
}

static int cv_do_error_test(Cv *cv, int *nflagPtr, double saved_t, int *nefPtr,
                            double *dsmPtr) {
  double dsm = cv->acnrm * cv->tq[2];
  *dsmPtr = dsm;
  if (dsm <= 1.0) return CV_SUCCESS;
  (*nefPtr)++;
  cv->netf++;
  *nflagPtr = PREV_ERR_FAIL;
  cv_restore(cv, saved_t);
  if (*nefPtr == cv->o->max_conv_fails) return CV_ERR_FAILURE;
  cv->etamax = 1.0;
  if (*nefPtr <= MXNEF1) {
    cv->eta = 1.0 / (rpowerr(BIAS2 * dsm, 1.0 / cv->L) + ADDON);
    cv->eta = fmax(ETAMIN, fmax(cv->eta, 0.0));
    if (*nefPtr >= SMALL_NEF) cv->eta = fmin(cv->eta, ETAMXF);
    cv_rescale(cv);
    return TRY_AGAIN;
  }
  if (cv->q > 1) {
    cv->eta = ETAMIN;
    cv_decrease_bdf(cv);
    cv->L = cv->q;
    cv->q--;
    cv->qwait = cv->L;
    cv_rescale(cv);
    return TRY_AGAIN;
  }
  cv->eta = ETAMIN;
  cv->h *= cv->eta;
  cv->next_h = cv->h;
  cv->hscale = cv->h;
  cv->qwait = LONG_WAIT;
  cv->nscon = 0;
  if (cv_f(cv, cv->zn[0], cv->tempv) > 0) return CV_UNREC_RHSFUNC_ERR;
  for (int i = 0; i < cv->n; ++i) cv->zn[1][i] = cv->h * cv->tempv[i];
  return TRY_AGAIN;
}

static void cv_complete_step(Cv *cv) {
  cv->nst++;
  cv->nscon++;
  cv->hu = cv->h;
  cv->qu = cv->q;
  for (int i = cv->q; i >= 2; --i) cv->tau[i] = cv->tau[i - 1];
  if ((cv->q == 1) && (cv->nst > 1)) cv->tau[2] = cv->tau[1];
  cv->tau[1] = cv->h;
  for (int j = 0; j <= cv->q; ++j)
    for (int i = 0; i < cv->n; ++i) cv->zn[j][i] += cv->l[j] * cv->acor[i];
  cv->qwait--;
  if ((cv->qwait == 1) && (cv->q != Q_MAX)) {
    memcpy(cv->zn[Q_MAX], cv->acor, cv->n * sizeof(double));
    cv->saved_tq5 = cv->tq[5];
    cv->indx_acor = Q_MAX;
  }
  memcpy(cv->last_yn, cv->zn[0], cv->n * sizeof(double));
}

static void cv_set_eta(Cv *cv) {
  if (cv->eta < THRESH) {
    cv->eta = 1.0;
    cv->hprime = cv->h;
  } else {
    cv->eta = fmin(cv->eta, cv->etamax);
    cv->hprime = cv->h * cv->eta;
    if (cv->qprime < cv->q) cv->nscon = 0;
  }
}

static void cv_prepare_next_step(Cv *cv, double dsm) {
  if (cv->etamax == 1.0) {
    cv->qwait = cv->qwait > 2 ? cv->qwait : 2;
    cv->qprime = cv->q;
    cv->hprime = cv->h;
    cv->eta = 1.0;
    return;
  }
  cv->etaq = 1.0 / (rpowerr(BIAS2 * dsm, 1.0 / cv->L) + ADDON);
  if (cv->qwait != 0) {
    cv->eta = cv->etaq;
    cv->qprime = cv->q;
    cv_set_eta(cv);
    return;
  }
  cv->qwait = 2;
  cv->etaqm1 = 0.0;
  if (cv->q > 1) {
    double ddn = wrms(cv, cv->zn[cv->q]) * cv->tq[1];
    cv->etaqm1 = 1.0 / (rpowerr(BIAS1 * ddn, 1.0 / cv->q) + ADDON);
  }
  cv->etaqp1 = 0.0;
  if (cv->q != Q_MAX && cv->saved_tq5 != 0.0) {
    double cquot = (cv->tq[5] / cv->saved_tq5) * rpoweri(cv->h / cv->tau[2], cv->L);
    for (int i = 0; i < cv->n; ++i) cv->tempv[i] = -cquot * cv->zn[Q_MAX][i] + cv->acor[i];
    double dup = wrms(cv, cv->tempv) * cv->tq[3];
    cv->etaqp1 = 1.0 / (rpowerr(BIAS3 * dup, 1.0 / (cv->L + 1)) + ADDON);
  }
  /* cvChooseEta */
  double etam = fmax(cv->etaqm1, fmax(cv->etaq, cv->etaqp1));
  if (etam < THRESH) {
    cv->eta = 1.0;
    cv->qprime = cv->q;
  } else if (etam == cv->etaq) {
    cv->eta = cv->etaq;
    cv->qprime = cv->q;
  } else if (etam == cv->etaqm1) {
    cv->eta = cv->etaqm1;
    cv->qprime = cv->q - 1;
  } else {
    cv->eta = cv->etaqp1;
    cv->qprime = cv->q + 1;
    memcpy(cv->zn[Q_MAX], cv->acor, cv->n * sizeof(double));
  }
  cv_set_eta(cv);
}

static int cv_step(Cv *cv) {
  double saved_t = cv->tn, dsm = 0.0;
  int ncf = 0, nef = 0, nflag = FIRST_CALL, kflag, eflag;
  if ((cv->nst > 0) && (cv->hprime != cv->h)) cv_adjust_params(cv);
  for (;;) {
    cv_predict(cv);
    cv_set(cv);
    nflag = cv_nls(cv, nflag);
    kflag = cv_handle_nflag(cv, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN) continue;
    if (kflag != DO_ERROR_TEST) return kflag;
    eflag = cv_do_error_test(cv, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != CV_SUCCESS) return eflag;
    break;
  }
  cv_complete_step(cv);
  cv_prepare_next_step(cv, dsm);
  cv->etamax = (cv->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  for (int i = 0; i < cv->n; ++i) cv->acor[i] *= cv->tq[2];
  return CV_SUCCESS;
}

/* CVodeReInit + CVodeSetInitStep + CVode(tout, CV_NORMAL).  y in/out. */
static int cv_solve(Cv *cv, double t0, double tout, double *y) {
  int n = cv->n;
  memcpy(cv->zn[0], y, n * sizeof(double));
  memcpy(cv->last_yn, y, n * sizeof(double));
  cv->tn = t0;
  cv->q = 1;
  cv->L = 2;
  cv->qwait = cv->L;
  cv->etamax = ETAMX1;
  cv->qu = 0;
  cv->hu = 0.0;
  cv->nst = cv->nfe = cv->ncfn = cv->netf = cv->nni = cv->nsetups = cv->nstlp = 0;
  cv->nje = cv->nstlj = 0;
  cv->nscon = 0;
  cv->indx_acor = Q_MAX;
  cv->saved_tq5 = 0.0;
  cv->crate = 1.0;
  cv->gamrat = 1.0;
  cv->jcur = 0;
  for (int i = 0; i < L_MAX + 2; ++i) cv->tau[i] = 0.0;
  for (int i = 0; i < 6; ++i) cv->tq[i] = 0.0;
  cv->init_h = tout - t0;
  cv->next_h = 0.0;
  if (ewt_set(cv, cv->zn[0])) return CV_ILL_INPUT;
  if (cv_f(cv, cv->zn[0], cv->zn[1]) > 0) return CV_FIRST_RHSFUNC_ERR;
  cv->h = cv->init_h;
  for (int i = 0; i < n; ++i) cv->zn[1][i] *= cv->h;
  cv->hscale = cv->h;
  cv->hprime = cv->h;
  long nstloc = 0;
  for (;;) {
    cv->next_h = cv->h;
    cv->next_q = cv->q;
    if (cv->nst > 0 && ewt_set(cv, cv->zn[0])) return CV_ILL_INPUT;
    if (nstloc >= cv->o->max_steps) return CV_TOO_MUCH_WORK;
    double nrm = wrms(cv, cv->zn[0]);
    if (DBL_EPSILON * nrm > 1.0) return CV_TOO_MUCH_ACC;
    int kflag = cv_step(cv);
    if (kflag != CV_SUCCESS) return kflag;
    nstloc++;
    if ((cv->tn - tout) * cv->h >= 0.0) {
      /* CVodeGetDky(tout, 0) */
      double s = (tout - cv->tn) / cv->h;
      for (int i = 0; i < n; ++i) y[i] = cv->zn[cv->q][i];
      for (int j = cv->q - 1; j >= 0; --j)
        for (int i = 0; i < n; ++i) y[i] = cv->zn[j][i] + s * y[i];
      cv->next_q = cv->qprime;
      cv->next_h = cv->hprime;
      return CV_SUCCESS;
    }
  }
}

static Cv *cv_alloc(Oracle *o, int nc) {
  Cv *cv = xcalloc(1, sizeof(Cv));
  cv->o = o;
  cv->nc = nc;
  cv->n = nc * o->nv;
  size_t n = cv->n ? cv->n : 1;
  for (int j = 0; j < L_MAX; ++j) cv->zn[j] = xcalloc(n, sizeof(double));
  cv->ewt = xcalloc(n, sizeof(double));
  cv->acor = xcalloc(n, sizeof(double));
  cv->y = xcalloc(n, sizeof(double));
  cv->tempv = xcalloc(n, sizeof(double));
  cv->ftemp = xcalloc(n, sizeof(double));
  cv->last_yn = xcalloc(n, sizeof(double));
  cv->gh_tmp = xcalloc(n, sizeof(double));
  cv->gh_corr = xcalloc(n, sizeof(double));
  cv->acor_init = xcalloc(n, sizeof(double));
  cv->Jw = xcalloc((size_t)nc * (o->nnz ? o->nnz : 1), sizeof(double));
  cv->lu = xcalloc((size_t)nc * (o->lu_nnz ? o->lu_nnz : 1), sizeof(double));
  return cv;
}

static void cv_free(Cv *cv) {
  for (int j = 0; j < L_MAX; ++j) free(cv->zn[j]);
  free(cv->ewt); free(cv->acor); free(cv->y); free(cv->tempv); free(cv->ftemp);
  free(cv->last_yn); free(cv->gh_tmp); free(cv->gh_corr); free(cv->acor_init);
  free(cv->Jw); free(cv->lu);
  free(cv);
}

static void stats_add(Oracle *o, const Cv *cv, int flag) {
  o->st[0] += cv->nst; o->st[1] += cv->nfe; o->st[2] += cv->nsetups;
  o->st[3] += cv->netf; o->st[4] += cv->nni; o->st[5] += cv->ncfn;
  o->st[6] += cv->nje;
  if (flag < 0) o->st[7] = flag;
  o->h_last = cv->hu;
  o->h_next = cv->next_h;
}

/* Run CVODE on cells [c0, c0+nc) as one problem (solver_run body after the
 * per-cell rate update). Returns the CVODE flag. */
static int run_problem(Oracle *o, Cv *cv, int c0, int nc, double t0, double t1) {
  int nv = o->nv, ns = o->ns;
  cv->c0 = c0;
  double *y = xmalloc((size_t)(nc * nv > 0 ? nc * nv : 1) * sizeof(double));
  for (int c = 0; c < nc; ++c)
    for (int d = 0; d < nv; ++d) {
      double v = o->state[(size_t)(c0 + c) * ns + o->dep_state[d]];
      y[c * nv + d] = v > TINY ? v : TINY;
    }
  int flag = CV_SUCCESS;
  if (o->n_rxn > 0) {
    /* is_anything_going_on_here: one f() whose result only matters when it
     * fails, which it cannot on y >= TINY. */
    cv->init_h = t1 - t0;
    cv->next_h = 0.0;
    double *tmp = xmalloc((size_t)(nc * nv > 0 ? nc * nv : 1) * sizeof(double));
    range_f(o, c0, nc, y, tmp, cv->init_h, 1);
    free(tmp);
    flag = cv_solve(cv, t0, t1, y);
  }
  stats_add(o, cv, flag);
  if (flag >= 0) {
    for (int c = 0; c < nc; ++c)
      for (int d = 0; d < nv; ++d)
        o->state[(size_t)(c0 + c) * ns + o->dep_state[d]] =
            y[c * nv + d] > 0.0 ? y[c * nv + d] : 0.0;
  }
  free(y);
  return flag;
}

static void clamp_constants(Oracle *o, double *state, int c0, int c1) {
  for (int c = c0; c < c1; ++c)
    for (int i = 0; i < o->ns; ++i)
      if (o->var_type[i] != VARIABLE) {
        double *v = &state[(size_t)c * o->ns + i];
        *v = *v > TINY ? *v : TINY;
      }
}

int oracle_solver_run(void *om, double *state, const double *env,
                      double t_initial, double t_final, int mode) {
  Oracle *o = om;
  for (int i = 0; i < 8; ++i) o->st[i] = 0;
  clamp_constants(o, state, 0, o->nc);
  o->state = state;
  o->env = env;
  for (int c = 0; c < o->nc; ++c) update_env_cell(o, c, &env[2 * c]);
  int fail = 0;
  if (mode == 0) {
    Cv *cv = cv_alloc(o, o->nc);
    fail = run_problem(o, cv, 0, o->nc, t_initial, t_final) < 0;
    cv_free(cv);
  } else {
    Cv *cv = cv_alloc(o, 1);
    for (int c = 0; c < o->nc; ++c)
      if (run_problem(o, cv, c, 1, t_initial, t_final) < 0) fail = 1;
    cv_free(cv);
  }
  o->state = NULL;
  return fail ? 1 : 0;
}

int oracle_solver_run_cells(void *om, double *state, const double *env,
                            double t_initial, double t_final, int c0, int c1) {
  Oracle *o = om;
  for (int i = 0; i < 8; ++i) o->st[i] = 0;
  clamp_constants(o, state, c0, c1);
  o->state = state;
  o->env = env;
  for (int c = c0; c < c1; ++c) update_env_cell(o, c, &env[2 * c]);
  int nfail = 0;
  Cv *cv = cv_alloc(o, 1);
  for (int c = c0; c < c1; ++c)
    if (run_problem(o, cv, c, 1, t_initial, t_final) < 0) nfail++;
  cv_free(cv);
  o->state = NULL;
  return nfail;
}

void oracle_get_stats(void *om, long *counters8, double *h2) {
  Oracle *o = om;
  for (int i = 0; i < 8; ++i) counters8[i] = o->st[i];
  h2[0] = o->h_last;
  h2[1] = o->h_next;
}

void oracle_set_guess_helper(void *om, int on) { ((Oracle *)om)->guess_helper_on = on; }
void oracle_set_neg_mode(void *om, int mode) { ((Oracle *)om)->neg_mode = mode; }

/* --------------------------------------------------------------------- */
/* Direct f/J evaluation for parity tests                                 */
/* --------------------------------------------------------------------- */
int oracle_eval_f(void *om, int cell, const double *state_cell, double *deriv,
                  double time_step, int use_deriv_est) {
  Oracle *o = om;
  double *st = xmalloc((size_t)o->ns * sizeof(double));
  double *y = xmalloc((size_t)(o->nv ? o->nv : 1) * sizeof(double));
  memcpy(st, state_cell, o->ns * sizeof(double));
  for (int d = 0; d < o->nv; ++d) y[d] = state_cell[o->dep_state[d]];
  int rc = update_cell_state(o, st, y);
  if (!rc) {
    if (use_deriv_est) cell_est(o, cell, y, o->est);
    cell_deriv(o, cell, st, deriv, use_deriv_est ? o->est : NULL, time_step);
  }
  free(st);
  free(y);
  return rc;
}

int oracle_eval_jac(void *om, int cell, const double *state_cell, double *jac,
                    double time_step) {
  Oracle *o = om;
  double *st = xmalloc((size_t)o->ns * sizeof(double));
  double *y = xmalloc((size_t)(o->nv ? o->nv : 1) * sizeof(double));
  double *fy = xmalloc((size_t)(o->nv ? o->nv : 1) * sizeof(double));
  memcpy(st, state_cell, o->ns * sizeof(double));
  for (int d = 0; d < o->nv; ++d) y[d] = state_cell[o->dep_state[d]];
  int rc = update_cell_state(o, st, y);
  if (!rc) {
    cell_deriv(o, cell, st, fy, NULL, time_step);
    cell_jac(o, cell, st, jac, time_step);
    memcpy(&o->J_solver[(size_t)cell * o->nnz], jac, o->nnz * sizeof(double));
    memcpy(&o->J_state[(size_t)cell * o->nv], y, o->nv * sizeof(double));
    memcpy(&o->J_deriv[(size_t)cell * o->nv], fy, o->nv * sizeof(double));
  }
  free(st);
  free(y);
  free(fy);
  return rc;
}

void oracle_reset_jac_history(void *om) {
  Oracle *o = om;
  memset(o->J_solver, 0, (size_t)o->nc * o->nnz * sizeof(double));
  memset(o->J_state, 0, (size_t)o->nc * o->nv * sizeof(double));
  memset(o->J_deriv, 0, (size_t)o->nc * o->nv * sizeof(double));
}

/* --------------------------------------------------------------------- */
/* Getters and LU checks                                                  */
/* --------------------------------------------------------------------- */
int oracle_n_dep(void *om) { return ((Oracle *)om)->nv; }
int oracle_nnz(void *om) { return ((Oracle *)om)->nnz; }
int oracle_nnz_rxn(void *om) { return ((Oracle *)om)->rjac.nnz; }
int oracle_n_mapped(void *om) { return ((Oracle *)om)->n_mapped; }
int oracle_rxn_int_size(void *om) { return ((Oracle *)om)->n_int_total; }
int oracle_n_env(void *om) { return ((Oracle *)om)->n_env; }
int oracle_lu_nnz(void *om) { return ((Oracle *)om)->lu_nnz; }

void oracle_get_pattern(void *om, int *rxn_colptr, int *rxn_rowval, int *colptr,
                        int *rowval, int *deriv_ids) {
  Oracle *o = om;
  memcpy(rxn_colptr, o->rjac.colptr, (o->ns + 1) * sizeof(int));
  memcpy(rxn_rowval, o->rjac.rowval, o->rjac.nnz * sizeof(int));
  memcpy(colptr, o->colptr, (o->nv + 1) * sizeof(int));
  memcpy(rowval, o->rowval, o->nnz * sizeof(int));
  memcpy(deriv_ids, o->deriv_ids, o->ns * sizeof(int));
}

void oracle_get_jac_map(void *om, int *solver_id, int *rxn_id, int *param_id) {
  Oracle *o = om;
  memcpy(solver_id, o->map_solver, o->n_mapped * sizeof(int));
  memcpy(rxn_id, o->map_rxn, o->n_mapped * sizeof(int));
  memcpy(param_id, o->map_param, o->n_mapped * sizeof(int));
}

void oracle_get_rxn_int_data(void *om, int *dst, int n) {
  Oracle *o = om;
  memcpy(dst, o->rxn_int, (n < o->n_int_total ? n : o->n_int_total) * sizeof(int));
}

void oracle_get_rxn_env(void *om, double *dst) {
  Oracle *o = om;
  memcpy(dst, o->rxn_env, (size_t)o->nc * o->n_env * sizeof(double));
}

int oracle_lu_solve_sparse(void *om, const double *jac, double gamma,
                           const double *b, double *x) {
  Oracle *o = om;
  double *lu = xmalloc((o->lu_nnz ? o->lu_nnz : 1) * sizeof(double));
  int rc = lu_factor(o, jac, gamma, lu);
  memcpy(x, b, o->nv * sizeof(double));
  if (!rc) lu_solve(o, lu, x);
  free(lu);
  return rc;
}

int oracle_lu_solve_dense(void *om, const double *jac, double gamma,
                          const double *b, double *x) {
  Oracle *o = om;
  int n = o->nv;
  double *A = xcalloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j)
    for (int k = o->colptr[j]; k < o->colptr[j + 1]; ++k)
      A[o->rowval[k] * n + j] = -gamma * jac[k];
  for (int i = 0; i < n; ++i) A[i * n + i] += 1.0;
  memcpy(x, b, n * sizeof(double));
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
    if (A[p * n + k] == 0.0) {
      free(A);
      return 1;
    }
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      double t = x[k];
      x[k] = x[p];
      x[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      double l = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
      x[i] -= l * x[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) x[i] -= A[i * n + j] * x[j];
    x[i] /= A[i * n + i];
  }
  free(A);
  return 0;
}

/* --------------------------------------------------------------------- */
/* Known-answer hooks for the pattern builder and the accumulators        */
/* --------------------------------------------------------------------- */
int oracle_kat_pattern(int n, int n_pairs, const int *dep, const int *ind,
                       int *colptr, int *rowval) {
  ColElems c;
  cols_init(&c, n);
  for (int k = 0; k < n_pairs; ++k) cols_register(&c, dep[k], ind[k]);
  Csc m = cols_build(&c);
  memcpy(colptr, m.colptr, (n + 1) * sizeof(int));
  memcpy(rowval, m.rowval, m.nnz * sizeof(int));
  int nnz = m.nnz;
  free(m.colptr);
  free(m.rowval);
  return nnz;
}

/* jacobian_add_value / jacobian_output (Jacobian.c:223-241): flag 0 adds to
 * the production sum, 1 to the loss sum; out = (double)(prod - loss). */
void oracle_kat_jac_accum(int nnz, int n_adds, const int *elem, const int *flag,
                          const double *val, double *out) {
  long double *p = xcalloc(nnz, sizeof(long double));
  long double *l = xcalloc(nnz, sizeof(long double));
  for (int k = 0; k < n_adds; ++k) {
    if (flag[k] == 0) p[elem[k]] += val[k];
    if (flag[k] == 1) l[elem[k]] += val[k];
  }
  for (int e = 0; e < nnz; ++e) out[e] = p[e] - l[e];
  free(p);
  free(l);
}

/* time_derivative_add_value / time_derivative_output with the oracle's
 * accumulators (time_derivative.c:46-97). */
void oracle_kat_time_derivative(int n_spec, int n_contrib, const int *spec,
                                const double *contrib, const double *est,
                                double *out) {
  Oracle o;
  memset(&o, 0, sizeof o);
  o.nv = n_spec;
  o.td_p = xcalloc(n_spec, sizeof(long double));
  o.td_l = xcalloc(n_spec, sizeof(long double));
  for (int k = 0; k < n_contrib; ++k) td_add(&o, spec[k], (long double)contrib[k]);
  for (int i = 0; i < n_spec; ++i) {
    long double rp = o.td_p[i], rl = o.td_l[i];
    if (rp + rl != 0.0) {
      if (est) {
        long double sf = 1.0 / (rp + rl) /
                         (1.0 / (rp + rl) + MAX_PRECISION_LOSS / fabsl(rp - rl));
        out[i] = sf * (rp - rl) + (1.0 - sf) * est[i];
      } else {
        out[i] = rp - rl;
      }
    } else {
      out[i] = 0.0;
    }
  }
  free(o.td_p);
  free(o.td_l);
}
