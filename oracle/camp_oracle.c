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
#define RXN_CMAQ_H2O2 3
#define RXN_CMAQ_OH_HNO3 4
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
  /* aerosol phases, representations and sub models (camp_solver.c:200-349) */
  int n_phase, n_added_phase, *ph_int, *ph_int_idx, *ph_flt_idx;
  double *ph_flt;
  int n_rep, n_added_rep, *rep_int, *rep_int_idx, *rep_flt_idx, *rep_env_idx, n_rep_env;
  double *rep_flt, *rep_env;
  int n_sub, n_added_sub, *sub_int, *sub_int_idx, *sub_flt_idx, *sub_env_idx, n_sub_env;
  double *sub_flt, *sub_env;
  double *J_params; /* [pjac.nnz], one cell */
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
 * rxn_CMAQ_H2O2.c:19-35, rxn_CMAQ_OH_HNO3.c:19-41,
 * rxn_photolysis.c:20-33). */
static int nprop_int(int type) { return type == RXN_PHOTOLYSIS ? 3 : 2; }
static int nprop_flt(int type) {
  switch (type) {
    case RXN_ARRHENIUS: return 6;
    case RXN_TROE: return 10;
    case RXN_CMAQ_H2O2: return 7;
    case RXN_CMAQ_OH_HNO3: return 11;
    default: return 1;
  }
}
#define NR(d) ((d)[0])
#define NP(d) ((d)[1])
#define REACT(d, t, x) ((d)[nprop_int(t) + (x)] - 1)
#define PROD(d, t, x) ((d)[nprop_int(t) + NR(d) + (x)] - 1)
#define DERIV_ID(d, t, x) ((d)[nprop_int(t) + NR(d) + NP(d) + (x)])
#define JAC_ID(d, t, x) ((d)[nprop_int(t) + 2 * (NR(d) + NP(d)) + (x)])
#define YIELD(f, t, x) ((f)[nprop_flt(t) + (x)])

static int is_mass_action(int type) {
  return type == RXN_ARRHENIUS || type == RXN_TROE || type == RXN_PHOTOLYSIS ||
         type == RXN_CMAQ_H2O2 || type == RXN_CMAQ_OH_HNO3;
}

static void td_add(Oracle *o, int i, long double c);
/* ===================================================================== */
/* Multi-phase chemistry: aerosol phases (aero_phase_solver.c), the       */
/* modal/binned-mass representation (aero_reps/aero_rep_modal_binned_     */
/* mass.c), SIMPOL phase transfer (rxns/rxn_SIMPOL_phase_transfer.c),     */
/* aqueous equilibrium (rxns/rxn_aqueous_equilibrium.c) and the ZSR       */
/* aerosol-water sub model (sub_models/sub_model_ZSR_aerosol_water.c).    */
/* Like the reference, these keep per-cell intermediate values in their   */
/* shared float arrays, so every use is preceded by the cell's update.    */
/* ===================================================================== */
#define RXN_AQUEOUS_EQUILIBRIUM 7
#define RXN_SIMPOL_PHASE_TRANSFER 10
#define AERO_REP_MODAL_BINNED_MASS 2
#define SUB_MODEL_ZSR_AEROSOL_WATER 2
#define CONSTANT 2
#define PSSA 3
#define UNIV_GAS_CONST 8.314472               /* util.h:17 */
#define PI_ 3.14159265358979323846
#define MINIMUM_MASS 1.0e-25L                 /* aero_phase_solver.c:19-24 */
#define MINIMUM_MW 0.1L
#define MINIMUM_DENSITY 1800.0L

#define PH_INT(o, p) (&(o)->ph_int[(o)->ph_int_idx[p]])
#define PH_FLT(o, p) (&(o)->ph_flt[(o)->ph_flt_idx[p]])

static int phase_spec_counts(int t) { return t == VARIABLE || t == CONSTANT || t == PSSA; }

/* aero_phase_get_used_jac_elem, aero_phase_solver.c:40-63 */
static int phase_used_jac_elem(Oracle *o, int p, int state_var_id, char *flags) {
  int *id = PH_INT(o, p);
  int n = 0;
  for (int i = 0; i < id[0]; ++i)
    if (phase_spec_counts(id[1 + i])) {
      flags[state_var_id + i] = 1;
      ++n;
    }
  return n;
}

/* aero_phase_get_mass__kg_m3, aero_phase_solver.c:78-113 */
static void phase_get_mass(Oracle *o, int p, const double *state, double *mass, double *mw,
                           double *jm, double *jmw) {
  int *id = PH_INT(o, p);
  double *fd = PH_FLT(o, p);
  long double l_mass = MINIMUM_MASS;
  long double moles = MINIMUM_MASS / MINIMUM_MW;
  int ij = 0;
  for (int i = 0; i < id[0]; ++i) {
    if (!phase_spec_counts(id[1 + i])) continue;
    l_mass += state[i];
    moles += state[i] / (long double)fd[i];
    if (jm) jm[ij] = 1.0L;
    if (jmw) jmw[ij] = 1.0L / fd[i];
    ij++;
  }
  *mw = (double)l_mass / moles;
  if (jmw)
    for (int j = 0; j < ij; ++j) jmw[j] = (moles - jmw[j] * l_mass) / (moles * moles);
  *mass = (double)l_mass;
}

/* aero_phase_get_volume__m3_m3, aero_phase_solver.c:128-150 */
static void phase_get_volume(Oracle *o, int p, const double *state, double *volume,
                             double *je) {
  int *id = PH_INT(o, p);
  double *fd = PH_FLT(o, p);
  int n = id[0];
  *volume = MINIMUM_MASS / MINIMUM_DENSITY;
  int ij = 0;
  for (int i = 0; i < n; ++i) {
    if (!phase_spec_counts(id[1 + i])) continue;
    *volume += state[i] / fd[n + i];
    if (je) je[ij++] = 1.0 / fd[n + i];
  }
}

/* modal/binned-mass layout, aero_rep_modal_binned_mass.c:27-81 (d: int data
 * after the type slot, f: float data, e: the cell's env data) */
#define MB_BINNED 1
#define MB_MODAL 2
#define MB_NSEC(d) ((d)[0])
#define MB_ID(d) ((d)[3])
#define MB_ILOC(d, x) ((d)[4 + (x)] - 1)
#define MB_FLOC(d, x) ((d)[4 + MB_NSEC(d) + (x)] - 1)
#define MB_TYPE(d, x) ((d)[MB_ILOC(d, x)])
#define MB_NBIN(d, x) ((d)[MB_ILOC(d, x) + 1])
#define MB_NPH(d, x) ((d)[MB_ILOC(d, x) + 2])
#define MB_PSTATE(d, x, y, b) ((d)[MB_ILOC(d, x) + 3 + (b) * MB_NPH(d, x) + (y)] - 1)
#define MB_PMODEL(d, x, y, b) \
  ((d)[MB_ILOC(d, x) + 3 + MB_NBIN(d, x) * MB_NPH(d, x) + (b) * MB_NPH(d, x) + (y)] - 1)
#define MB_PNJAC(d, x, y, b) \
  ((d)[MB_ILOC(d, x) + 3 + 2 * MB_NBIN(d, x) * MB_NPH(d, x) + (b) * MB_NPH(d, x) + (y)])
#define MB_DP(d, f, x, b) ((f)[MB_FLOC(d, x) + (b) * 3])
#define MB_GMD(d, e, x) ((e)[x])
#define MB_GSD(d, e, x) ((e)[MB_NSEC(d) + (x)])
#define MB_NCONC(d, f, x, b) ((f)[MB_FLOC(d, x) + (b) * 3 + 1])
#define MB_RAD(d, f, x, b) ((f)[MB_FLOC(d, x) + (b) * 3 + 2])
#define MB_PMASS(d, f, x, y, b) ((f)[MB_FLOC(d, x) + 3 * MB_NBIN(d, x) + (b) * MB_NPH(d, x) + (y)])
#define MB_PMW(d, f, x, y, b) \
  ((f)[MB_FLOC(d, x) + (3 + MB_NPH(d, x)) * MB_NBIN(d, x) + (b) * MB_NPH(d, x) + (y)])

typedef struct {
  int *d;
  double *f, *e;
} RepData;

static RepData rep_data(Oracle *o, int r, int cell) {
  RepData x;
  int *id = &o->rep_int[o->rep_int_idx[r]];
  if (id[0] != AERO_REP_MODAL_BINNED_MASS) {
    fprintf(stderr, "oracle: unsupported aerosol representation type %d\n", id[0]);
    exit(EXIT_FAILURE);
  }
  x.d = id + 1;
  x.f = &o->rep_flt[o->rep_flt_idx[r]];
  x.e = &o->rep_env[(size_t)cell * o->n_rep_env + o->rep_env_idx[r]];
  return x;
}

/* aero_rep_modal_binned_mass_get_used_jac_elem, :84-113 */
static int rep_used_jac_elem(Oracle *o, int r, int aero_phase_idx, char *flags) {
  RepData x = rep_data(o, r, 0);
  int *d = x.d;
  int n = 0;
  for (int s = 0; s < MB_NSEC(d) && aero_phase_idx >= 0; s++)
    for (int b = 0; b < MB_NBIN(d, s) && aero_phase_idx >= 0; b++)
      for (int p = 0; p < MB_NPH(d, s) && aero_phase_idx >= 0; p++) {
        if (aero_phase_idx == 0)
          for (int j = 0; j < MB_NPH(d, s); j++) {
            MB_PNJAC(d, s, j, b) =
                phase_used_jac_elem(o, MB_PMODEL(d, s, j, b), MB_PSTATE(d, s, j, b), flags);
            n += MB_PNJAC(d, s, j, b);
          }
        aero_phase_idx -= 1;
      }
  return n;
}

/* aero_rep_modal_binned_mass_update_state, :183-266 */
static void rep_update_state(Oracle *o, int cell, const double *state) {
  for (int r = 0; r < o->n_rep; ++r) {
    RepData x = rep_data(o, r, cell);
    int *d = x.d;
    double *f = x.f, *e = x.e;
    for (int s = 0; s < MB_NSEC(d); s++) {
      double volume, mass;
      switch (MB_TYPE(d, s)) {
        case MB_MODAL:
          volume = 0.0;
          for (int p = 0; p < MB_NPH(d, s); p++) {
            const double *st = state + MB_PSTATE(d, s, p, 0);
            phase_get_mass(o, MB_PMODEL(d, s, p, 0), st, &MB_PMASS(d, f, s, p, 0),
                           &MB_PMW(d, f, s, p, 0), NULL, NULL);
            double pv = 0.0;
            phase_get_volume(o, MB_PMODEL(d, s, p, 0), st, &pv, NULL);
            volume += pv;
          }
          MB_NCONC(d, f, s, 0) = volume * 6.0 /
                                 (PI_ * pow(MB_GMD(d, e, s), 3) *
                                  exp(9.0 / 2.0 * pow(log(MB_GSD(d, e, s)), 2)));
          break;
        case MB_BINNED:
          for (int b = 0; b < MB_NBIN(d, s); b++) {
            volume = 0.0;
            for (int p = 0; p < MB_NPH(d, s); p++) {
              const double *st = state + MB_PSTATE(d, s, p, b);
              phase_get_mass(o, MB_PMODEL(d, s, p, b), st, &MB_PMASS(d, f, s, p, b),
                             &MB_PMW(d, f, s, p, b), NULL, NULL);
              double pv = 0.0;
              phase_get_volume(o, MB_PMODEL(d, s, p, b), st, &pv, NULL);
              volume += pv;
            }
            MB_NCONC(d, f, s, b) =
                volume * 3.0 / (4.0 * PI_) / pow(MB_DP(d, f, s, b) / 2.0, 3);
          }
          break;
      }
      (void)mass;
    }
  }
}

/* aero_rep_modal_binned_mass_get_effective_radius__m, :282-309 */
static void rep_eff_radius(Oracle *o, int r, int cell, int aero_phase_idx, double *radius,
                           double *partial) {
  RepData x = rep_data(o, r, cell);
  int *d = x.d;
  double *f = x.f;
  for (int s = 0; s < MB_NSEC(d); s++)
    for (int b = 0; b < MB_NBIN(d, s); b++) {
      aero_phase_idx -= MB_NPH(d, s);
      if (aero_phase_idx < 0) {
        *radius = MB_RAD(d, f, s, b);
        if (partial)
          for (int p = 0; p < MB_NPH(d, s); ++p)
            for (int k = 0; k < MB_PNJAC(d, s, p, b); ++k) *(partial++) = 0.0;
        s = MB_NSEC(d);
        break;
      }
    }
}

/* aero_rep_modal_binned_mass_get_number_conc__n_m3, :325-372 */
static void rep_number_conc(Oracle *o, int r, int cell, const double *state,
                            int aero_phase_idx, double *nconc, double *partial) {
  RepData x = rep_data(o, r, cell);
  int *d = x.d;
  double *f = x.f, *e = x.e;
  for (int s = 0; s < MB_NSEC(d) && aero_phase_idx >= 0; s++)
    for (int b = 0; b < MB_NBIN(d, s) && aero_phase_idx >= 0; b++) {
      aero_phase_idx -= MB_NPH(d, s);
      if (aero_phase_idx < 0) {
        *nconc = MB_NCONC(d, f, s, b);
        if (partial)
          for (int p = 0; p < MB_NPH(d, s); ++p) {
            const double *st = state + MB_PSTATE(d, s, p, b);
            double pv = 0.0;
            phase_get_volume(o, MB_PMODEL(d, s, p, b), st, &pv, partial);
            for (int k = 0; k < MB_PNJAC(d, s, p, b); ++k) {
              switch (MB_TYPE(d, s)) {
                case MB_MODAL:
                  *(partial++) *= 6.0 / (PI_ * pow(MB_GMD(d, e, s), 3) *
                                         exp(9.0 / 2.0 * pow(log(MB_GSD(d, e, s)), 2)));
                  break;
                case MB_BINNED:
                  *(partial++) *= 3.0 / (4.0 * PI_) / pow(MB_DP(d, f, s, b) / 2.0, 3);
                  break;
              }
            }
          }
        s = MB_NSEC(d);
        break;
      }
    }
}

/* get_aero_phase_mass (:402-446) and get_aero_phase_avg_MW (:459-503):
 * which = 0 -> mass, 1 -> average MW */
static void rep_phase_prop(Oracle *o, int r, int cell, const double *state, int aero_phase_idx,
                           int which, double *value, double *partial) {
  RepData x = rep_data(o, r, cell);
  int *d = x.d;
  double *f = x.f;
  for (int s = 0; s < MB_NSEC(d) && aero_phase_idx >= 0; ++s)
    for (int b = 0; b < MB_NBIN(d, s) && aero_phase_idx >= 0; ++b) {
      if (aero_phase_idx < 0 || aero_phase_idx >= MB_NPH(d, s)) {
        aero_phase_idx -= MB_NPH(d, s);
        continue;
      }
      for (int p = 0; p < MB_NPH(d, s); ++p) {
        if (aero_phase_idx == 0) {
          *value = which == 0 ? MB_PMASS(d, f, s, p, b) : MB_PMW(d, f, s, p, b);
          if (partial) {
            const double *st = state + MB_PSTATE(d, s, p, b);
            double mass, mw;
            phase_get_mass(o, MB_PMODEL(d, s, p, b), st, &mass, &mw,
                           which == 0 ? partial : NULL, which == 1 ? partial : NULL);
            partial += MB_PNJAC(d, s, p, b);
          }
        } else if (partial) {
          for (int k = 0; k < MB_PNJAC(d, s, p, b); ++k) *(partial++) = 0.0;
        }
        aero_phase_idx -= 1;
      }
    }
}

/* util.h:19-68 */
static double mean_speed(double T, double mw) {
  return sqrt(8.0 * UNIV_GAS_CONST * T / (PI_ * mw));
}
static double mean_free_path(double D, double T, double mw) {
  return 3.0 * D / mean_speed(T, mw);
}
static double transition_corr(double mfp, double r, double alpha) {
  double Kn = mfp / r;
  return (0.75 * alpha * (1.0 + Kn)) / (Kn * Kn + (1.0 + 0.283 * alpha) * Kn + 0.75 * alpha);
}
static double d_transition_corr_d_radius(double mfp, double r, double alpha) {
  double Kn = mfp / r;
  return (0.75 * alpha * mfp * (Kn * Kn + 2.0 * Kn + 1.0 + (0.283 - 0.75) * alpha)) /
         pow(r * (Kn * Kn + (1.0 + 0.283 * alpha) * Kn + 0.75 * alpha), 2);
}
static double transition_rate(double D, double mfp, double r, double alpha) {
  return 4.0 * PI_ * r * D * transition_corr(mfp, r, alpha);
}
static double d_transition_rate_d_radius(double D, double mfp, double r, double alpha) {
  return 4.0 * PI_ * D * (transition_corr(mfp, r, alpha) + r * d_transition_corr_d_radius(mfp, r, alpha));
}

/* SIMPOL layout, rxn_SIMPOL_phase_transfer.c:31-74 (d after the type slot) */
#define SP_N(d) ((d)[0])
#define SP_GAS(d) ((d)[1] - 1)
#define SP_AERO(d, x) ((d)[2 + (x)] - 1)
#define SP_ACT(d, x) ((d)[2 + SP_N(d) + (x)] - 1)
#define SP_PHASE(d, x) ((d)[2 + 2 * SP_N(d) + (x)] - 1)
#define SP_REP(d, x) ((d)[2 + 3 * SP_N(d) + (x)] - 1)
#define SP_DERIV(d, x) ((d)[2 + 4 * SP_N(d) + (x)])
#define SP_GAS_ACT_JAC(d, x) ((d)[3 + 5 * SP_N(d) + (x)])
#define SP_AERO_ACT_JAC(d, x) ((d)[3 + 6 * SP_N(d) + (x)])
#define SP_JAC(d, x) ((d)[3 + 7 * SP_N(d) + (x)])
#define SP_ILOC(d, x) ((d)[4 + 10 * SP_N(d) + (x)] - 1)
#define SP_FLOC(d, x) ((d)[4 + 11 * SP_N(d) + (x)] - 1)
#define SP_NJAC(d, x) ((d)[SP_ILOC(d, x)])
#define SP_PJAC(d, x, s, k) ((d)[SP_ILOC(d, x) + 1 + (s) * SP_NJAC(d, x) + (k)])
#define SP_RAD_J(d, f, x, k) ((f)[SP_FLOC(d, x) + (k)])
#define SP_NUM_J(d, f, x, k) ((f)[SP_FLOC(d, x) + SP_NJAC(d, x) + (k)])
#define SP_MASS_J(d, f, x, k) ((f)[SP_FLOC(d, x) + 2 * SP_NJAC(d, x) + (k)])
#define SP_MW_J(d, f, x, k) ((f)[SP_FLOC(d, x) + 3 * SP_NJAC(d, x) + (k)])
#define SP_JAC_GAS 0
#define SP_JAC_AERO 1
/* float: DELTA_H, DELTA_S, DIFF_COEFF, PRE_C_AVG, B1..B4, CONV, MW;
 * env: MFP_M, ALPHA, EQUIL_CONST, KGM3_TO_PPM */

/* rxn_SIMPOL_phase_transfer_get_used_jac_elem, :86-149 */
static void simpol_used_jac_elem(Oracle *o, int *d, ColElems *jac) {
  char *flags = xcalloc(o->ns, 1);
  cols_register(jac, SP_GAS(d), SP_GAS(d));
  for (int i = 0; i < SP_N(d); i++) {
    cols_register(jac, SP_AERO(d, i), SP_GAS(d));
    cols_register(jac, SP_GAS(d), SP_AERO(d, i));
    cols_register(jac, SP_AERO(d, i), SP_AERO(d, i));
    if (SP_ACT(d, i) > 0) {
      cols_register(jac, SP_GAS(d), SP_ACT(d, i));
      cols_register(jac, SP_AERO(d, i), SP_ACT(d, i));
    }
    memset(flags, 0, o->ns);
    int n_elem = rep_used_jac_elem(o, SP_REP(d, i), SP_PHASE(d, i), flags);
    if (n_elem > SP_NJAC(d, i)) {
      fprintf(stderr, "oracle: SIMPOL got more Jacobian elements than expected\n");
      exit(EXIT_FAILURE);
    }
    int used = 0;
    for (int k = 0; k < o->ns; ++k)
      if (flags[k]) {
        cols_register(jac, SP_GAS(d), k);
        cols_register(jac, SP_AERO(d, i), k);
        SP_PJAC(d, i, SP_JAC_GAS, used) = k;
        SP_PJAC(d, i, SP_JAC_AERO, used) = k;
        ++used;
      }
    for (; used < SP_NJAC(d, i); ++used) {
      SP_PJAC(d, i, SP_JAC_GAS, used) = -1;
      SP_PJAC(d, i, SP_JAC_AERO, used) = -1;
    }
  }
  free(flags);
}

/* rxn_SIMPOL_phase_transfer_update_ids, :160-207 */
static void simpol_update_ids(Oracle *o, int *d) {
  SP_DERIV(d, 0) = o->deriv_ids[SP_GAS(d)];
  for (int i = 0; i < SP_N(d); i++) SP_DERIV(d, i + 1) = o->deriv_ids[SP_AERO(d, i)];
  int j = 0;
  SP_JAC(d, j++) = csc_find(&o->rjac, SP_GAS(d), SP_GAS(d));
  for (int i = 0; i < SP_N(d); i++) {
    SP_JAC(d, j++) = csc_find(&o->rjac, SP_AERO(d, i), SP_GAS(d));
    SP_JAC(d, j++) = csc_find(&o->rjac, SP_GAS(d), SP_AERO(d, i));
    SP_JAC(d, j++) = csc_find(&o->rjac, SP_AERO(d, i), SP_AERO(d, i));
    if (SP_ACT(d, i) > 0) {
      SP_GAS_ACT_JAC(d, i) = csc_find(&o->rjac, SP_GAS(d), SP_ACT(d, i));
      SP_AERO_ACT_JAC(d, i) = csc_find(&o->rjac, SP_AERO(d, i), SP_ACT(d, i));
    } else {
      SP_GAS_ACT_JAC(d, i) = -1;
      SP_AERO_ACT_JAC(d, i) = -1;
    }
    for (int k = 0; k < SP_NJAC(d, i); ++k) {
      if (SP_PJAC(d, i, SP_JAC_GAS, k) > 0)
        SP_PJAC(d, i, SP_JAC_GAS, k) = csc_find(&o->rjac, SP_GAS(d), SP_PJAC(d, i, SP_JAC_GAS, k));
      if (SP_PJAC(d, i, SP_JAC_AERO, k) > 0)
        SP_PJAC(d, i, SP_JAC_AERO, k) =
            csc_find(&o->rjac, SP_AERO(d, i), SP_PJAC(d, i, SP_JAC_AERO, k));
    }
  }
}

/* rxn_SIMPOL_phase_transfer_update_env_state, :220-266 */
static void simpol_env(const double *f, double *e, double T, double P) {
  e[1] = 0.1;
  if (f[0] != 0.0 || f[1] != 0.0) {
    double del_G = f[0] - T * f[1];
    e[1] = exp(-del_G / (UNIV_GAS_CONST * T));
    e[1] = e[1] / (1.0 + e[1]);
  }
  e[0] = mean_free_path(f[2], T, f[9]);
  double vp = f[4] / T + f[5] + f[6] * T + f[7] * log(T);
  vp = 101325.0 * pow(10, vp);
  e[3] = f[8] * T / P;
  e[2] = vp / P / f[9] * 1.0e6;
}

/* rxn_SIMPOL_phase_transfer_calc_deriv_contrib, :279-402 */
static void simpol_deriv(Oracle *o, int cell, const double *state, int *d, const double *fl,
                         const double *e) {
  for (int i = 0; i < SP_N(d); i++) {
    double radius, number_conc, mass, avg_mw;
    rep_eff_radius(o, SP_REP(d, i), cell, SP_PHASE(d, i), &radius, NULL);
    int conc_type = 1; /* modal/binned mass: TOTAL_PARTICLE_MASS (:375-385) */
    rep_number_conc(o, SP_REP(d, i), cell, state, SP_PHASE(d, i), &number_conc, NULL);
    rep_phase_prop(o, SP_REP(d, i), cell, state, SP_PHASE(d, i), 0, &mass, NULL);
    rep_phase_prop(o, SP_REP(d, i), cell, state, SP_PHASE(d, i), 1, &avg_mw, NULL);
    long double cond = transition_rate(fl[2], e[0], radius, e[1]);
    long double evap = cond * (e[2] * avg_mw / mass);
    long double act = 1.0;
    if (SP_ACT(d, i) > -1) act = state[SP_ACT(d, i)];
    evap *= act;
    cond *= state[SP_GAS(d)];
    evap *= state[SP_AERO(d, i)];
    if (conc_type == 0) {
      if (SP_DERIV(d, 0) >= 0) {
        td_add(o, SP_DERIV(d, 0), number_conc * evap);
        td_add(o, SP_DERIV(d, 0), -number_conc * cond);
      }
      if (SP_DERIV(d, 1 + i) >= 0) {
        td_add(o, SP_DERIV(d, 1 + i), -evap / e[3]);
        td_add(o, SP_DERIV(d, 1 + i), cond / e[3]);
      }
    } else {
      if (SP_DERIV(d, 0) >= 0) {
        td_add(o, SP_DERIV(d, 0), number_conc * evap);
        td_add(o, SP_DERIV(d, 0), -number_conc * cond);
      }
      if (SP_DERIV(d, 1 + i) >= 0) {
        td_add(o, SP_DERIV(d, 1 + i), -number_conc * evap / e[3]);
        td_add(o, SP_DERIV(d, 1 + i), number_conc * cond / e[3]);
      }
    }
  }
}

static void jac_add(Oracle *o, int elem, int production, long double v) {
  if (production) o->jp[elem] += v;
  else o->jl[elem] += v;
}

/* rxn_SIMPOL_phase_transfer_calc_jac_contrib, :415-~740 */
static void simpol_jac(Oracle *o, int cell, const double *state, int *d, double *fl,
                       const double *e) {
  for (int i = 0; i < SP_N(d); i++) {
    double radius, number_conc, mass, avg_mw;
    rep_eff_radius(o, SP_REP(d, i), cell, SP_PHASE(d, i), &radius, &SP_RAD_J(d, fl, i, 0));
    int conc_type = 1;
    rep_number_conc(o, SP_REP(d, i), cell, state, SP_PHASE(d, i), &number_conc,
                    &SP_NUM_J(d, fl, i, 0));
    rep_phase_prop(o, SP_REP(d, i), cell, state, SP_PHASE(d, i), 0, &mass, &SP_MASS_J(d, fl, i, 0));
    rep_phase_prop(o, SP_REP(d, i), cell, state, SP_PHASE(d, i), 1, &avg_mw, &SP_MW_J(d, fl, i, 0));
    long double cond = transition_rate(fl[2], e[0], radius, e[1]);
    long double evap = cond * (e[2] * avg_mw / mass);
    long double act = 1.0;
    if (SP_ACT(d, i) > -1) act = state[SP_ACT(d, i)];
    const double kp = e[3];
    if (conc_type == 0) {
      if (SP_JAC(d, 1 + i * 3 + 1) >= 0)
        jac_add(o, SP_JAC(d, 1 + i * 3 + 1), 1, number_conc * evap * act);
      if (SP_JAC(d, 0) >= 0) jac_add(o, SP_JAC(d, 0), 0, number_conc * cond);
      if (SP_JAC(d, 1 + i * 3) >= 0) jac_add(o, SP_JAC(d, 1 + i * 3), 1, cond / kp);
      if (SP_JAC(d, 1 + i * 3 + 2) >= 0)
        jac_add(o, SP_JAC(d, 1 + i * 3 + 2), 0, evap * act / kp);
      if (SP_GAS_ACT_JAC(d, i) > 0)
        jac_add(o, SP_GAS_ACT_JAC(d, i), 1, number_conc * evap * state[SP_AERO(d, i)]);
      if (SP_AERO_ACT_JAC(d, i) > 0)
        jac_add(o, SP_AERO_ACT_JAC(d, i), 0, evap / kp * state[SP_AERO(d, i)]);
    } else {
      if (SP_JAC(d, 1 + i * 3 + 1) >= 0)
        jac_add(o, SP_JAC(d, 1 + i * 3 + 1), 1, number_conc * evap * act);
      if (SP_JAC(d, 0) >= 0) jac_add(o, SP_JAC(d, 0), 0, number_conc * cond);
      if (SP_JAC(d, 1 + i * 3) >= 0) jac_add(o, SP_JAC(d, 1 + i * 3), 1, number_conc * cond / kp);
      if (SP_JAC(d, 1 + i * 3 + 2) >= 0)
        jac_add(o, SP_JAC(d, 1 + i * 3 + 2), 0, number_conc * evap * act / kp);
      if (SP_GAS_ACT_JAC(d, i) > 0)
        jac_add(o, SP_GAS_ACT_JAC(d, i), 1, number_conc * evap * state[SP_AERO(d, i)]);
      if (SP_AERO_ACT_JAC(d, i) > 0)
        jac_add(o, SP_AERO_ACT_JAC(d, i), 0, number_conc * evap / kp * state[SP_AERO(d, i)]);
    }
    evap *= act;
    cond *= state[SP_GAS(d)];
    evap *= state[SP_AERO(d, i)];
    double d_cond_d_radius =
        d_transition_rate_d_radius(fl[2], e[0], radius, e[1]) * state[SP_GAS(d)];
    double d_evap_d_radius = d_cond_d_radius / state[SP_GAS(d)] * e[2] * avg_mw / mass *
                             state[SP_AERO(d, i)];
    double d_evap_d_mass = -evap / mass;
    double d_evap_d_mw = evap / avg_mw;
    for (int k = 0; k < SP_NJAC(d, i); ++k) {
      int g = SP_PJAC(d, i, SP_JAC_GAS, k), a = SP_PJAC(d, i, SP_JAC_AERO, k);
      double rj = SP_RAD_J(d, fl, i, k), nj = SP_NUM_J(d, fl, i, k);
      double mj = SP_MASS_J(d, fl, i, k), wj = SP_MW_J(d, fl, i, k);
      if (g > 0) {
        jac_add(o, g, 1, number_conc * d_evap_d_radius * rj);
        jac_add(o, g, 0, number_conc * d_cond_d_radius * rj);
        jac_add(o, g, 1, evap * nj);
        jac_add(o, g, 0, cond * nj);
        jac_add(o, g, 1, number_conc * d_evap_d_mass * mj);
        jac_add(o, g, 1, number_conc * d_evap_d_mw * wj);
      }
      if (a > 0) {
        if (conc_type == 0) {
          jac_add(o, a, 0, d_evap_d_radius / kp * rj);
          jac_add(o, a, 1, d_cond_d_radius / kp * rj);
          jac_add(o, a, 0, d_evap_d_mass / kp * mj);
          jac_add(o, a, 0, d_evap_d_mw / kp * wj);
        } else {
          jac_add(o, a, 0, number_conc * d_evap_d_radius / kp * rj);
          jac_add(o, a, 1, number_conc * d_cond_d_radius / kp * rj);
          jac_add(o, a, 0, evap / kp * nj);
          jac_add(o, a, 1, cond / kp * nj);
          jac_add(o, a, 0, number_conc * d_evap_d_mass / kp * mj);
          jac_add(o, a, 0, number_conc * d_evap_d_mw / kp * wj);
        }
      }
    }
  }
}

/* aqueous equilibrium layout, rxn_aqueous_equilibrium.c:27-68 */
#define AQ_NR(d) ((d)[0])
#define AQ_NP(d) ((d)[1])
#define AQ_NPH(d) ((d)[2])
#define AQ_REACT(d, x) ((d)[3 + (x)] - 1)
#define AQ_PROD(d, x) ((d)[3 + AQ_NR(d) * AQ_NPH(d) + (x)] - 1)
#define AQ_WATER(d, x) ((d)[3 + (AQ_NR(d) + AQ_NP(d)) * AQ_NPH(d) + (x)] - 1)
#define AQ_ACT(d, x) ((d)[3 + (AQ_NR(d) + AQ_NP(d) + 1) * AQ_NPH(d) + (x)] - 1)
#define AQ_DERIV(d, x) ((d)[3 + (AQ_NR(d) + AQ_NP(d) + 2) * AQ_NPH(d) + (x)])
#define AQ_JAC(d, x) ((d)[3 + (2 * (AQ_NR(d) + AQ_NP(d)) + 2) * AQ_NPH(d) + (x)])
#define AQ_A(f) ((f)[0])
#define AQ_C(f) ((f)[1])
#define AQ_KREV(f) ((f)[2])
#define AQ_WATER_CONC(f) ((f)[3])
#define AQ_ACT_VAL(f) ((f)[4])
#define AQ_MFTM(d, f, x) ((f)[5 + (x)])
#define AQ_REACT_CONC(d, f, x) ((f)[5 + AQ_NR(d) + AQ_NP(d) + (x)])
#define AQ_PROD_CONC(d, f, x) ((f)[5 + 2 * AQ_NR(d) + AQ_NP(d) + (x)])
#define AQ_SMALL_WATER(d, f, x) ((f)[5 + 2 * AQ_NR(d) + 2 * AQ_NP(d) + (x)])
#define AQ_SMALL_CONC(d, f, x) ((f)[5 + 2 * AQ_NR(d) + 2 * AQ_NP(d) + AQ_NPH(d) + (x)])
#define MIN_WATER 1.0e-4

/* rxn_aqueous_equilibrium_get_used_jac_elem, :80-126 */
static void aqeq_used_jac_elem(int *d, ColElems *jac) {
  int nr = AQ_NR(d), np = AQ_NP(d);
  for (int p = 0; p < AQ_NPH(d); p++) {
    for (int ri = p * nr; ri < (p + 1) * nr; ri++) {
      for (int rd = p * nr; rd < (p + 1) * nr; rd++) cols_register(jac, AQ_REACT(d, rd), AQ_REACT(d, ri));
      for (int pd = p * np; pd < (p + 1) * np; pd++) cols_register(jac, AQ_PROD(d, pd), AQ_REACT(d, ri));
    }
    for (int pi = p * np; pi < (p + 1) * np; pi++) {
      for (int rd = p * nr; rd < (p + 1) * nr; rd++) cols_register(jac, AQ_REACT(d, rd), AQ_PROD(d, pi));
      for (int pd = p * np; pd < (p + 1) * np; pd++) cols_register(jac, AQ_PROD(d, pd), AQ_PROD(d, pi));
    }
    for (int rd = p * nr; rd < (p + 1) * nr; rd++) cols_register(jac, AQ_REACT(d, rd), AQ_WATER(d, p));
    for (int pd = p * np; pd < (p + 1) * np; pd++) cols_register(jac, AQ_PROD(d, pd), AQ_WATER(d, p));
    if (AQ_ACT(d, p) < 0) continue;
    for (int rd = p * nr; rd < (p + 1) * nr; ++rd) cols_register(jac, AQ_REACT(d, rd), AQ_ACT(d, p));
    for (int pd = p * np; pd < (p + 1) * np; ++pd) cols_register(jac, AQ_PROD(d, pd), AQ_ACT(d, p));
  }
}

/* rxn_aqueous_equilibrium_update_ids, :138-230 */
static void aqeq_update_ids(Oracle *o, int *d, double *f) {
  int nr = AQ_NR(d), np = AQ_NP(d);
  for (int p = 0, k = 0; p < AQ_NPH(d); p++) {
    for (int r = 0; r < nr; r++) AQ_DERIV(d, k++) = o->deriv_ids[AQ_REACT(d, p * nr + r)];
    for (int q = 0; q < np; q++) AQ_DERIV(d, k++) = o->deriv_ids[AQ_PROD(d, p * np + q)];
  }
  const Csc *J = &o->rjac;
  for (int p = 0, j = 0; p < AQ_NPH(d); p++) {
    for (int ri = p * nr; ri < (p + 1) * nr; ri++) {
      for (int rd = p * nr; rd < (p + 1) * nr; rd++) AQ_JAC(d, j++) = csc_find(J, AQ_REACT(d, rd), AQ_REACT(d, ri));
      for (int pd = p * np; pd < (p + 1) * np; pd++) AQ_JAC(d, j++) = csc_find(J, AQ_PROD(d, pd), AQ_REACT(d, ri));
    }
    for (int pi = p * np; pi < (p + 1) * np; pi++) {
      for (int rd = p * nr; rd < (p + 1) * nr; rd++) AQ_JAC(d, j++) = csc_find(J, AQ_REACT(d, rd), AQ_PROD(d, pi));
      for (int pd = p * np; pd < (p + 1) * np; pd++) AQ_JAC(d, j++) = csc_find(J, AQ_PROD(d, pd), AQ_PROD(d, pi));
    }
    for (int rd = p * nr; rd < (p + 1) * nr; rd++) AQ_JAC(d, j++) = csc_find(J, AQ_REACT(d, rd), AQ_WATER(d, p));
    for (int pd = p * np; pd < (p + 1) * np; pd++) AQ_JAC(d, j++) = csc_find(J, AQ_PROD(d, pd), AQ_WATER(d, p));
    if (AQ_ACT(d, p) < 0) {
      for (int rd = p * nr; rd < (p + 1) * nr; ++rd) AQ_JAC(d, j++) = -1;
      for (int pd = p * np; pd < (p + 1) * np; ++pd) AQ_JAC(d, j++) = -1;
    } else {
      for (int rd = p * nr; rd < (p + 1) * nr; ++rd) AQ_JAC(d, j++) = csc_find(J, AQ_REACT(d, rd), AQ_ACT(d, p));
      for (int pd = p * np; pd < (p + 1) * np; ++pd) AQ_JAC(d, j++) = csc_find(J, AQ_PROD(d, pd), AQ_ACT(d, p));
    }
  }
  for (int p = 0; p < AQ_NPH(d); p++) {
    AQ_SMALL_CONC(d, f, p) = 99999.0;
    for (int r = 0; r < nr; r++)
      if (AQ_SMALL_CONC(d, f, p) > o->abs_tol[AQ_REACT(d, p * nr + r)] / 100.0)
        AQ_SMALL_CONC(d, f, p) = o->abs_tol[AQ_REACT(d, p * nr + r)] / 100.0;
    for (int q = 0; q < np; q++)
      if (AQ_SMALL_CONC(d, f, p) > o->abs_tol[AQ_PROD(d, p * np + q)] / 100.0)
        AQ_SMALL_CONC(d, f, p) = o->abs_tol[AQ_PROD(d, p * np + q)] / 100.0;
  }
  for (int p = 0; p < AQ_NPH(d); p++) AQ_SMALL_WATER(d, f, p) = o->abs_tol[AQ_WATER(d, p)] / 10.0;
}

/* rxn_aqueous_equilibrium_update_env_state, :244-265 */
static void aqeq_env(const double *f, double *e, double T) {
  double K = AQ_C(f) == 0.0 ? AQ_A(f) : AQ_A(f) * exp(AQ_C(f) * (1.0 / T - 1.0 / 298.0));
  e[0] = K * AQ_KREV(f);
}

/* calc_standard_rate, :278-310 */
static long double aqeq_standard_rate(int *d, double *f, const double *e, int water_partial,
                                      long double *rf, long double *rr) {
  long double water = AQ_WATER_CONC(f);
  long double react = (long double)AQ_REACT_CONC(d, f, 0) * AQ_MFTM(d, f, 0);
  for (int r = 1; r < AQ_NR(d); r++) react *= AQ_REACT_CONC(d, f, r) * AQ_MFTM(d, f, r) / water;
  long double prod = (long double)AQ_PROD_CONC(d, f, 0) * AQ_MFTM(d, f, AQ_NR(d));
  prod *= (long double)AQ_ACT_VAL(f);
  for (int q = 1; q < AQ_NP(d); q++)
    prod *= AQ_PROD_CONC(d, f, q) * AQ_MFTM(d, f, AQ_NR(d) + q) / water;
  *rf = e[0] * react;
  *rr = AQ_KREV(f) * prod;
  if (water_partial) return *rf * (AQ_NR(d) - 1) - *rr * (AQ_NP(d) - 1);
  return *rf - *rr;
}

/* rxn_aqueous_equilibrium_calc_deriv_contrib, :323-390 */
static void aqeq_deriv(Oracle *o, const double *state, int *d, double *f, const double *e) {
  int nr = AQ_NR(d), np = AQ_NP(d);
  for (int p = 0, k = 0; p < AQ_NPH(d); p++) {
    long double water = state[AQ_WATER(d, p)];
    if (water < MIN_WATER * AQ_SMALL_WATER(d, f, p)) {
      k += nr + np;
      continue;
    }
    for (int r = 0; r < nr; ++r) AQ_REACT_CONC(d, f, r) = state[AQ_REACT(d, p * nr + r)];
    for (int q = 0; q < np; ++q) AQ_PROD_CONC(d, f, q) = state[AQ_PROD(d, p * np + q)];
    AQ_WATER_CONC(f) = state[AQ_WATER(d, p)];
    AQ_ACT_VAL(f) = AQ_ACT(d, p) >= 0 ? state[AQ_ACT(d, p)] : 1.0;
    long double rf, rr;
    long double rate = aqeq_standard_rate(d, f, e, 0, &rf, &rr);
    if (rate == 0.0) {
      k += nr + np;
      continue;
    }
    for (int r = 0; r < nr; r++) {
      if (AQ_DERIV(d, k) < 0) {
        k++;
        continue;
      }
      td_add(o, AQ_DERIV(d, k), -rf / AQ_MFTM(d, f, r));
      td_add(o, AQ_DERIV(d, k++), rr / AQ_MFTM(d, f, r));
    }
    for (int q = 0; q < np; q++) {
      if (AQ_DERIV(d, k) < 0) {
        k++;
        continue;
      }
      td_add(o, AQ_DERIV(d, k), rf / AQ_MFTM(d, f, nr + q));
      td_add(o, AQ_DERIV(d, k++), -rr / AQ_MFTM(d, f, nr + q));
    }
  }
}

/* rxn_aqueous_equilibrium_calc_jac_contrib, :403-~548 */
static void aqeq_jac(Oracle *o, const double *state, int *d, double *f, const double *e) {
  int nr = AQ_NR(d), np = AQ_NP(d);
  for (int p = 0, j = 0; p < AQ_NPH(d); p++) {
    long double water = state[AQ_WATER(d, p)];
    if (water < MIN_WATER * AQ_SMALL_WATER(d, f, p)) {
      j += (nr + np) * (nr + np + 2);
      continue;
    }
    long double fwd = e[0];
    for (int r = 0; r < nr; r++) fwd *= state[AQ_REACT(d, p * nr + r)] * AQ_MFTM(d, f, r) / water;
    long double rev = AQ_KREV(f);
    for (int q = 0; q < np; q++) rev *= state[AQ_PROD(d, p * np + q)] * AQ_MFTM(d, f, nr + q) / water;
    if (AQ_ACT(d, p) >= 0) rev *= state[AQ_ACT(d, p)];
    for (int ri = 0; ri < nr; ri++) {
      for (int rd = 0; rd < nr; rd++) {
        if (AQ_JAC(d, j) < 0 || fwd == 0.0) {
          j++;
          continue;
        }
        jac_add(o, AQ_JAC(d, j++), 0,
                fwd / state[AQ_REACT(d, p * nr + ri)] / AQ_MFTM(d, f, rd) * water);
      }
      for (int pd = 0; pd < np; pd++) {
        if (AQ_JAC(d, j) < 0 || fwd == 0.0) {
          j++;
          continue;
        }
        jac_add(o, AQ_JAC(d, j++), 1,
                fwd / state[AQ_REACT(d, p * nr + ri)] / AQ_MFTM(d, f, nr + pd) * water);
      }
    }
    for (int pi = 0; pi < np; pi++) {
      for (int rd = 0; rd < nr; rd++) {
        if (AQ_JAC(d, j) < 0 || rev == 0.0) {
          j++;
          continue;
        }
        jac_add(o, AQ_JAC(d, j++), 1,
                rev / state[AQ_PROD(d, p * np + pi)] / AQ_MFTM(d, f, rd) * water);
      }
      for (int pd = 0; pd < np; pd++) {
        if (AQ_JAC(d, j) < 0 || rev == 0.0) {
          j++;
          continue;
        }
        jac_add(o, AQ_JAC(d, j++), 0,
                rev / state[AQ_PROD(d, p * np + pi)] / AQ_MFTM(d, f, nr + pd) * water);
      }
    }
    for (int rd = 0; rd < nr; rd++) {
      if (AQ_JAC(d, j) < 0) {
        j++;
        continue;
      }
      jac_add(o, AQ_JAC(d, j), 0, -fwd / AQ_MFTM(d, f, rd) * (nr - 1));
      jac_add(o, AQ_JAC(d, j++), 1, -rev / AQ_MFTM(d, f, rd) * (np - 1));
    }
    for (int pd = 0; pd < np; pd++) {
      if (AQ_JAC(d, j) < 0) {
        j++;
        continue;
      }
      jac_add(o, AQ_JAC(d, j), 1, -fwd / AQ_MFTM(d, f, nr + pd) * (nr - 1));
      jac_add(o, AQ_JAC(d, j++), 0, -rev / AQ_MFTM(d, f, nr + pd) * (np - 1));
    }
    if (AQ_ACT(d, p) < 0) {
      j += nr + np;
      continue;
    }
    for (int rd = 0; rd < nr; rd++) {
      if (AQ_JAC(d, j) < 0) {
        j++;
        continue;
      }
      jac_add(o, AQ_JAC(d, j++), 1, rev / state[AQ_ACT(d, p)] / AQ_MFTM(d, f, rd) * water);
    }
    for (int pd = 0; pd < np; pd++) {
      if (AQ_JAC(d, j) < 0) {
        j++;
        continue;
      }
      jac_add(o, AQ_JAC(d, j++), 0, rev / state[AQ_ACT(d, p)] / AQ_MFTM(d, f, nr + pd) * water);
    }
  }
}

/* ZSR layout, sub_model_ZSR_aerosol_water.c:30-68 (d after the type slot) */
#define ZS_NPH(d) ((d)[0])
#define ZS_GAS_WATER(d) ((d)[1] - 1)
#define ZS_NPAIR(d) ((d)[2])
#define ZS_PHASE(d, p) ((d)[5 + (p)] - 1)
#define ZS_PILOC(d, x) ((d)[5 + ZS_NPH(d) + (x)] - 1)
#define ZS_PFLOC(d, x) ((d)[5 + ZS_NPH(d) + ZS_NPAIR(d) + (x)] - 1)
#define ZS_TYPE(d, x) ((d)[ZS_PILOC(d, x)])
#define ZS_J_NCAT(d, x) ((d)[ZS_PILOC(d, x) + 1])
#define ZS_J_NAN(d, x) ((d)[ZS_PILOC(d, x) + 2])
#define ZS_J_CAT(d, x) ((d)[ZS_PILOC(d, x) + 3])
#define ZS_J_AN(d, x) ((d)[ZS_PILOC(d, x) + 4])
#define ZS_J_NY(d, x) ((d)[ZS_PILOC(d, x) + 5])
#define ZS_J_GW_JAC(d, p, x) ((d)[ZS_PILOC(d, x) + 6 + (p)])
#define ZS_J_CAT_JAC(d, p, x) ((d)[ZS_PILOC(d, x) + 6 + ZS_NPH(d) + (p)])
#define ZS_J_AN_JAC(d, p, x) ((d)[ZS_PILOC(d, x) + 6 + 2 * ZS_NPH(d) + (p)])
#define ZS_J_LOW_RH(d, f, x) ((f)[ZS_PFLOC(d, x)])
#define ZS_J_CAT_MW(d, f, x) ((f)[ZS_PFLOC(d, x) + 1])
#define ZS_J_AN_MW(d, f, x) ((f)[ZS_PFLOC(d, x) + 2])
#define ZS_J_Y(d, f, x, y) ((f)[ZS_PFLOC(d, x) + 3 + (y)])
#define ZS_ALPHA (-100.0)
#define ZS_MIN_WATER 1.0e-30L
#define ZS_JACOBSON 1

static void zsr_check_type(int *d, int x) {
  if (ZS_TYPE(d, x) != ZS_JACOBSON) {
    fprintf(stderr, "oracle: ZSR ion-pair type %d is not restated\n", ZS_TYPE(d, x));
    exit(EXIT_FAILURE);
  }
}

/* sub_model_ZSR_aerosol_water_get_used_jac_elem, :90-122 */
static void zsr_used_jac_elem(int *d, ColElems *jac) {
  for (int p = 0; p < ZS_NPH(d); ++p) {
    cols_register(jac, ZS_PHASE(d, p), ZS_GAS_WATER(d));
    for (int x = 0; x < ZS_NPAIR(d); ++x) {
      zsr_check_type(d, x);
      cols_register(jac, ZS_PHASE(d, p), ZS_PHASE(d, p) + ZS_J_CAT(d, x));
      cols_register(jac, ZS_PHASE(d, p), ZS_PHASE(d, p) + ZS_J_AN(d, x));
    }
  }
}

/* sub_model_ZSR_aerosol_water_update_ids, :134-173 */
static void zsr_update_ids(Oracle *o, int *d) {
  for (int p = 0; p < ZS_NPH(d); ++p)
    for (int x = 0; x < ZS_NPAIR(d); ++x) {
      ZS_J_GW_JAC(d, p, x) = csc_find(&o->pjac, ZS_PHASE(d, p), ZS_GAS_WATER(d));
      ZS_J_CAT_JAC(d, p, x) = csc_find(&o->pjac, ZS_PHASE(d, p), ZS_PHASE(d, p) + ZS_J_CAT(d, x));
      ZS_J_AN_JAC(d, p, x) = csc_find(&o->pjac, ZS_PHASE(d, p), ZS_PHASE(d, p) + ZS_J_AN(d, x));
    }
}

/* sub_model_ZSR_aerosol_water_update_env_state, :189-207 */
static void zsr_env(double *e, double T, double P) {
  double t_steam = 373.15;
  double a = 1.0 - t_steam / T;
  a = (((-0.1299 * a - 0.6445) * a - 1.976) * a + 13.3185) * a;
  double water_vp = 101325.0 * exp(a);
  e[0] = P / water_vp / 1.0e6;
}

/* sub_model_ZSR_aerosol_water_calculate, :216-301 (JACOBSON pairs) */
static void zsr_calculate(int *d, double *f, const double *e, double *state) {
  long double a_w = e[0] * state[ZS_GAS_WATER(d)];
  for (int p = 0; p < ZS_NPH(d); p++) {
    double *water = &state[ZS_PHASE(d, p)];
    *water = ZS_MIN_WATER;
    for (int x = 0; x < ZS_NPAIR(d); x++) {
      long double molality, conc;
      long double j_aw = a_w > ZS_J_LOW_RH(d, f, x) ? a_w : ZS_J_LOW_RH(d, f, x);
      molality = 0.0;
      for (int k = 0; k < ZS_J_NY(d, x); k++) molality += ZS_J_Y(d, f, x, k) * pow(j_aw, k);
      molality *= molality;
      long double cation = state[ZS_PHASE(d, p) + ZS_J_CAT(d, x)] / ZS_J_NCAT(d, x) /
                           ZS_J_CAT_MW(d, f, x) / 1000.0;
      long double anion = state[ZS_PHASE(d, p) + ZS_J_AN(d, x)] / ZS_J_NAN(d, x) /
                          ZS_J_AN_MW(d, f, x) / 1000.0;
      long double e_ac = exp(ZS_ALPHA * cation);
      long double e_aa = exp(ZS_ALPHA * anion);
      conc = (cation * e_ac + anion * e_aa) / (e_ac + e_aa);
      *water += conc / molality * 1000.0;
    }
  }
}

/* sub_model_ZSR_aerosol_water_get_jac_contrib, :314-~440 (JACOBSON pairs) */
static void zsr_jac(int *d, double *f, const double *e, const double *state, double *J) {
  long double a_w = e[0] * state[ZS_GAS_WATER(d)];
  long double d_aw_d_wg = e[0];
  for (int p = 0; p < ZS_NPH(d); p++)
    for (int x = 0; x < ZS_NPAIR(d); x++) {
      long double molality, d_molal_d_wg, conc;
      long double j_aw = a_w > ZS_J_LOW_RH(d, f, x) ? a_w : ZS_J_LOW_RH(d, f, x);
      long double d_jaw_d_wg = a_w > ZS_J_LOW_RH(d, f, x) ? d_aw_d_wg : 0.0;
      molality = ZS_J_Y(d, f, x, 0);
      d_molal_d_wg = 0.0;
      for (int k = 1; k < ZS_J_NY(d, x); k++) {
        molality += ZS_J_Y(d, f, x, k) * pow(j_aw, k);
        d_molal_d_wg += ZS_J_Y(d, f, x, k) * k * pow(j_aw, (k - 1));
      }
      d_molal_d_wg *= d_jaw_d_wg;
      long double cation = state[ZS_PHASE(d, p) + ZS_J_CAT(d, x)] / ZS_J_NCAT(d, x) /
                           ZS_J_CAT_MW(d, f, x) / 1000.0;
      long double d_cation_d_C = 1.0 / ZS_J_NCAT(d, x) / ZS_J_CAT_MW(d, f, x) / 1000.0;
      long double anion = state[ZS_PHASE(d, p) + ZS_J_AN(d, x)] / ZS_J_NAN(d, x) /
                          ZS_J_AN_MW(d, f, x) / 1000.0;
      long double d_anion_d_A = 1.0 / ZS_J_NAN(d, x) / ZS_J_AN_MW(d, f, x) / 1000.0;
      long double e_ac = exp(ZS_ALPHA * cation);
      long double e_aa = exp(ZS_ALPHA * anion);
      conc = (cation * e_ac + anion * e_aa) / (e_ac + e_aa);
      long double denom = (e_ac + e_aa) * (e_ac + e_aa);
      long double d_conc_d_cation =
          (e_ac * e_ac + e_ac * e_aa * (1.0 - ZS_ALPHA * anion + ZS_ALPHA * cation)) / denom;
      long double d_conc_d_anion =
          (e_aa * e_aa + e_ac * e_aa * (1.0 - ZS_ALPHA * cation + ZS_ALPHA * anion)) / denom;
      J[ZS_J_GW_JAC(d, p, x)] += -2.0 * conc / pow(molality, 3) * 1000.0 * d_molal_d_wg;
      J[ZS_J_AN_JAC(d, p, x)] += 1.0 / pow(molality, 2) * 1000.0 * d_conc_d_anion * d_anion_d_A;
      J[ZS_J_CAT_JAC(d, p, x)] += 1.0 / pow(molality, 2) * 1000.0 * d_conc_d_cation * d_cation_d_C;
    }
}

/* aero_phase/aero_rep/sub_model _add_condensed_data (aero_phase_solver.c:
 * 165-194, aero_rep_solver.c:443-482, sub_model_solver.c:413-453), with
 * the array sizes solver_new takes (camp_solver.c:200-349). */
void oracle_set_aero_sizes(void *om, int n_phase, int n_phase_int, int n_phase_float,
                           int n_rep, int n_rep_int, int n_rep_float, int n_rep_env,
                           int n_sub, int n_sub_int, int n_sub_float, int n_sub_env) {
  Oracle *o = om;
  o->n_phase = n_phase;
  o->ph_int = xcalloc(n_phase_int + 1, sizeof(int));
  o->ph_flt = xcalloc(n_phase_float + 1, sizeof(double));
  o->ph_int_idx = xcalloc(n_phase + 1, sizeof(int));
  o->ph_flt_idx = xcalloc(n_phase + 1, sizeof(int));
  o->n_rep = n_rep;
  o->rep_int = xcalloc(n_rep_int + n_rep + 1, sizeof(int));
  o->rep_flt = xcalloc(n_rep_float + 1, sizeof(double));
  o->rep_env = xcalloc((size_t)o->nc * n_rep_env + 1, sizeof(double));
  o->rep_int_idx = xcalloc(n_rep + 1, sizeof(int));
  o->rep_flt_idx = xcalloc(n_rep + 1, sizeof(int));
  o->rep_env_idx = xcalloc(n_rep + 1, sizeof(int));
  o->n_sub = n_sub;
  o->sub_int = xcalloc(n_sub_int + n_sub + 1, sizeof(int));
  o->sub_flt = xcalloc(n_sub_float + 1, sizeof(double));
  o->sub_env = xcalloc((size_t)o->nc * n_sub_env + 1, sizeof(double));
  o->sub_int_idx = xcalloc(n_sub + 1, sizeof(int));
  o->sub_flt_idx = xcalloc(n_sub + 1, sizeof(int));
  o->sub_env_idx = xcalloc(n_sub + 1, sizeof(int));
}

void oracle_add_aero_phase(void *om, int n_int, int n_float, const int *ip, const double *fp) {
  Oracle *o = om;
  int k = o->n_added_phase++;
  o->ph_int_idx[k + 1] = o->ph_int_idx[k] + n_int;
  o->ph_flt_idx[k + 1] = o->ph_flt_idx[k] + n_float;
  memcpy(&o->ph_int[o->ph_int_idx[k]], ip, n_int * sizeof(int));
  memcpy(&o->ph_flt[o->ph_flt_idx[k]], fp, n_float * sizeof(double));
}

void oracle_add_aero_rep(void *om, int type, int n_int, int n_float, int n_env, const int *ip,
                         const double *fp) {
  Oracle *o = om;
  int k = o->n_added_rep++;
  o->rep_int_idx[k + 1] = o->rep_int_idx[k] + n_int + 1;
  o->rep_flt_idx[k + 1] = o->rep_flt_idx[k] + n_float;
  o->rep_env_idx[k + 1] = o->rep_env_idx[k] + n_env;
  int *id = &o->rep_int[o->rep_int_idx[k]];
  id[0] = type;
  memcpy(id + 1, ip, n_int * sizeof(int));
  memcpy(&o->rep_flt[o->rep_flt_idx[k]], fp, n_float * sizeof(double));
  o->n_rep_env += n_env;
}

void oracle_add_sub_model(void *om, int type, int n_int, int n_float, int n_env, const int *ip,
                          const double *fp) {
  Oracle *o = om;
  int k = o->n_added_sub++;
  o->sub_int_idx[k + 1] = o->sub_int_idx[k] + n_int + 1;
  o->sub_flt_idx[k + 1] = o->sub_flt_idx[k] + n_float;
  o->sub_env_idx[k + 1] = o->sub_env_idx[k] + n_env;
  int *id = &o->sub_int[o->sub_int_idx[k]];
  id[0] = type;
  memcpy(id + 1, ip, n_int * sizeof(int));
  memcpy(&o->sub_flt[o->sub_flt_idx[k]], fp, n_float * sizeof(double));
  o->n_sub_env += n_env;
}

#define SUB_INT(o, k) (&(o)->sub_int[(o)->sub_int_idx[k]])
#define SUB_FLT(o, k) (&(o)->sub_flt[(o)->sub_flt_idx[k]])
#define SUB_ENV(o, k, cell) (&(o)->sub_env[(size_t)(cell) * (o)->n_sub_env + (o)->sub_env_idx[k]])

static void sub_check(int *id) {
  if (id[0] != SUB_MODEL_ZSR_AEROSOL_WATER) {
    fprintf(stderr, "oracle: unsupported sub model type %d\n", id[0]);
    exit(EXIT_FAILURE);
  }
}

/* sub_model_calculate, sub_model_solver.c:308-345 */
static void sub_calculate(Oracle *o, int cell, double *state) {
  for (int k = 0; k < o->n_sub; ++k) {
    int *id = SUB_INT(o, k);
    sub_check(id);
    zsr_calculate(id + 1, SUB_FLT(o, k), SUB_ENV(o, k, cell), state);
  }
}

/* sub_model_get_jac_contrib, sub_model_solver.c:354-395 */
static void sub_jac(Oracle *o, int cell, const double *state, double *J) {
  for (int k = 0; k < o->n_sub; ++k) {
    int *id = SUB_INT(o, k);
    sub_check(id);
    zsr_jac(id + 1, SUB_FLT(o, k), SUB_ENV(o, k, cell), state, J);
  }
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
    } else if (t == RXN_AQUEOUS_EQUILIBRIUM) {
      aqeq_used_jac_elem(id, jac);
    } else if (t == RXN_SIMPOL_PHASE_TRANSFER) {
      simpol_used_jac_elem(o, id, jac);
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
    } else if (t == RXN_AQUEOUS_EQUILIBRIUM) {
      aqeq_update_ids(o, id, RX_FLT(o, r));
    } else if (t == RXN_SIMPOL_PHASE_TRANSFER) {
      simpol_update_ids(o, id);
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

  /* sub-model parameter Jacobian: the dummy (0,0) element, then the sub
   * models' elements (camp_solver.c:1509-1550) */
  ColElems pj;
  cols_init(&pj, ns);
  cols_register(&pj, 0, 0);
  for (int k = 0; k < o->n_sub; ++k) {
    int *id = SUB_INT(o, k);
    sub_check(id);
    zsr_used_jac_elem(id + 1, &pj);
  }
  o->pjac = cols_build(&pj);
  for (int k = 0; k < o->n_sub; ++k) zsr_update_ids(o, SUB_INT(o, k) + 1);
  o->J_params = xcalloc(o->pjac.nnz, sizeof(double));

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
  free(o->ph_int); free(o->ph_flt); free(o->ph_int_idx); free(o->ph_flt_idx);
  free(o->rep_int); free(o->rep_flt); free(o->rep_env); free(o->rep_int_idx);
  free(o->rep_flt_idx); free(o->rep_env_idx);
  free(o->sub_int); free(o->sub_flt); free(o->sub_env); free(o->sub_int_idx);
  free(o->sub_flt_idx); free(o->sub_env_idx);
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
    free(o->J_params);
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
  /* aero_rep_update_env_state (nothing for modal/binned mass, :137-146) and
   * sub_model_update_env_state (camp_solver.c:599-601) */
  for (int k = 0; k < o->n_sub; ++k) {
    sub_check(SUB_INT(o, k));
    zsr_env(SUB_ENV(o, k, cell), T, P);
  }
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
      case RXN_CMAQ_H2O2: {
        /* rxn_CMAQ_H2O2.c:106-124: k = (k1 + k2 [M]) conv^(nr-1), [M] in k2_A */
        double conv = f[6] * P / T;
        e[0] = (f[0] * (f[2] == 0.0 ? 1.0 : exp(f[2] / T)) *
                    (f[1] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[1])) +
                f[3] * (f[5] == 0.0 ? 1.0 : exp(f[5] / T)) *
                    (f[4] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[4])) * conv) *
               pow(conv, NR(id) - 1);
        break;
      }
      case RXN_CMAQ_OH_HNO3: {
        /* rxn_CMAQ_OH_HNO3.c:110-133: k = (k0 + k3 / (1 + k3 / k2))
         * conv^(nr-1) SCALING, [M] in k3_A */
        double conv = f[10] * P / T;
        double k2 = f[3] * (f[5] == 0.0 ? 1.0 : exp(f[5] / T)) *
                    (f[4] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[4]));
        double k3 = f[6] * (f[8] == 0.0 ? 1.0 : exp(f[8] / T)) *
                    (f[7] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[7])) * conv;
        e[0] = (f[0] * (f[2] == 0.0 ? 1.0 : exp(f[2] / T)) *
                    (f[1] == 0.0 ? 1.0 : pow(T / ((double)300.0), f[1])) +
                k3 / (((double)1.0) + k3 / k2)) *
               pow(conv, NR(id) - 1) * f[9];
        break;
      }
      case RXN_PHOTOLYSIS:
      case RXN_EMISSION:
      case RXN_FIRST_ORDER_LOSS:
        e[0] = f[0] * e[1];
        break;
      case RXN_AQUEOUS_EQUILIBRIUM:
        aqeq_env(f, e, T);
        break;
      case RXN_SIMPOL_PHASE_TRANSFER:
        simpol_env(f, e, T, P);
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
static void cell_deriv(Oracle *o, int cell, double *state, double *deriv,
                       const double *est, double time_step) {
  /* aero_rep_update_state and sub_model_calculate (camp_solver.c:913-917) */
  rep_update_state(o, cell, state);
  sub_calculate(o, cell, state);
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
    } else if (t == RXN_AQUEOUS_EQUILIBRIUM) {
      aqeq_deriv(o, state, id, f, e);
    } else if (t == RXN_SIMPOL_PHASE_TRANSFER) {
      simpol_deriv(o, cell, state, id, f, e);
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
static void cell_jac(Oracle *o, int cell, double *state, double *J,
                     double time_step) {
  for (int e = 0; e < o->pjac.nnz; ++e) o->J_params[e] = 0.0;
  for (int e = 0; e < o->rjac.nnz; ++e) o->jp[e] = o->jl[e] = 0.0;
  /* aero_rep_update_state, sub_model_calculate, sub-model Jacobian
   * (camp_solver.c:1065-1072) */
  rep_update_state(o, cell, state);
  sub_calculate(o, cell, state);
  sub_jac(o, cell, state, o->J_params);
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
    } else if (t == RXN_AQUEOUS_EQUILIBRIUM) {
      aqeq_jac(o, state, id, f, e);
    } else if (t == RXN_SIMPOL_PHASE_TRANSFER) {
      simpol_jac(o, cell, state, id, f, e);
    }
  }
  for (int e = 0; e < o->rjac.nnz; ++e) o->J_rxn[e] = o->jp[e] - o->jl[e];
  for (int k = 0; k < o->nnz; ++k) J[k] = 0.0;
  /* J += J_rxn * J_params over the map, J_params[0] = 1 (:1100-1106) */
  o->J_params[0] = 1.0;
  for (int m = 0; m < o->n_mapped; ++m)
    J[o->map_solver[m]] += o->J_rxn[o->map_rxn[m]] * o->J_params[o->map_param[m]];
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
  int q, qprime, next_q, L, qwait, qu, indx_acor, jcur, use_acor_init, projected;
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

/* Non-negativity safeguard (not in CVODE).  CAMP's f() rejects
 * y < -SMALL (camp_solver.c:816-822) as a recoverable failure, and CVODE
 * answers that only by cutting the step, so a species driven to ~0 by a
 * fast loss stalls the integration (a converged iterate just below -SMALL
 * even makes every later prediction fail).  Wherever f() would see the
 * iterate y = zn0 + corr (the prediction, each Newton iterate, and the
 * converged value), components below -SMALL by at most their tolerance
 * (-y_i * ewt_i <= 1) are projected to TINY, the floor solver_run puts
 * under every concentration (camp_solver.c:554-567), through corr.
 * Exact zeros are avoided because the aerosol Jacobians divide by
 * concentrations.  Returns -1 (nothing below -SMALL), 0 (projected) or 1
 * (a larger excursion: the recoverable f() failure it implies). */
static int cv_project(Cv *cv, double *corr) {
  double worst = -1.0;
  for (int i = 0; i < cv->n; ++i)
    if (cv->y[i] < -SMALL) {
      double r = -cv->y[i] * cv->ewt[i];
      worst = r > worst ? r : worst;
    }
  if (worst > 1.0) return 1;
  if (worst < 0.0) return -1;
  for (int i = 0; i < cv->n; ++i)
    if (cv->y[i] < -SMALL) {
      corr[i] = TINY - cv->zn[0][i];
      cv->y[i] = TINY;
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
      int projected = 0;
      if (cv->o->neg_mode == 1) {
        int pr = cv_project(cv, cv->acor);
        if (pr > 0) return cv->jcur ? RHSFUNC_RECVR : TRY_AGAIN;
        projected = pr == 0;
      }
      cv->projected = projected;
      cv->acnrm = (m == 0 && !projected) ? del : wrms(cv, cv->acor);
      cv->jcur = 0;
      return CV_SUCCESS;
    }
    m++;
    if ((m == MAXCOR) || ((m >= 2) && (del > RDIV * delp)))
      return cv->jcur ? CONV_FAIL : TRY_AGAIN;
    delp = del;
    if (cv->o->neg_mode == 1 && cv_project(cv, cv->acor) > 0)
      return cv->jcur ? RHSFUNC_RECVR : TRY_AGAIN;
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
    if (cv->o->neg_mode == 1) {
      double mn = cv->y[0];
      for (int i = 1; i < cv->n; ++i) mn = cv->y[i] < mn ? cv->y[i] : mn;
      if (mn < -SMALL) {
        if (!cv->use_acor_init) {
          for (int i = 0; i < cv->n; ++i) cv->acor_init[i] = 0.0;
          cv->use_acor_init = 1;
        }
        cv_project(cv, cv->acor_init);
      }
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
  /* a component projected to TINY restarts flat at TINY: its history is
   * cleared so that the next prediction is TINY rather than a negative
   * extrapolation */
  if (cv->projected)
    for (int i = 0; i < cv->n; ++i)
      if (cv->y[i] == TINY) {
        cv->zn[0][i] = TINY;
        for (int j = 1; j <= cv->q; ++j) cv->zn[j][i] = 0.0;
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
    /* the post-solve sub_model_calculate runs on the cell the last f()/Jac()
     * loop left current, the problem's last cell (camp_solver.c:685-687) */
    if (o->n_rxn > 0)
      sub_calculate(o, c0 + nc - 1, &o->state[(size_t)(c0 + nc - 1) * ns]);
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

/* The Jacobian-extrapolated derivative alone (the `est` argument that
 * camp_solver.c:877-879 hands to time_derivative_output), from the cell's
 * current J history. */
void oracle_eval_est(void *om, int cell, const double *state_cell, double *est) {
  Oracle *o = om;
  double *y = xmalloc((size_t)(o->nv ? o->nv : 1) * sizeof(double));
  for (int d = 0; d < o->nv; ++d) y[d] = state_cell[o->dep_state[d]];
  cell_est(o, cell, y, est);
  free(y);
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

/* --------------------------------------------------------------------- */
/* Multi-phase introspection for the parity tests                         */
/* --------------------------------------------------------------------- */
int oracle_param_pattern(void *om, int *colptr, int *rowval) {
  Oracle *o = om;
  if (colptr) memcpy(colptr, o->pjac.colptr, (o->ns + 1) * sizeof(int));
  if (rowval) memcpy(rowval, o->pjac.rowval, o->pjac.nnz * sizeof(int));
  return o->pjac.nnz;
}

/* which: 0 rxn float data, 1 rep int data (with type slots), 2 rep float
 * data, 3 sub-model int data (with type slots), 4 sub-model env data */
void oracle_get_arrays(void *om, int which, void *dst, int n) {
  Oracle *o = om;
  switch (which) {
    case 0: memcpy(dst, o->rxn_float, n * sizeof(double)); break;
    case 1: memcpy(dst, o->rep_int, n * sizeof(int)); break;
    case 2: memcpy(dst, o->rep_flt, n * sizeof(double)); break;
    case 3: memcpy(dst, o->sub_int, n * sizeof(int)); break;
    case 4: memcpy(dst, o->sub_env, n * sizeof(double)); break;
  }
}

/* aero_rep_update_state + sub_model_calculate on one cell's state */
void oracle_cell_update(void *om, int cell, double *state) {
  Oracle *o = om;
  rep_update_state(o, cell, state);
  sub_calculate(o, cell, state);
}

/* which: 0 radius, 1 number concentration, 2 phase mass, 3 phase avg MW */
void oracle_rep_get(void *om, int rep, int cell, const double *state, int phase, int which,
                    double *value, double *partial) {
  Oracle *o = om;
  switch (which) {
    case 0: rep_eff_radius(o, rep, cell, phase, value, partial); break;
    case 1: rep_number_conc(o, rep, cell, state, phase, value, partial); break;
    case 2: rep_phase_prop(o, rep, cell, state, phase, 0, value, partial); break;
    case 3: rep_phase_prop(o, rep, cell, state, phase, 1, value, partial); break;
  }
}

/* The two factors of the solver Jacobian for one cell: the reaction
 * Jacobian (rjac order) and the sub-model parameter Jacobian (pjac order,
 * dummy element included), as Jac() forms them (camp_solver.c:1052-1106). */
int oracle_eval_jac_parts(void *om, int cell, const double *state_cell, double *j_rxn,
                          double *j_params) {
  Oracle *o = om;
  double *st = xmalloc((size_t)o->ns * sizeof(double));
  double *y = xmalloc((size_t)(o->nv ? o->nv : 1) * sizeof(double));
  double *fy = xmalloc((size_t)(o->nv ? o->nv : 1) * sizeof(double));
  double *jac = xmalloc((size_t)(o->nnz ? o->nnz : 1) * sizeof(double));
  memcpy(st, state_cell, o->ns * sizeof(double));
  for (int d = 0; d < o->nv; ++d) y[d] = state_cell[o->dep_state[d]];
  int rc = update_cell_state(o, st, y);
  if (!rc) {
    cell_deriv(o, cell, st, fy, NULL, 1.0);
    cell_jac(o, cell, st, jac, 1.0);
    memcpy(j_rxn, o->J_rxn, o->rjac.nnz * sizeof(double));
    memcpy(j_params, o->J_params, o->pjac.nnz * sizeof(double));
  }
  free(st); free(y); free(fy); free(jac);
  return rc;
}
