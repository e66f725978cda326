// Host-side model of the CAMP chemistry hot path and the device "program"
// compiled from it.
//
// The model holds what CAMP's ModelData holds for the gas-phase path
// (src/camp_common.h:67-179): state-variable types, the packed reaction
// parameters appended by rxn_add_condensed_data (rxn_solver.c:658-693) and
// the per-cell environment-dependent reaction data.  At solver_initialize
// it derives the bit-exact CAMP sparsity pattern and ids (camp_solver.c:
// 1417-1712) and compiles them into flat tables the HIP kernels read:
//   * rate terms (one per reaction) and per-species gather lists for f(y),
//   * per-Jacobian-element gather lists for J(y),
//   * a static-pivot sparse LU program (ordering, fill, per-pivot lists).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace camp {

// Rows of 64 LU update records the per-cell factorisation processes
// together: a pivot's first kLuRowBatch rows, then its further rows in
// batches of kLuRowBatch (the program pads those to whole batches).
constexpr int kLuRowBatch = 5;
// Row-view entries of the f estimate's ELL a lane reads per batch.
constexpr int kEstBatch = 8;

// src/camp_common.h:34-45, src/time_derivative.h:18
constexpr int kVariable = 1;
constexpr double kSmall = 1.0e-30;
constexpr double kTiny = 1.0e-60;
constexpr double kMaxPrecisionLoss = 1.0e-14;

// src/rxn_solver.c:19-35
constexpr int kRxnArrhenius = 1;
constexpr int kRxnTroe = 2;
constexpr int kRxnCmaqH2O2 = 3;
constexpr int kRxnCmaqOhHno3 = 4;
constexpr int kRxnPhotolysis = 5;
constexpr int kRxnFirstOrderLoss = 12;
constexpr int kRxnEmission = 13;

// Flags of a rate term.
// lanes per cell of the lane-group f/J kernels (the tunable widths)
constexpr int kGroupWidths[3] = {8, 16, 32};
constexpr int kTermMassAction = 1;   // skip contributions when the rate is 0
constexpr int kTermRoundRate = 2;    // rate product rounded to double (loss)
constexpr int kTermSlot = 4;         // contribution read from a per-cell slot

// Sparse matrix with per-column sorted rows (CSC), as built by Jacobian.c.
struct Csc {
  int n = 0;
  std::vector<int> colptr, rowval;
  int nnz() const { return (int)rowval.size(); }
  int find(int dep, int ind) const;
};

// Flat tables consumed by the kernels.  Index conventions: "state" indices
// address a cell's full state vector (n_state entries), "dep" indices its
// solver variables, "element" indices the per-cell solver Jacobian in CSC
// order, "pos" indices the per-cell LU value array.
struct Program {
  int ns = 0, nv = 0, nnz = 0, n_env = 0, n_terms = 0, lu_nnz = 0;
  int max_f_len = 0, max_j_len = 0;
  std::vector<int> dep_state;                       // [nv]
  std::vector<int> t_kidx, t_react, t_nreact, t_flags;  // terms; t_react[3*t+j]
  std::vector<int> f_ptr, f_term, f_guard;          // per dep species
  std::vector<double> f_coef;
  std::vector<int> j_ptr, j_term, j_skip, j_guard, j_rind;  // per element
  std::vector<double> j_coef;
  // The J gather lists as self-contained records: je_a = {k index, the
  // other reactants (-1 = none) x2, the differentiated reactant}, je_b =
  // guard product state id, je_c = coefficient
  std::vector<int> je_a, je_b;
  std::vector<double> je_c;
  // The f records per species (f_ptr order): fe_tb = {term index, or slot
  // index for slot records; product state id of the negative-yield guard
  // (-1 = none) | flags << 24} and fe_c = coefficient, combined as one
  // 16-byte record each in fe_rec {term or slot, guard | flags << 24,
  // coefficient (low, high word)}, from which the balanced gather is built
  std::vector<int> fe_tb;
  std::vector<double> fe_c;
  std::vector<int> fe_rec;
  // The per-species pass of f(y) (estimate and blend) in lane-slot order:
  // species are assigned to slots by row length, longest first (slot s =
  // pass * 64 + lane), fs_species[slot] is the species, and per pass fl_hdr
  // = {0, 0, offset and width of re_ell ({element, column}, padding -1,
  // rows padded to a multiple of 4 plus 4)}; entry j of lane l is at
  // offset + j * 64 + l, so a wavefront's loads are contiguous.
  std::vector<int> fs_species, fl_hdr, re_ell;
  // The f gather in balanced bins: the species' record lists are cut into
  // pieces of at most fb_rows records (a species longer than that is split
  // into near-equal pieces) and the pieces packed longest first into 64
  // lane bins of fb_rows records (a piece that no bin holds is split across
  // bins), so every lane walks fb_rows records.
  // fb_rec holds row j of lane l at j * 64 + l (rows beyond fb_rows, up to
  // a multiple of 4 plus 4, are padding) as {term or slot index |
  // target << 16, guard | flags << 24, coefficient (low, high word)}; flag
  // 16 ends a piece, flag 32 marks a piece of a split species.  A piece's
  // running sum is kept at its target in the gather's LDS scratch: a whole
  // species' value at [species] and magnitude sum at [nv + species], the
  // p-th partial's (hi, lo, magnitudes) at [2 nv + 3 p].  fb_fin[i] =
  // {first partial, partial count} per species (count 0: one piece, -1: no
  // records).
  int fb_rows = 0, fb_nparts = 0;
  std::vector<int> fb_rec, fb_fin;
  // rate terms as one 16-byte record: {k index | flags << 28, reactant state
  // ids x3 (-1 = none)}
  std::vector<int> t_rec;
  std::vector<int> colptr, rowval;                  // solver pattern (dep ids)
  std::vector<int> r_ptr, r_elem, r_col;            // row view of the pattern
  // LU (positions in the permuted LU value array)
  std::vector<int> perm, iperm;                     // perm[k] = dep index of pivot k
  std::vector<int> m_to_lu, diag_pos, e_diag;       // [nnz], [nv], [nnz]
  std::vector<int> l_ptr, l_pos, l_row;             // per pivot k: L(:,k) below k
  std::vector<int> u_ptr, u_pos, u_row;             // per pivot k: U(:,k) above k
  std::vector<int> up_ptr;                          // per pivot k: update range
  // the column solve's copies of L(:,k) and U(:,k): each column padded to
  // whole rows of 64 entries with (row -1, position 0), so that a column's
  // first 64 entries load with no bounds check; lu_hdr's ranges index them
  std::vector<int> cs_lpos, cs_lrow, cs_upos, cs_urow;
  std::vector<int> lu_hdr;  // per pivot k: {l_ptr[k], l_ptr[k+1], up_ptr[k], up_ptr[k+1],
                            //  diag_pos[k], u_ptr[k], u_ptr[k+1], 0} (two 16-byte
                            //  loads), then 2 empty pivots
  std::vector<uint16_t> up_dst, up_l, up_u;         // LU[dst] -= LU[l] * LU[u]
  // the factorisation's program (built in Model::initialize): per pivot L
  // positions (lf_pos) and update records dst | l << 16 | u << 32 (up_rec),
  // each padded to rows of 64 with writes to the lanes' sink slots past the
  // LU values, and lf_hdr = 4 ints per pivot
  std::vector<int> lf_pos, lf_hdr;
  // the register triangular solve's program (built in Model::initialize):
  // per block of 8 pivots and lane, 8 words pos(row t) | pos(row t+64) << 16
  std::vector<int> ts_tbl;
  int ts_fwd_blocks = 0, ts_bwd_blocks = 0;
  std::vector<uint64_t> up_rec;
  long n_updates() const { return (long)up_dst.size(); }

  // ---- lane-group f and J (gas-phase programs) ------------------------------
  // G lanes integrate one cell's f(y) or J(y), 64 / G cells per wavefront,
  // G one of kGroupWidths.  The rate constants the terms use are read from a
  // compact per-cell array, k_slot[i] being the rxn_env slot of compact
  // index i.  gf_rec: the f records of whole species, the species packed
  // longest first into G bins (the least-loaded bin takes the next
  // species); row r of lane j at (r * G + j), 8 ints each:
  // {target species | last record of its species << 30, guard state id (-1
  // none), coefficient (low, high word), compact k index | term flags << 28,
  // reactant state ids x3 (-1 none)}; padding rows target species nv with
  // coefficient 0.  gj_rec: the J records of contiguous element runs of
  // near-equal record counts per lane, 8 ints each: {target element | last
  // record of its element << 30, guard state id, coefficient (low, high),
  // compact k index, other reactants x2 (-1 none), the differentiated
  // reactant's state id}; padding rows target element nnz.  f_zero: the
  // species without records (their f is 0).
  // One program per group width kGroupWidths[w] (lanes per cell).
  int g_f_rows[3] = {0, 0, 0}, g_j_rows[3] = {0, 0, 0}, n_k = 0;
  std::vector<int> gf_rec[3], gj_rec[3], k_slot, f_count, f_zero;
  // The per-cell kernels' J program (gas-phase programs): gj_rec's layout
  // for the 64 lanes of a wavefront, with the rate constant's rxn_env slot
  // in place of the compact k index; cj_rows rows (a multiple of 4) plus 4
  // rows of padding.
  int cj_rows = 0;
  std::vector<int> cj_rec;

  // ---- multi-phase tables (empty for gas-phase models) -------------------
  bool mp = false;
  int rnnz = 0;                          // reaction-pattern elements (J gather space)
  int n_pi = 0, n_zi = 0, n_sp = 0, n_aq = 0;  // instance counts of the tables below
  // aerosol phase instances (one per bin and phase): first state id, phase
  // model, bin diameter; phase models: counted flag, MW, density per species
  std::vector<int> pi_start, pi_model;
  std::vector<double> pi_dp;
  std::vector<int> pm_ptr, pm_counted;
  std::vector<double> pm_mw, pm_dens;
  // ZSR items (one per sub model and phase) and their JACOBSON ion pairs
  std::vector<int> zi_water, zi_gw, zi_sub, zi_pair_ptr;
  std::vector<int> zp_cat, zp_an, zp_ncat, zp_nan, zp_ny, zp_jgw, zp_jcat, zp_jan;
  std::vector<double> zp_catmw, zp_anmw, zp_lowrh, zp_y;   // zp_y: 8 per pair
  int n_sub_env = 0;
  // SIMPOL instances (one per reaction and phase)
  std::vector<int> sp_gas, sp_aero, sp_dgas, sp_daero, sp_pi, sp_env, sp_flt, sp_slot;
  std::vector<double> sp_radius;
  // aqueous-equilibrium instances: species (3 + 3 slots), water, ids, slots
  std::vector<int> aq_nr, aq_np, aq_sp, aq_water, aq_env, aq_flt, aq_deriv, aq_slot;
  std::vector<double> aq_small_water;
  int n_slots = 0;
  // special Jacobian records per reaction element: {kind, instance, p1, p2}, c
  std::vector<int> js_ptr, js_a;
  std::vector<double> js_c;
  // solver element e = sum over [jm_ptr[e], jm_ptr[e+1]) of J_rxn[jm_rxn] * J_params[jm_param]
  std::vector<int> jm_ptr, jm_rxn, jm_param;
  int pnnz = 0;
  // the order in which the per-cell kernels assemble J_rxn: mj_rows rows of
  // 64 reaction elements (lane l of row r takes mj_elist[64 r + l], -1 past
  // the end), elements sorted by the kinds of their records so that the
  // lanes of a row run the same record code
  int mj_rows = 0;
  std::vector<int> mj_elist;
  // each row's records as one stream, lane-major: record row k of row r at
  // mj_rhdr[2 r] + k holds, per lane, the lane's element's k-th record --
  // its mass-action records, then its special records -- as 8 ints
  // {kind (-1 mass action: k slot, reactants y1, y2 | the special record's
  // kind, instance, p1, p2; 0 padding), then (mass action) the
  // differentiated reactant and the guard, and the coefficient (low, high
  // word)}; mj_rhdr[2 r + 1] records in row r (at least 1), and one
  // padding record row past the last
  std::vector<int> mj_rec, mj_rhdr;
};


struct Model {
  // solver_new arguments (camp_solver.c:83-362)
  int ns = 0, nc = 0, nv = 0;
  std::vector<int> var_type, deriv_ids, dep_state;
  int n_rxn = 0, n_added = 0, n_env = 0;
  std::vector<int> rxn_int, rxn_int_idx, rxn_float_idx, rxn_env_idx;
  std::vector<double> rxn_float;
  std::vector<double> rxn_env;  // [nc][n_env], calloc-zeroed as in CAMP
  // aerosol phases, representations and sub models (camp_solver.c:200-349);
  // reps and sub models carry a type slot like reactions
  int n_phase = 0, n_added_phase = 0, n_rep = 0, n_added_rep = 0, n_sub = 0, n_added_sub = 0;
  int n_rep_env = 0, n_sub_env = 0;
  std::vector<int> ph_int, ph_int_idx, ph_flt_idx, rep_int, rep_int_idx, rep_flt_idx, rep_env_idx,
      sub_int, sub_int_idx, sub_flt_idx, sub_env_idx;
  std::vector<double> ph_flt, rep_flt, sub_flt;
  // solver_initialize arguments
  std::vector<double> abs_tol;  // [ns]
  double rel_tol = 1.0e-8;
  int max_steps = 10000, max_conv_fails = 1000;
  // pattern (camp_solver.c:1417-1712)
  Csc rjac, pjac, sjac;
  std::vector<int> map_solver, map_rxn, map_param;
  Program prog;
  bool initialized = false;

  void add_rxn(int type, int n_int, int n_float, int n_env_param,
               const int* int_param, const double* float_param);
  void size_aero(int n_phase, int n_phase_int, int n_phase_float, int n_rep, int n_rep_int,
                 int n_rep_float, int n_sub, int n_sub_int, int n_sub_float);
  void add_phase(int n_int, int n_float, const int* int_param, const double* float_param);
  void add_rep(int type, int n_int, int n_float, int n_env_param, const int* int_param,
               const double* float_param);
  void add_sub(int type, int n_int, int n_float, int n_env_param, const int* int_param,
               const double* float_param);
  bool multiphase() const { return n_phase + n_rep + n_sub > 0; }
  // multi-phase tables of the kernel program; f_slots[i] receives the
  // contribution slots that feed solver variable i
  std::string compile_multiphase(std::vector<std::vector<int>>& f_slots);
  // Builds pattern, ids, the kernel program and the LU ordering.
  // Returns an empty string on success, else an error message.
  std::string initialize();
  bool update_data(int cell, int* rxn_id, int type, int uid, double rate);
  int* rint(int r) { return &rxn_int[rxn_int_idx[r]]; }
  double* rflt(int r) { return &rxn_float[rxn_float_idx[r]]; }
};

// Markowitz ordering of a square pattern (dep-indexed CSC).
std::vector<int> markowitz_order(int n, const std::vector<int>& colptr,
                                 const std::vector<int>& rowval);

}  // namespace camp
