// Host model: packed-parameter store, CAMP sparsity pattern and ids, and
// compilation into the flat tables the HIP kernels read.
#include "camp_model.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace camp {

int Csc::find(int dep, int ind) const {
  for (int e = colptr[ind]; e < colptr[ind + 1]; ++e)
    if (rowval[e] == dep) return e;
  return -1;
}

namespace {

// Registration-then-build semantics of Jacobian.c:105-214: unique (dep,
// ind) pairs per column, each column sorted ascending at build time.
struct PatternBuilder {
  std::vector<std::vector<int>> cols;
  explicit PatternBuilder(int n) : cols(n) {}
  void add(int dep, int ind) {
    auto& c = cols[ind];
    if (std::find(c.begin(), c.end(), dep) == c.end()) c.push_back(dep);
  }
  Csc build() {
    Csc m;
    m.n = (int)cols.size();
    m.colptr.assign(m.n + 1, 0);
    for (int j = 0; j < m.n; ++j) {
      std::sort(cols[j].begin(), cols[j].end());
      m.colptr[j + 1] = m.colptr[j] + (int)cols[j].size();
      m.rowval.insert(m.rowval.end(), cols[j].begin(), cols[j].end());
    }
    return m;
  }
};

bool mass_action(int t) {
  return t == kRxnArrhenius || t == kRxnTroe || t == kRxnPhotolysis;
}

// Packed layout of the mass-action types (rxn_arrhenius.c:20-36,
// rxn_troe.c:20-39, rxn_photolysis.c:20-33); d points after the type slot.
struct MassActionView {
  int* d;
  double* f;
  int t;
  int npi() const { return t == kRxnPhotolysis ? 3 : 2; }
  int npf() const { return t == kRxnArrhenius ? 6 : t == kRxnTroe ? 10 : 1; }
  int nr() const { return d[0]; }
  int np() const { return d[1]; }
  int react(int x) const { return d[npi() + x] - 1; }
  int prod(int x) const { return d[npi() + nr() + x] - 1; }
  int& deriv_id(int x) { return d[npi() + nr() + np() + x]; }
  int& jac_id(int x) { return d[npi() + 2 * (nr() + np()) + x]; }
  double yield(int x) const { return f[npf() + x]; }
};

}  // namespace

void Model::add_rxn(int type, int n_int, int n_float, int n_env_param,
                    const int* int_param, const double* float_param) {
  int k = n_added;
  if (k >= n_rxn) return;
  rxn_int_idx[k + 1] = rxn_int_idx[k] + n_int + 1;
  rxn_float_idx[k + 1] = rxn_float_idx[k] + n_float;
  rxn_env_idx[k + 1] = rxn_env_idx[k] + n_env_param;
  int* id = &rxn_int[rxn_int_idx[k]];
  id[0] = type;
  std::memcpy(id + 1, int_param, sizeof(int) * n_int);
  if (n_float) std::memcpy(&rxn_float[rxn_float_idx[k]], float_param, sizeof(double) * n_float);
  n_env += n_env_param;
  ++n_added;
}

bool Model::update_data(int cell, int* rxn_id, int type, int uid, double rate) {
  for (; *rxn_id < n_rxn; (*rxn_id)++) {
    int* id = rint(*rxn_id);
    int t = id[0];
    if (t != type) continue;
    int my_id;
    if (t == kRxnPhotolysis) my_id = id[3];
    else if (t == kRxnEmission || t == kRxnFirstOrderLoss) my_id = id[1];
    else continue;
    if (uid == my_id && my_id > 0) {
      double* e = &rxn_env[(size_t)cell * n_env + rxn_env_idx[*rxn_id]];
      e[1] = rate;
      e[0] = rflt(*rxn_id)[0] * e[1];
      return true;
    }
  }
  return false;
}

std::vector<int> markowitz_order(int n, const std::vector<int>& colptr,
                                 const std::vector<int>& rowval) {
  std::vector<char> A((size_t)n * n, 0);
  for (int j = 0; j < n; ++j)
    for (int k = colptr[j]; k < colptr[j + 1]; ++k) A[(size_t)rowval[k] * n + j] = 1;
  std::vector<char> alive(n, 1);
  std::vector<int> order;
  for (int s = 0; s < n; ++s) {
    long best = -1;
    int bj = -1;
    for (int j = 0; j < n; ++j) {
      if (!alive[j]) continue;
      long r = -1, c = -1;
      for (int u = 0; u < n; ++u)
        if (alive[u]) {
          r += A[(size_t)j * n + u];
          c += A[(size_t)u * n + j];
        }
      long m = r * c;
      if (best < 0 || m < best) best = m, bj = j;
    }
    for (int i = 0; i < n; ++i) {
      if (!alive[i] || i == bj || !A[(size_t)i * n + bj]) continue;
      for (int c = 0; c < n; ++c)
        if (alive[c] && c != bj && A[(size_t)bj * n + c]) A[(size_t)i * n + c] = 1;
    }
    order.push_back(bj);
    alive[bj] = 0;
  }
  return order;
}

std::string Model::initialize() {
  if (n_added != n_rxn) return "not all reactions were added before solver_initialize";
  // --- reaction Jacobian pattern: diagonal + per-type used elements -------
  PatternBuilder rb(ns);
  for (int i = 0; i < ns; ++i) rb.add(i, i);
  for (int r = 0; r < n_rxn; ++r) {
    int* id = rint(r);
    int t = id[0];
    if (mass_action(t)) {
      MassActionView v{id + 1, rflt(r), t};
      for (int i = 0; i < v.nr(); ++i) {
        for (int d = 0; d < v.nr(); ++d) rb.add(v.react(d), v.react(i));
        for (int d = 0; d < v.np(); ++d) rb.add(v.prod(d), v.react(i));
      }
    } else if (t == kRxnFirstOrderLoss) {
      rb.add(id[2] - 1, id[2] - 1);
    } else if (t != kRxnEmission) {
      char buf[96];
      std::snprintf(buf, sizeof buf, "unsupported reaction type %d", t);
      return buf;
    }
  }
  rjac = rb.build();

  // --- derivative ids and in-place id update (rxn_*_update_ids) ----------
  deriv_ids.assign(ns, -1);
  dep_state.clear();
  for (int i = 0; i < ns; ++i)
    if (var_type[i] == kVariable) {
      deriv_ids[i] = (int)dep_state.size();
      dep_state.push_back(i);
    }
  nv = (int)dep_state.size();
  for (int r = 0; r < n_rxn; ++r) {
    int* id = rint(r);
    int t = id[0];
    if (mass_action(t)) {
      MassActionView v{id + 1, rflt(r), t};
      for (int i = 0; i < v.nr(); ++i) v.deriv_id(i) = deriv_ids[v.react(i)];
      for (int i = 0; i < v.np(); ++i) v.deriv_id(v.nr() + i) = deriv_ids[v.prod(i)];
      int e = 0;
      for (int i = 0; i < v.nr(); ++i) {
        for (int d = 0; d < v.nr(); ++d) v.jac_id(e++) = rjac.find(v.react(d), v.react(i));
        for (int d = 0; d < v.np(); ++d) v.jac_id(e++) = rjac.find(v.prod(d), v.react(i));
      }
    } else if (t == kRxnEmission) {
      id[3] = deriv_ids[id[2] - 1];
    } else if (t == kRxnFirstOrderLoss) {
      id[3] = deriv_ids[id[2] - 1];
      id[4] = rjac.find(id[2] - 1, id[2] - 1);
    }
  }

  // --- parameter (sub-model) pattern: only the dummy (0,0) element -------
  PatternBuilder pb(ns);
  pb.add(0, 0);
  pjac = pb.build();

  // --- solver Jacobian pattern and jac_map --------------------------------
  PatternBuilder sb(ns);
  for (int i_ind = 0; i_ind < ns; ++i_ind)
    for (int i_dep = 0; i_dep < ns; ++i_dep) {
      if (var_type[i_dep] != kVariable || rjac.find(i_dep, i_ind) == -1) continue;
      if (var_type[i_ind] == kVariable) {
        sb.add(i_dep, i_ind);
        continue;
      }
      for (int j_ind = 0; j_ind < ns; ++j_ind)
        if (pjac.find(i_ind, j_ind) != -1 && var_type[j_ind] == kVariable) sb.add(i_dep, j_ind);
    }
  sjac = sb.build();
  map_solver.clear();
  map_rxn.clear();
  map_param.clear();
  for (int i_ind = 0; i_ind < ns; ++i_ind)
    for (int e = rjac.colptr[i_ind]; e < rjac.colptr[i_ind + 1]; ++e) {
      int i_dep = rjac.rowval[e];
      if (var_type[i_dep] != kVariable) continue;
      if (var_type[i_ind] == kVariable) {
        map_solver.push_back(sjac.find(i_dep, i_ind));
        map_rxn.push_back(e);
        map_param.push_back(0);
        continue;
      }
      for (int j_ind = 0; j_ind < ns; ++j_ind)
        if (pjac.find(i_ind, j_ind) != -1 && var_type[j_ind] == kVariable) {
          map_solver.push_back(sjac.find(i_dep, j_ind));
          map_rxn.push_back(e);
          map_param.push_back(pjac.find(i_ind, j_ind));
        }
    }
  for (int p : map_param)
    if (p != 0) return "sub-model Jacobian parameters are not supported on this path";

  // --- compile the kernel program ----------------------------------------
  Program& P = prog;
  P = Program();
  P.ns = ns;
  P.nv = nv;
  P.n_env = n_env;
  P.nnz = sjac.nnz();
  P.dep_state = dep_state;
  P.colptr.assign(nv + 1, 0);
  for (int c = 0; c < ns; ++c)
    if (deriv_ids[c] >= 0) P.colptr[deriv_ids[c]] = sjac.colptr[c];
  P.colptr[nv] = P.nnz;
  P.rowval.resize(P.nnz);
  for (int e = 0; e < P.nnz; ++e) P.rowval[e] = deriv_ids[sjac.rowval[e]];
  std::vector<int> rxn_to_solver(rjac.nnz(), -1);
  for (size_t m = 0; m < map_rxn.size(); ++m) rxn_to_solver[map_rxn[m]] = map_solver[m];

  // rate terms: one per reaction
  P.n_terms = n_rxn;
  P.t_kidx.resize(n_rxn);
  P.t_react.assign(3 * n_rxn, -1);
  P.t_nreact.resize(n_rxn);
  P.t_flags.resize(n_rxn);
  std::vector<std::vector<int>> f_term(nv), f_guard(nv), j_term(P.nnz), j_skip(P.nnz),
      j_guard(P.nnz), j_rind(P.nnz);
  std::vector<std::vector<double>> f_coef(nv), j_coef(P.nnz);
  for (int r = 0; r < n_rxn; ++r) {
    int* id = rint(r);
    int t = id[0];
    P.t_kidx[r] = rxn_env_idx[r];
    if (mass_action(t)) {
      MassActionView v{id + 1, rflt(r), t};
      if (v.nr() > 3) return "reactions with more than 3 reactants are not supported";
      P.t_nreact[r] = v.nr();
      P.t_flags[r] = kTermMassAction;
      for (int i = 0; i < v.nr(); ++i) P.t_react[3 * r + i] = v.react(i);
      int dv = 0;
      for (int i = 0; i < v.nr(); ++i, ++dv)
        if (v.deriv_id(dv) >= 0) {
          f_term[v.deriv_id(dv)].push_back(r);
          f_coef[v.deriv_id(dv)].push_back(-1.0);
          f_guard[v.deriv_id(dv)].push_back(-1);
        }
      for (int i = 0; i < v.np(); ++i, ++dv)
        if (v.deriv_id(dv) >= 0) {
          f_term[v.deriv_id(dv)].push_back(r);
          f_coef[v.deriv_id(dv)].push_back(v.yield(i));
          f_guard[v.deriv_id(dv)].push_back(v.prod(i));
        }
      int el = 0;
      for (int i_ind = 0; i_ind < v.nr(); ++i_ind) {
        for (int d = 0; d < v.nr(); ++d, ++el) {
          int je = v.jac_id(el);
          if (je < 0 || rxn_to_solver[je] < 0) continue;
          int se = rxn_to_solver[je];
          j_term[se].push_back(r);
          j_skip[se].push_back(i_ind);
          j_coef[se].push_back(-1.0);
          j_guard[se].push_back(-1);
          j_rind[se].push_back(v.react(i_ind));
        }
        for (int d = 0; d < v.np(); ++d, ++el) {
          int je = v.jac_id(el);
          if (je < 0 || rxn_to_solver[je] < 0) continue;
          int se = rxn_to_solver[je];
          j_term[se].push_back(r);
          j_skip[se].push_back(i_ind);
          j_coef[se].push_back(v.yield(d));
          j_guard[se].push_back(v.prod(d));
          j_rind[se].push_back(v.react(i_ind));
        }
      }
    } else if (t == kRxnEmission) {
      P.t_nreact[r] = 0;
      P.t_flags[r] = 0;
      if (id[3] >= 0) {
        f_term[id[3]].push_back(r);
        f_coef[id[3]].push_back(1.0);
        f_guard[id[3]].push_back(-1);
      }
    } else if (t == kRxnFirstOrderLoss) {
      P.t_nreact[r] = 1;
      P.t_flags[r] = kTermRoundRate;
      P.t_react[3 * r] = id[2] - 1;
      if (id[3] >= 0) {
        f_term[id[3]].push_back(r);
        f_coef[id[3]].push_back(-1.0);
        f_guard[id[3]].push_back(-1);
      }
      if (id[4] >= 0 && rxn_to_solver[id[4]] >= 0) {
        int se = rxn_to_solver[id[4]];
        j_term[se].push_back(r);
        j_skip[se].push_back(0);
        j_coef[se].push_back(-1.0);
        j_guard[se].push_back(-1);
        j_rind[se].push_back(id[2] - 1);
      }
    }
  }
  P.f_ptr.assign(1, 0);
  for (int i = 0; i < nv; ++i) {
    P.f_term.insert(P.f_term.end(), f_term[i].begin(), f_term[i].end());
    P.f_coef.insert(P.f_coef.end(), f_coef[i].begin(), f_coef[i].end());
    P.f_guard.insert(P.f_guard.end(), f_guard[i].begin(), f_guard[i].end());
    P.f_ptr.push_back((int)P.f_term.size());
    P.max_f_len = std::max(P.max_f_len, (int)f_term[i].size());
  }
  P.j_ptr.assign(1, 0);
  for (int e = 0; e < P.nnz; ++e) {
    P.j_term.insert(P.j_term.end(), j_term[e].begin(), j_term[e].end());
    P.j_skip.insert(P.j_skip.end(), j_skip[e].begin(), j_skip[e].end());
    P.j_coef.insert(P.j_coef.end(), j_coef[e].begin(), j_coef[e].end());
    P.j_guard.insert(P.j_guard.end(), j_guard[e].begin(), j_guard[e].end());
    P.j_rind.insert(P.j_rind.end(), j_rind[e].begin(), j_rind[e].end());
    P.j_ptr.push_back((int)P.j_term.size());
    P.max_j_len = std::max(P.max_j_len, (int)j_term[e].size());
  }
  for (size_t q = 0; q < P.f_term.size(); ++q) {
    int t = P.f_term[q];
    P.fe_a.push_back(P.t_kidx[t]);
    for (int j = 0; j < 3; ++j) P.fe_a.push_back(j < P.t_nreact[t] ? P.t_react[3 * t + j] : -1);
    P.fe_b.push_back((P.f_guard[q] & 0xFFFFFF) | (P.t_flags[t] << 24));
    P.fe_c.push_back(P.f_coef[q]);
  }
  for (size_t q = 0; q < P.j_term.size(); ++q) {
    int t = P.j_term[q], skip = P.j_skip[q];
    int others[2] = {-1, -1}, n_o = 0;
    for (int j = 0; j < P.t_nreact[t]; ++j)
      if (j != skip) others[n_o++] = P.t_react[3 * t + j];
    P.je_a.push_back(P.t_kidx[t]);
    P.je_a.push_back(others[0]);
    P.je_a.push_back(others[1]);
    P.je_a.push_back(P.j_rind[q]);
    P.je_b.push_back(P.j_guard[q]);
    P.je_c.push_back(P.j_coef[q]);
  }
  if (P.fe_c.empty()) {
    P.fe_a.assign(4, -1); P.fe_b.push_back(-1); P.fe_c.push_back(0.0);
  }
  if (P.je_c.empty()) {
    P.je_a.assign(4, -1); P.je_b.push_back(-1); P.je_c.push_back(0.0);
  }
  if (P.f_term.empty()) { P.f_term.push_back(0); P.f_coef.push_back(0); P.f_guard.push_back(-1); }
  if (P.j_term.empty()) {
    P.j_term.push_back(0); P.j_skip.push_back(0); P.j_coef.push_back(0);
    P.j_guard.push_back(-1); P.j_rind.push_back(0);
  }

  // row view of the solver pattern (entries of each row in column order)
  P.r_ptr.assign(nv + 1, 0);
  for (int e = 0; e < P.nnz; ++e) P.r_ptr[P.rowval[e] + 1]++;
  for (int i = 0; i < nv; ++i) P.r_ptr[i + 1] += P.r_ptr[i];
  P.r_elem.resize(P.nnz);
  P.r_col.resize(P.nnz);
  {
    std::vector<int> fill(P.r_ptr.begin(), P.r_ptr.end() - 1);
    for (int j = 0; j < nv; ++j)
      for (int e = P.colptr[j]; e < P.colptr[j + 1]; ++e) {
        int i = P.rowval[e];
        P.r_elem[fill[i]] = e;
        P.r_col[fill[i]++] = j;
      }
  }

  // --- LU: Markowitz ordering, symbolic fill, per-pivot programs ----------
  P.perm = markowitz_order(nv, P.colptr, P.rowval);
  P.iperm.assign(nv, 0);
  for (int k = 0; k < nv; ++k) P.iperm[P.perm[k]] = k;
  std::vector<char> F((size_t)nv * nv, 0);
  for (int j = 0; j < nv; ++j)
    for (int e = P.colptr[j]; e < P.colptr[j + 1]; ++e)
      F[(size_t)P.iperm[P.rowval[e]] * nv + P.iperm[j]] = 1;
  for (int k = 0; k < nv; ++k)
    for (int i = k + 1; i < nv; ++i)
      if (F[(size_t)i * nv + k])
        for (int j = k + 1; j < nv; ++j)
          if (F[(size_t)k * nv + j]) F[(size_t)i * nv + j] = 1;
  std::vector<int> pos((size_t)nv * nv, -1);
  int p = 0;
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < nv; ++j)
      if (F[(size_t)i * nv + j]) pos[(size_t)i * nv + j] = p++;
  P.lu_nnz = p;
  if (P.lu_nnz > 65535) return "LU fill exceeds 65535 entries per cell";
  P.m_to_lu.resize(P.nnz);
  P.e_diag.assign(P.nnz, 0);
  for (int j = 0; j < nv; ++j)
    for (int e = P.colptr[j]; e < P.colptr[j + 1]; ++e) {
      P.m_to_lu[e] = pos[(size_t)P.iperm[P.rowval[e]] * nv + P.iperm[j]];
      P.e_diag[e] = P.rowval[e] == j;
    }
  P.diag_pos.resize(nv);
  for (int k = 0; k < nv; ++k) P.diag_pos[k] = pos[(size_t)k * nv + k];
  P.lu_rowptr.assign(nv + 1, 0);
  P.lu_col.assign(P.lu_nnz, 0);
  P.row_off.assign((size_t)nv * nv, 255);
  for (int i = 0; i < nv; ++i) {
    int first = -1, cnt = 0;
    for (int j = 0; j < nv; ++j) {
      int q = pos[(size_t)i * nv + j];
      if (q < 0) continue;
      if (first < 0) first = q;
      P.lu_col[q] = j;
      if (cnt < 255) P.row_off[(size_t)i * nv + j] = (uint8_t)(q - first);
      ++cnt;
    }
    P.lu_rowptr[i] = first;
    if (cnt >= 255) P.row_off.clear();  // rows too long for 8-bit offsets
  }
  P.lu_rowptr[nv] = P.lu_nnz;
  P.l_ptr.assign(1, 0);
  P.u_ptr.assign(1, 0);
  P.up_ptr.assign(1, 0);
  for (int k = 0; k < nv; ++k) {
    for (int i = k + 1; i < nv; ++i)
      if (F[(size_t)i * nv + k]) {
        P.l_pos.push_back(pos[(size_t)i * nv + k]);
        P.l_row.push_back(i);
      }
    for (int i = 0; i < k; ++i)
      if (F[(size_t)i * nv + k]) {
        P.u_pos.push_back(pos[(size_t)i * nv + k]);
        P.u_row.push_back(i);
      }
    for (int i = k + 1; i < nv; ++i) {
      if (!F[(size_t)i * nv + k]) continue;
      for (int j = k + 1; j < nv; ++j) {
        if (!F[(size_t)k * nv + j]) continue;
        P.up_dst.push_back((uint16_t)pos[(size_t)i * nv + j]);
        P.up_l.push_back((uint16_t)pos[(size_t)i * nv + k]);
        P.up_u.push_back((uint16_t)pos[(size_t)k * nv + j]);
      }
    }
    P.l_ptr.push_back((int)P.l_pos.size());
    P.u_ptr.push_back((int)P.u_pos.size());
    P.up_ptr.push_back((int)P.up_dst.size());
  }
  if (P.l_pos.empty()) { P.l_pos.push_back(0); P.l_row.push_back(0); }
  if (P.u_pos.empty()) { P.u_pos.push_back(0); P.u_row.push_back(0); }
  if (P.up_dst.empty()) { P.up_dst.push_back(0); P.up_l.push_back(0); P.up_u.push_back(0); }
  initialized = true;
  return std::string();
}

}  // namespace camp
