// Host model: packed-parameter store, CAMP sparsity pattern and ids, and
// compilation into the flat tables the HIP kernels read.
#include "camp_model.hpp"

#include "camp_aero.hpp"

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
  return t == kRxnArrhenius || t == kRxnTroe || t == kRxnPhotolysis || t == kRxnCmaqH2O2 ||
         t == kRxnCmaqOhHno3;
}

// Packed layout of the mass-action types (rxn_arrhenius.c:20-36,
// rxn_troe.c:20-39, rxn_CMAQ_H2O2.c:19-35, rxn_CMAQ_OH_HNO3.c:19-41,
// rxn_photolysis.c:20-33); d points after the type slot.
struct MassActionView {
  int* d;
  double* f;
  int t;
  int npi() const { return t == kRxnPhotolysis ? 3 : 2; }
  int npf() const {
    return t == kRxnArrhenius ? 6 : t == kRxnTroe ? 10 : t == kRxnCmaqH2O2 ? 7
                                  : t == kRxnCmaqOhHno3 ? 11 : 1;
  }
  int nr() const { return d[0]; }
  int np() const { return d[1]; }
  int react(int x) const { return d[npi() + x] - 1; }
  int prod(int x) const { return d[npi() + nr() + x] - 1; }
  int& deriv_id(int x) { return d[npi() + nr() + np() + x]; }
  int& jac_id(int x) { return d[npi() + 2 * (nr() + np()) + x]; }
  double yield(int x) const { return f[npf() + x]; }
};

// aero_phase_get_used_jac_elem (aero_phase_solver.c:40-63)
int phase_used_jac_elem(const Model& m, int p, int state_id, std::vector<char>& flags) {
  PhaseView v{&m.ph_int[m.ph_int_idx[p]], &m.ph_flt[m.ph_flt_idx[p]]};
  int n = 0;
  for (int i = 0; i < v.n(); ++i)
    if (v.counted(i)) {
      flags[state_id + i] = 1;
      ++n;
    }
  return n;
}

BinnedRepView rep_view(Model& m, int r) {
  return BinnedRepView{&m.rep_int[m.rep_int_idx[r]] + 1, &m.rep_flt[m.rep_flt_idx[r]]};
}

// aero_rep_modal_binned_mass_get_used_jac_elem (:84-113): flags the species
// of every phase in the bin that holds phase instance aero_phase_idx and
// records each phase's count in the representation's int data
int rep_used_jac_elem(Model& m, int r, int aero_phase_idx, std::vector<char>& flags) {
  BinnedRepView v = rep_view(m, r);
  int n = 0;
  for (int s = 0; s < v.nsec() && aero_phase_idx >= 0; s++)
    for (int b = 0; b < v.nbin(s) && aero_phase_idx >= 0; b++)
      for (int p = 0; p < v.nph(s) && aero_phase_idx >= 0; p++) {
        if (aero_phase_idx == 0)
          for (int j = 0; j < v.nph(s); j++) {
            v.phase_njac(s, j, b) =
                phase_used_jac_elem(m, v.phase_model(s, j, b), v.phase_state(s, j, b), flags);
            n += v.phase_njac(s, j, b);
          }
        aero_phase_idx -= 1;
      }
  return n;
}

}  // namespace

void Model::size_aero(int n_ph, int n_ph_int, int n_ph_flt, int n_rp, int n_rp_int,
                      int n_rp_flt, int n_sb, int n_sb_int, int n_sb_flt) {
  n_phase = n_ph;
  ph_int.assign(std::max(n_ph_int, 0), 0);
  ph_flt.assign(std::max(n_ph_flt, 0), 0.0);
  ph_int_idx.assign(n_ph + 1, 0);
  ph_flt_idx.assign(n_ph + 1, 0);
  n_rep = n_rp;
  rep_int.assign(std::max(n_rp_int, 0) + n_rp, 0);
  rep_flt.assign(std::max(n_rp_flt, 0), 0.0);
  rep_int_idx.assign(n_rp + 1, 0);
  rep_flt_idx.assign(n_rp + 1, 0);
  rep_env_idx.assign(n_rp + 1, 0);
  n_sub = n_sb;
  sub_int.assign(std::max(n_sb_int, 0) + n_sb, 0);
  sub_flt.assign(std::max(n_sb_flt, 0), 0.0);
  sub_int_idx.assign(n_sb + 1, 0);
  sub_flt_idx.assign(n_sb + 1, 0);
  sub_env_idx.assign(n_sb + 1, 0);
}

// aero_phase_solver.c:165-194
void Model::add_phase(int n_int, int n_float, const int* ip, const double* fp) {
  int k = n_added_phase;
  if (k >= n_phase) return;
  ph_int_idx[k + 1] = ph_int_idx[k] + n_int;
  ph_flt_idx[k + 1] = ph_flt_idx[k] + n_float;
  std::memcpy(&ph_int[ph_int_idx[k]], ip, sizeof(int) * n_int);
  if (n_float) std::memcpy(&ph_flt[ph_flt_idx[k]], fp, sizeof(double) * n_float);
  ++n_added_phase;
}

// aero_rep_solver.c:443-482
void Model::add_rep(int type, int n_int, int n_float, int n_env_param, const int* ip,
                    const double* fp) {
  int k = n_added_rep;
  if (k >= n_rep) return;
  rep_int_idx[k + 1] = rep_int_idx[k] + n_int + 1;
  rep_flt_idx[k + 1] = rep_flt_idx[k] + n_float;
  rep_env_idx[k + 1] = rep_env_idx[k] + n_env_param;
  int* id = &rep_int[rep_int_idx[k]];
  id[0] = type;
  std::memcpy(id + 1, ip, sizeof(int) * n_int);
  if (n_float) std::memcpy(&rep_flt[rep_flt_idx[k]], fp, sizeof(double) * n_float);
  n_rep_env += n_env_param;
  ++n_added_rep;
}

// sub_model_solver.c:413-453
void Model::add_sub(int type, int n_int, int n_float, int n_env_param, const int* ip,
                    const double* fp) {
  int k = n_added_sub;
  if (k >= n_sub) return;
  sub_int_idx[k + 1] = sub_int_idx[k] + n_int + 1;
  sub_flt_idx[k + 1] = sub_flt_idx[k] + n_float;
  sub_env_idx[k + 1] = sub_env_idx[k] + n_env_param;
  int* id = &sub_int[sub_int_idx[k]];
  id[0] = type;
  std::memcpy(id + 1, ip, sizeof(int) * n_int);
  if (n_float) std::memcpy(&sub_flt[sub_flt_idx[k]], fp, sizeof(double) * n_float);
  n_sub_env += n_env_param;
  ++n_added_sub;
}

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

// J records of contiguous element runs with near-equal record counts, one
// run per lane of G (the layout of Program::gj_rec); an element without
// records gets one record of coefficient 0, so that every element is
// written.  kc maps the rxn_env slot of a record's rate constant to its
// compact index; nullptr keeps the slot itself.
static void build_j_runs(const Program& P, int G, const std::vector<int>* kc,
                         std::vector<int>& gj, int& rows_out) {
  auto put_coef = [](int* r, double c) {
    uint64_t bits;
    std::memcpy(&bits, &c, sizeof(bits));
    r[0] = (int)(uint32_t)(bits & 0xFFFFFFFFu);
    r[1] = (int)(uint32_t)(bits >> 32);
  };
  const int nnz = P.nnz;
  auto nrec = [&](int e) { return std::max(1, P.j_ptr[e + 1] - P.j_ptr[e]); };
  std::vector<long> acc(nnz + 1, 0);
  for (int e = 0; e < nnz; ++e) acc[e + 1] = acc[e] + nrec(e);
  const long total = acc[nnz];
  std::vector<int> first(G + 1, nnz);
  first[0] = 0;
  for (int j = 1; j < G; ++j) {
    long target = total * j / G;
    int e = first[j - 1];
    while (e < nnz && acc[e] < target) ++e;
    first[j] = e;
  }
  int rows = 1;
  for (int j = 0; j < G; ++j) rows = std::max(rows, (int)(acc[first[j + 1]] - acc[first[j]]));
  rows_out = (rows + 3) / 4 * 4;
  // plus 4 rows that the kernels' look-ahead loads but never processes
  const int j_alloc = rows_out + 4;
  gj.assign((size_t)8 * G * j_alloc, 0);
  for (int j = 0; j < G; ++j) {
    int r = 0;
    for (int e = first[j]; e < first[j + 1]; ++e) {
      if (P.j_ptr[e] == P.j_ptr[e + 1]) {
        int* d = &gj[(size_t)8 * ((size_t)r * G + j)];
        d[0] = e | (1 << 30);
        d[1] = -1;
        put_coef(&d[2], 0.0);
        d[4] = 0;
        d[5] = d[6] = -1;
        d[7] = 0;
        ++r;
        continue;
      }
      for (int q = P.j_ptr[e]; q < P.j_ptr[e + 1]; ++q, ++r) {
        int* d = &gj[(size_t)8 * ((size_t)r * G + j)];
        d[0] = e | ((q + 1 == P.j_ptr[e + 1]) ? (1 << 30) : 0);
        d[1] = P.je_b[q];
        put_coef(&d[2], P.je_c[q]);
        d[4] = kc ? (*kc)[P.je_a[4 * q]] : P.je_a[4 * q];
        d[5] = P.je_a[4 * q + 1];
        d[6] = P.je_a[4 * q + 2];
        d[7] = P.je_a[4 * q + 3];
      }
    }
    for (; r < j_alloc; ++r) {
      int* d = &gj[(size_t)8 * ((size_t)r * G + j)];
      d[0] = nnz | (1 << 30);
      d[1] = -1;
      put_coef(&d[2], 0.0);
      d[4] = 0;
      d[5] = d[6] = -1;
      d[7] = 0;
    }
  }
}

// The lane-group f and J program (see Program::gf_rec): whole species
// packed longest first into G bins for f, contiguous element runs
// of near-equal record counts for J, and the compact k index map.
static void build_group_program(Program& P, int w) {
  const int G = kGroupWidths[w], nv = P.nv;
  // compact k: the rxn_env slot of each term, in term order
  std::vector<int> kc_of_slot(std::max(P.n_env, 1), -1);
  P.k_slot.clear();
  for (int t = 0; t < P.n_terms; ++t) {
    int sl = P.t_kidx[t];
    if (kc_of_slot[sl] < 0) {
      kc_of_slot[sl] = (int)P.k_slot.size();
      P.k_slot.push_back(sl);
    }
  }
  P.n_k = (int)P.k_slot.size();
  if (P.k_slot.empty()) P.k_slot.push_back(0);
  auto put_coef = [](int* r, double c) {
    uint64_t bits;
    std::memcpy(&bits, &c, sizeof(bits));
    r[0] = (int)(uint32_t)(bits & 0xFFFFFFFFu);
    r[1] = (int)(uint32_t)(bits >> 32);
  };
  // f: species into bins, longest first, each to the least-loaded bin
  P.f_count.assign(nv, 0);
  P.f_zero.clear();
  std::vector<int> sp(nv);
  for (int i = 0; i < nv; ++i) {
    sp[i] = i;
    P.f_count[i] = P.f_ptr[i + 1] - P.f_ptr[i];
    if (P.f_count[i] == 0) P.f_zero.push_back(i);
  }
  std::stable_sort(sp.begin(), sp.end(),
                   [&](int a, int b) { return P.f_count[a] > P.f_count[b]; });
  std::vector<std::vector<int>> fbin(G);
  std::vector<int> load(G, 0);
  for (int i : sp) {
    if (P.f_count[i] == 0) continue;
    int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    fbin[b].push_back(i);
    load[b] += P.f_count[i];
  }
  // each bin walks its species in index order
  for (auto& b : fbin) std::sort(b.begin(), b.end());
  // rows: the longest bin, rounded up to a multiple of 4, plus 4 rows that
  // the kernels' look-ahead loads but never processes
  P.g_f_rows[w] = (std::max(1, *std::max_element(load.begin(), load.end())) + 3) / 4 * 4;
  const int f_alloc = P.g_f_rows[w] + 4;
  std::vector<int>& gf = P.gf_rec[w];
  gf.assign((size_t)8 * G * f_alloc, 0);
  for (int j = 0; j < G; ++j) {
    int r = 0;
    for (int i : fbin[j])
      for (int q = P.f_ptr[i]; q < P.f_ptr[i + 1]; ++q, ++r) {
        int* d = &gf[(size_t)8 * ((size_t)r * G + j)];
        int t = P.f_term[q];
        d[0] = i | ((q + 1 == P.f_ptr[i + 1]) ? (1 << 30) : 0);
        d[1] = P.f_guard[q];
        put_coef(&d[2], P.f_coef[q]);
        d[4] = kc_of_slot[P.t_kidx[t]] | (P.t_flags[t] << 28);
        for (int m = 0; m < 3; ++m) d[5 + m] = m < P.t_nreact[t] ? P.t_react[3 * t + m] : -1;
      }
    for (; r < f_alloc; ++r) {
      int* d = &gf[(size_t)8 * ((size_t)r * G + j)];
      d[0] = nv | (1 << 30);
      d[1] = -1;
      put_coef(&d[2], 0.0);
      d[4] = 0;
      d[5] = d[6] = d[7] = -1;
    }
  }
  build_j_runs(P, G, &kc_of_slot, P.gj_rec[w], P.g_j_rows[w]);
}

// The per-cell kernels' J_rxn assembly order for multi-phase programs:
// reaction elements grouped by their record signature (the number of
// mass-action records, then the kinds of the special records in order), so
// that a row of 64 lanes runs one code path.
static void build_mp_j_order(Program& P) {
  const int rn = P.rnnz;
  std::vector<std::vector<int>> sig(rn);
  for (int e = 0; e < rn; ++e) {
    sig[e].push_back(P.j_ptr[e + 1] - P.j_ptr[e]);
    for (int q = P.js_ptr[e]; q < P.js_ptr[e + 1]; ++q) sig[e].push_back(P.js_a[4 * q]);
  }
  std::vector<int> order(rn);
  for (int e = 0; e < rn; ++e) order[e] = e;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sig[x] < sig[y]; });
  P.mj_rows = (rn + 63) / 64;
  P.mj_elist.assign((size_t)64 * std::max(P.mj_rows, 1), -1);
  for (int i = 0; i < rn; ++i) P.mj_elist[i] = order[i];
  // the rows' record streams
  auto put = [&](int* d, int k0, int k1, int k2, int k3, int k4, int k5, double c) {
    uint64_t bits;
    std::memcpy(&bits, &c, sizeof(bits));
    d[0] = k0; d[1] = k1; d[2] = k2; d[3] = k3; d[4] = k4; d[5] = k5;
    d[6] = (int)(uint32_t)(bits & 0xFFFFFFFFu);
    d[7] = (int)(uint32_t)(bits >> 32);
  };
  P.mj_rec.clear();
  P.mj_rhdr.clear();
  int rec_rows = 0;
  for (int r = 0; r < P.mj_rows; ++r) {
    int len = 1;
    for (int l = 0; l < 64; ++l) {
      int e = P.mj_elist[(size_t)64 * r + l];
      if (e >= 0)
        len = std::max(len, (P.j_ptr[e + 1] - P.j_ptr[e]) + (P.js_ptr[e + 1] - P.js_ptr[e]));
    }
    P.mj_rhdr.push_back(rec_rows);
    P.mj_rhdr.push_back(len);
    P.mj_rec.resize((size_t)8 * 64 * (rec_rows + len), 0);  // zeros: padding records
    for (int l = 0; l < 64; ++l) {
      int e = P.mj_elist[(size_t)64 * r + l];
      if (e < 0) continue;
      int k = 0;
      for (int q = P.j_ptr[e]; q < P.j_ptr[e + 1]; ++q, ++k)
        put(&P.mj_rec[(size_t)8 * (64 * (size_t)(rec_rows + k) + l)], -1, P.je_a[4 * q],
            P.je_a[4 * q + 1], P.je_a[4 * q + 2], P.je_a[4 * q + 3], P.je_b[q], P.je_c[q]);
      for (int q = P.js_ptr[e]; q < P.js_ptr[e + 1]; ++q, ++k)
        put(&P.mj_rec[(size_t)8 * (64 * (size_t)(rec_rows + k) + l)], P.js_a[4 * q],
            P.js_a[4 * q + 1], P.js_a[4 * q + 2], P.js_a[4 * q + 3], 0, 0, P.js_c[q]);
    }
    rec_rows += len;
  }
  P.mj_rec.resize((size_t)8 * 64 * (rec_rows + 1), 0);  // the look-ahead's padding row
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
    } else if (t == kRxnAqueousEquilibrium) {
      // rxn_aqueous_equilibrium_get_used_jac_elem (:80-126)
      AqEqView v{id + 1, rflt(r)};
      int nr = v.nr(), np = v.np();
      for (int p = 0; p < v.nph(); p++) {
        for (int ri = p * nr; ri < (p + 1) * nr; ri++) {
          for (int rd = p * nr; rd < (p + 1) * nr; rd++) rb.add(v.react(rd), v.react(ri));
          for (int pd = p * np; pd < (p + 1) * np; pd++) rb.add(v.prod(pd), v.react(ri));
        }
        for (int pi = p * np; pi < (p + 1) * np; pi++) {
          for (int rd = p * nr; rd < (p + 1) * nr; rd++) rb.add(v.react(rd), v.prod(pi));
          for (int pd = p * np; pd < (p + 1) * np; pd++) rb.add(v.prod(pd), v.prod(pi));
        }
        for (int rd = p * nr; rd < (p + 1) * nr; rd++) rb.add(v.react(rd), v.water(p));
        for (int pd = p * np; pd < (p + 1) * np; pd++) rb.add(v.prod(pd), v.water(p));
        if (v.act(p) < 0) continue;
        for (int rd = p * nr; rd < (p + 1) * nr; ++rd) rb.add(v.react(rd), v.act(p));
        for (int pd = p * np; pd < (p + 1) * np; ++pd) rb.add(v.prod(pd), v.act(p));
      }
    } else if (t == kRxnSimpolPhaseTransfer) {
      // rxn_SIMPOL_phase_transfer_get_used_jac_elem (:86-149)
      SimpolView v{id + 1, rflt(r)};
      std::vector<char> flags(ns);
      rb.add(v.gas(), v.gas());
      for (int i = 0; i < v.n(); i++) {
        rb.add(v.aero(i), v.gas());
        rb.add(v.gas(), v.aero(i));
        rb.add(v.aero(i), v.aero(i));
        if (v.act(i) > 0) {
          rb.add(v.gas(), v.act(i));
          rb.add(v.aero(i), v.act(i));
        }
        std::fill(flags.begin(), flags.end(), 0);
        if (v.rep(i) < 0 || v.rep(i) >= n_rep || rep_int[rep_int_idx[v.rep(i)]] !=
                                                      kAeroRepModalBinnedMass)
          return "SIMPOL phase transfer needs a modal/binned-mass aerosol representation";
        int n_elem = rep_used_jac_elem(*this, v.rep(i), v.phase(i), flags);
        if (n_elem > v.njac(i))
          return "SIMPOL phase transfer: more Jacobian elements than expected";
        int used = 0;
        for (int k = 0; k < ns; ++k)
          if (flags[k]) {
            rb.add(v.gas(), k);
            rb.add(v.aero(i), k);
            v.phase_jac(i, 0, used) = k;
            v.phase_jac(i, 1, used) = k;
            ++used;
          }
        for (; used < v.njac(i); ++used) v.phase_jac(i, 0, used) = v.phase_jac(i, 1, used) = -1;
      }
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
    } else if (t == kRxnAqueousEquilibrium) {
      // rxn_aqueous_equilibrium_update_ids (:138-230)
      AqEqView v{id + 1, rflt(r)};
      int nr = v.nr(), np = v.np();
      for (int p = 0, k = 0; p < v.nph(); p++) {
        for (int q = 0; q < nr; q++) v.deriv_id(k++) = deriv_ids[v.react(p * nr + q)];
        for (int q = 0; q < np; q++) v.deriv_id(k++) = deriv_ids[v.prod(p * np + q)];
      }
      for (int p = 0, j = 0; p < v.nph(); p++) {
        for (int ri = p * nr; ri < (p + 1) * nr; ri++) {
          for (int rd = p * nr; rd < (p + 1) * nr; rd++) v.jac_id(j++) = rjac.find(v.react(rd), v.react(ri));
          for (int pd = p * np; pd < (p + 1) * np; pd++) v.jac_id(j++) = rjac.find(v.prod(pd), v.react(ri));
        }
        for (int pi = p * np; pi < (p + 1) * np; pi++) {
          for (int rd = p * nr; rd < (p + 1) * nr; rd++) v.jac_id(j++) = rjac.find(v.react(rd), v.prod(pi));
          for (int pd = p * np; pd < (p + 1) * np; pd++) v.jac_id(j++) = rjac.find(v.prod(pd), v.prod(pi));
        }
        for (int rd = p * nr; rd < (p + 1) * nr; rd++) v.jac_id(j++) = rjac.find(v.react(rd), v.water(p));
        for (int pd = p * np; pd < (p + 1) * np; pd++) v.jac_id(j++) = rjac.find(v.prod(pd), v.water(p));
        for (int rd = p * nr; rd < (p + 1) * nr; ++rd)
          v.jac_id(j++) = v.act(p) < 0 ? -1 : rjac.find(v.react(rd), v.act(p));
        for (int pd = p * np; pd < (p + 1) * np; ++pd)
          v.jac_id(j++) = v.act(p) < 0 ? -1 : rjac.find(v.prod(pd), v.act(p));
      }
      for (int p = 0; p < v.nph(); p++) {
        v.small_conc(p) = 99999.0;
        for (int q = 0; q < nr; q++)
          if (v.small_conc(p) > abs_tol[v.react(p * nr + q)] / 100.0)
            v.small_conc(p) = abs_tol[v.react(p * nr + q)] / 100.0;
        for (int q = 0; q < np; q++)
          if (v.small_conc(p) > abs_tol[v.prod(p * np + q)] / 100.0)
            v.small_conc(p) = abs_tol[v.prod(p * np + q)] / 100.0;
      }
      for (int p = 0; p < v.nph(); p++) v.small_water(p) = abs_tol[v.water(p)] / 10.0;
    } else if (t == kRxnSimpolPhaseTransfer) {
      // rxn_SIMPOL_phase_transfer_update_ids (:160-207)
      SimpolView v{id + 1, rflt(r)};
      v.deriv_id(0) = deriv_ids[v.gas()];
      for (int i = 0; i < v.n(); i++) v.deriv_id(i + 1) = deriv_ids[v.aero(i)];
      int j = 0;
      v.jac_id(j++) = rjac.find(v.gas(), v.gas());
      for (int i = 0; i < v.n(); i++) {
        v.jac_id(j++) = rjac.find(v.aero(i), v.gas());
        v.jac_id(j++) = rjac.find(v.gas(), v.aero(i));
        v.jac_id(j++) = rjac.find(v.aero(i), v.aero(i));
        v.gas_act_jac(i) = v.act(i) > 0 ? rjac.find(v.gas(), v.act(i)) : -1;
        v.aero_act_jac(i) = v.act(i) > 0 ? rjac.find(v.aero(i), v.act(i)) : -1;
        for (int k = 0; k < v.njac(i); ++k) {
          if (v.phase_jac(i, 0, k) > 0) v.phase_jac(i, 0, k) = rjac.find(v.gas(), v.phase_jac(i, 0, k));
          if (v.phase_jac(i, 1, k) > 0)
            v.phase_jac(i, 1, k) = rjac.find(v.aero(i), v.phase_jac(i, 1, k));
        }
      }
    }
  }

  // --- parameter (sub-model) pattern: the dummy (0,0) element, then the
  // sub models' elements (camp_solver.c:1509-1552) --------------------------
  PatternBuilder pb(ns);
  pb.add(0, 0);
  for (int k = 0; k < n_sub; ++k) {
    int* id = &sub_int[sub_int_idx[k]];
    if (id[0] != kSubModelZsrAerosolWater) return "unsupported sub model type";
    ZsrView z{id + 1, &sub_flt[sub_flt_idx[k]]};
    for (int p = 0; p < z.nph(); ++p) {
      pb.add(z.phase(p), z.gas_water());
      for (int x = 0; x < z.npair(); ++x) {
        if (z.type(x) != kZsrJacobson) return "ZSR ion pairs other than JACOBSON are not supported";
        pb.add(z.phase(p), z.phase(p) + z.cation(x));
        pb.add(z.phase(p), z.phase(p) + z.anion(x));
      }
    }
  }
  pjac = pb.build();
  for (int k = 0; k < n_sub; ++k) {
    ZsrView z{&sub_int[sub_int_idx[k]] + 1, &sub_flt[sub_flt_idx[k]]};
    for (int p = 0; p < z.nph(); ++p)
      for (int x = 0; x < z.npair(); ++x) {
        z.gw_jac(p, x) = pjac.find(z.phase(p), z.gas_water());
        z.cation_jac(p, x) = pjac.find(z.phase(p), z.phase(p) + z.cation(x));
        z.anion_jac(p, x) = pjac.find(z.phase(p), z.phase(p) + z.anion(x));
      }
  }

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
  bool chained = false;
  for (int p : map_param) chained |= p != 0;
  if (chained && !multiphase()) return "sub-model Jacobian parameters without sub models";

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
  // Jacobian gather space: solver elements directly when the map is the
  // identity (gas phase), reaction-pattern elements otherwise (the solver
  // Jacobian is then J_rxn * J_params over the map, camp_solver.c:1100-1106)
  const bool mp = multiphase();
  const int jn = mp ? rjac.nnz() : P.nnz;
  std::vector<int> rxn_to_solver(rjac.nnz(), -1);
  if (mp) {
    for (int e = 0; e < rjac.nnz(); ++e) rxn_to_solver[e] = e;
  } else {
    for (size_t m = 0; m < map_rxn.size(); ++m) rxn_to_solver[map_rxn[m]] = map_solver[m];
  }
  // multi-phase instances: slots of f contributions per solver variable
  std::vector<std::vector<int>> f_slots(nv);
  if (mp) {
    std::string err = compile_multiphase(f_slots);
    if (!err.empty()) return err;
  }

  // rate terms: one per reaction
  P.n_terms = n_rxn;
  P.t_kidx.resize(n_rxn);
  P.t_react.assign(3 * n_rxn, -1);
  P.t_nreact.resize(n_rxn);
  P.t_flags.resize(n_rxn);
  std::vector<std::vector<int>> f_term(nv), f_guard(nv), j_term(jn), j_skip(jn), j_guard(jn),
      j_rind(jn);
  std::vector<std::vector<double>> f_coef(nv), j_coef(jn);
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
  // slot records (term -1 - slot): contributions of the multi-phase types
  for (int i = 0; i < nv; ++i)
    for (int sl : f_slots[i]) {
      f_term[i].push_back(-1 - sl);
      f_coef[i].push_back(1.0);
      f_guard[i].push_back(-1);
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
  for (int e = 0; e < jn; ++e) {
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
    if (t < 0) {  // slot record
      P.fe_c.push_back(1.0);
      P.fe_tb.push_back(-1 - t);
      P.fe_tb.push_back(0xFFFFFF | (kTermSlot << 24));
      continue;
    }
    P.fe_tb.push_back(t);
    P.fe_tb.push_back((P.f_guard[q] & 0xFFFFFF) | (P.t_flags[t] << 24));
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
    P.fe_c.push_back(0.0);
    P.fe_tb.assign(2, -1);
  }
  if (P.je_c.empty()) {
    P.je_a.assign(4, -1); P.je_b.push_back(-1); P.je_c.push_back(0.0);
  }
  if (P.f_term.empty()) { P.f_term.push_back(0); P.f_coef.push_back(0); P.f_guard.push_back(-1); }
  if (P.j_term.empty()) {
    P.j_term.push_back(0); P.j_skip.push_back(0); P.j_coef.push_back(0);
    P.j_guard.push_back(-1); P.j_rind.push_back(0);
  }

  P.t_rec.assign(4 * (size_t)std::max(P.n_terms, 1), -1);
  for (int t = 0; t < P.n_terms; ++t) {
    P.t_rec[4 * t] = P.t_kidx[t] | (P.t_flags[t] << 28);
    for (int j = 0; j < 3; ++j) P.t_rec[4 * t + 1 + j] = j < P.t_nreact[t] ? P.t_react[3 * t + j] : -1;
  }
  P.fe_rec.resize(2 * P.fe_tb.size());
  for (size_t q = 0; q < P.fe_c.size(); ++q) {
    uint64_t bits;
    std::memcpy(&bits, &P.fe_c[q], sizeof(bits));
    P.fe_rec[4 * q] = P.fe_tb[2 * q];
    P.fe_rec[4 * q + 1] = P.fe_tb[2 * q + 1];
    P.fe_rec[4 * q + 2] = (int)(uint32_t)(bits & 0xFFFFFFFFu);
    P.fe_rec[4 * q + 3] = (int)(uint32_t)(bits >> 32);
  }

  if (!P.mp) {
    for (int w = 0; w < 3; ++w) build_group_program(P, w);
    build_j_runs(P, 64, nullptr, P.cj_rec, P.cj_rows);
  } else {
    build_mp_j_order(P);
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
  {
    const int W = 64;
    std::vector<int> order(nv);
    for (int i = 0; i < nv; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return P.r_ptr[a + 1] - P.r_ptr[a] > P.r_ptr[b + 1] - P.r_ptr[b];
    });
    P.fs_species = order;
    P.fl_hdr.clear();
    P.re_ell.clear();
    int npass = (nv + W - 1) / W;
    for (int pass = 0; pass < npass; ++pass) {
      int cr = 0;
      for (int l = 0; l < W && pass * W + l < nv; ++l) {
        int i = order[pass * W + l];
        cr = std::max(cr, P.r_ptr[i + 1] - P.r_ptr[i]);
      }
      P.fl_hdr.push_back(0);
      P.fl_hdr.push_back(0);
      P.fl_hdr.push_back((int)(P.re_ell.size() / 2));
      P.fl_hdr.push_back(cr);
      // rows padded (with -1 entries) to a multiple of kEstBatch plus
      // kEstBatch, so that the estimate's batches and its loads a batch ahead
      // need no checks
      int cr_alloc = (cr + kEstBatch - 1) / kEstBatch * kEstBatch + kEstBatch;
      for (int j = 0; j < cr_alloc; ++j)
        for (int l = 0; l < W; ++l) {
          int slot = pass * W + l, q = -1;
          if (slot < nv) {
            int i = order[slot];
            if (j < P.r_ptr[i + 1] - P.r_ptr[i]) q = P.r_ptr[i] + j;
          }
          P.re_ell.push_back(q >= 0 ? P.r_elem[q] : -1);
          P.re_ell.push_back(q >= 0 ? P.r_col[q] : -1);
        }
    }
    if (P.re_ell.empty()) P.re_ell.assign(2, -1);
  }
  {
    // balanced bins of the f gather
    const int W = 64;
    int total = P.f_ptr[nv];
    if (P.n_terms >= 65536 || P.n_slots >= 65536) return "more than 65535 f terms or slots";
    struct Piece { int species, first, count, part; };
    // Pieces: species longer than L split into near-equal pieces, placed
    // longest first into the fullest bin that holds them; a piece that no
    // bin holds fills the emptiest bin and its rest goes on.  L grows until
    // the pieces fit with at most kMaxParts partials (LDS of the gather).
    const int kMaxParts = 64;
    std::vector<std::vector<Piece>> bins;
    int L = std::max(1, (total + W - 1) / W);
    for (;; ++L) {
      std::vector<Piece> pieces;
      for (int i = 0; i < nv; ++i) {
        int cnt = P.f_ptr[i + 1] - P.f_ptr[i];
        if (cnt == 0) continue;
        int np = (cnt + L - 1) / L, at = P.f_ptr[i];
        for (int k = 0; k < np; ++k) {
          int c = cnt / np + (k < cnt % np ? 1 : 0);
          pieces.push_back({i, at, c, -1});
          at += c;
        }
      }
      std::stable_sort(pieces.begin(), pieces.end(),
                       [](const Piece& a, const Piece& b) { return a.count > b.count; });
      bins.assign(W, {});
      std::vector<int> fill(W, 0);
      bool ok = true;
      for (Piece pc : pieces) {
        while (pc.count > 0) {
          int best = -1, most = -1;
          for (int b = 0; b < W; ++b) {
            int room = L - fill[b];
            if (room >= pc.count && (best < 0 || room < L - fill[best])) best = b;
            if (room > 0 && (most < 0 || room > L - fill[most])) most = b;
          }
          if (best >= 0) {
            bins[best].push_back(pc);
            fill[best] += pc.count;
            break;
          }
          if (most < 0) {
            ok = false;
            break;
          }
          int r = L - fill[most];
          bins[most].push_back({pc.species, pc.first, r, -1});
          fill[most] += r;
          pc.first += r;
          pc.count -= r;
        }
        if (!ok) break;
      }
      if (!ok) continue;
      // partials: the pieces of a species in record order get consecutive ids
      std::vector<std::vector<Piece*>> of(nv);
      for (auto& b : bins)
        for (Piece& pc : b) of[pc.species].push_back(&pc);
      int np_total = 0;
      for (int i = 0; i < nv; ++i) {
        if (of[i].size() < 2) continue;
        std::sort(of[i].begin(), of[i].end(),
                  [](const Piece* a, const Piece* b) { return a->first < b->first; });
        for (Piece* pc : of[i]) pc->part = np_total++;
      }
      if (np_total <= kMaxParts || L >= P.max_f_len) {
        P.fb_nparts = np_total;
        break;
      }
    }
    if (2 * nv + 3 * P.fb_nparts + 3 >= 65536) return "more than 65535 f gather targets";
    P.fb_rows = L;
    // rows padded to a multiple of 4 plus 4 more, so that the gather's
    // batches of 4 and its loads 4 rows ahead need no bounds checks
    const int L_alloc = (L + 3) / 4 * 4 + 4;
    P.fb_rec.assign((size_t)4 * W * L_alloc, 0);
    for (int l = 0; l < W; ++l) {
      int j = 0;
      for (const Piece& pc : bins[l])
        for (int r = 0; r < pc.count; ++r, ++j) {
          const int* src = &P.fe_rec[4 * (size_t)(pc.first + r)];
          int* dst = &P.fb_rec[4 * ((size_t)j * W + l)];
          // the piece's sum goes to s_fb[target]: a whole species' value at
          // [species] (its magnitude sum at [nv + species]), a partial's
          // (hi, lo, magnitudes) at [2 nv + 3 part ..] (flag 32)
          int target = pc.part >= 0 ? 2 * nv + 3 * pc.part : pc.species;
          dst[0] = (src[0] & 0xFFFF) | (target << 16);
          dst[1] = src[1] | (r + 1 == pc.count ? (16 << 24) : 0) | (pc.part >= 0 ? (32 << 24) : 0);
          dst[2] = src[2];
          dst[3] = src[3];
        }
      // padding: no contribution, running sums stored to a spare partial
      for (; j < L_alloc; ++j) {
        int* dst = &P.fb_rec[4 * ((size_t)j * W + l)];
        dst[0] = (2 * nv + 3 * P.fb_nparts) << 16;
        dst[1] = ((8 | 32) << 24) | 0xFFFFFF;
        dst[2] = dst[3] = 0;
      }
    }
    P.fb_fin.assign(2 * (size_t)std::max(nv, 1), 0);
    for (int i = 0; i < nv; ++i) {
      int cnt = P.f_ptr[i + 1] - P.f_ptr[i];
      P.fb_fin[2 * i] = 0;
      P.fb_fin[2 * i + 1] = cnt == 0 ? -1 : 0;
    }
    for (const auto& b : bins)
      for (const Piece& pc : b)
        if (pc.part >= 0) {
          int* f = &P.fb_fin[2 * pc.species];
          if (f[1] == 0) f[0] = pc.part;
          f[0] = std::min(f[0], pc.part);
          f[1] += 1;
        }
    // the packing must cover every species' records exactly once, in pieces
    // of consecutive records, with consecutive partial ids in record order
    std::vector<std::vector<const Piece*>> by(nv);
    for (const auto& b : bins) {
      int fill = 0;
      for (const Piece& pc : b) {
        by[pc.species].push_back(&pc);
        fill += pc.count;
      }
      if (fill > L) return "f gather packing: a bin exceeds its rows";
    }
    for (int i = 0; i < nv; ++i) {
      std::sort(by[i].begin(), by[i].end(),
                [](const Piece* a, const Piece* b) { return a->first < b->first; });
      int at = P.f_ptr[i];
      for (size_t k = 0; k < by[i].size(); ++k) {
        const Piece* pc = by[i][k];
        bool ok = pc->first == at && pc->count > 0 &&
                  (by[i].size() == 1 ? pc->part < 0 : pc->part == P.fb_fin[2 * i] + (int)k);
        if (!ok) return "f gather packing: species records not covered exactly once";
        at += pc->count;
      }
      if (at != P.f_ptr[i + 1]) return "f gather packing: species records not covered";
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
  // the column solve's padded columns: {start, start + entries} per pivot
  P.cs_lpos.clear(); P.cs_lrow.clear(); P.cs_upos.clear(); P.cs_urow.clear();
  std::vector<int> cl(2 * nv), cu(2 * nv);
  auto pad_cols = [&](const std::vector<int>& ptr, const std::vector<int>& vpos,
                      const std::vector<int>& vrow, std::vector<int>& opos, std::vector<int>& orow,
                      std::vector<int>& rng) {
    for (int k = 0; k < nv; ++k) {
      int n = ptr[k + 1] - ptr[k], start = (int)opos.size();
      rng[2 * k] = start;
      rng[2 * k + 1] = start + n;
      for (int q = 0; q < std::max(1, (n + 63) / 64) * 64; ++q) {
        opos.push_back(q < n ? vpos[ptr[k] + q] : 0);
        orow.push_back(q < n ? vrow[ptr[k] + q] : -1);
      }
    }
    // entries read for the pivots past the end (header zero): the first row
    if (opos.size() < 64) { opos.resize(64, 0); orow.resize(64, -1); }
  };
  pad_cols(P.l_ptr, P.l_pos, P.l_row, P.cs_lpos, P.cs_lrow, cl);
  pad_cols(P.u_ptr, P.u_pos, P.u_row, P.cs_upos, P.cs_urow, cu);
  P.lu_hdr.assign((size_t)8 * (nv + 2), 0);
  for (int k = 0; k < nv; ++k) {
    int* h = &P.lu_hdr[(size_t)8 * k];
    h[0] = cl[2 * k]; h[1] = cl[2 * k + 1]; h[2] = P.up_ptr[k]; h[3] = P.up_ptr[k + 1];
    h[4] = P.diag_pos[k]; h[5] = cu[2 * k]; h[6] = cu[2 * k + 1];
  }
  // The register triangular solve's program (nv <= 128): per block of 8
  // pivots, in the order the device sweeps visit them, and per lane t, the
  // LU positions of the entries (row t, pivot) and (row t + 64, pivot) as
  // pos0 | pos1 << 16 -- the zero slot lu_nnz + 64 where the row has no
  // entry or the pivot is outside the sweep -- two int4 per lane and block.
  P.ts_tbl.clear();
  P.ts_fwd_blocks = P.ts_bwd_blocks = 0;
  if (nv <= 128) {
    if (P.lu_nnz + 65 > 65535) return "LU fill too large for the register solve";
    const int W = 64, zero = P.lu_nnz + 64;
    auto block = [&](int kb, int dir, int lo, int hi) {  // pivots kb + dir*u in [lo, hi)
      for (int l = 0; l < W; ++l)
        for (int u = 0; u < 8; ++u) {
          int k = kb + dir * u, w = 0;
          for (int h = 0; h < 2; ++h) {
            int r = l + h * W, p = zero;
            bool in_sweep = k >= lo && k < hi && r < nv;
            if (in_sweep && (dir > 0 ? r > k : r < k) && F[(size_t)r * nv + k])
              p = pos[(size_t)r * nv + k];
            w |= p << (16 * h);
          }
          P.ts_tbl.push_back(w);
        }
    };
    const int kf = nv - 1, split = std::min(kf, W);
    for (int kb = 0; kb < split; kb += 8, ++P.ts_fwd_blocks) block(kb, 1, 0, split);
    for (int kb = split; kb < kf; kb += 8, ++P.ts_fwd_blocks) block(kb, 1, split, kf);
    for (int kb = nv - 1; kb >= W; kb -= 8, ++P.ts_bwd_blocks) block(kb, -1, W, nv);
    for (int kb = std::min(nv - 1, W - 1); kb >= 0; kb -= 8, ++P.ts_bwd_blocks) block(kb, -1, 0, W);
    P.ts_tbl.insert(P.ts_tbl.end(), (size_t)2 * 8 * W, zero | (zero << 16));  // look-ahead padding
  }
  // The factorisation's program: per pivot, the L positions to scale and the
  // update records {dst, l, u}, each padded to whole rows of 64 (at least
  // one row of L positions) with entries that write lane l's sink slot
  // lu_nnz + l, so the device loops need no per-lane bounds checks; every
  // pivot's updates are padded with rows of sinks to a whole number (at
  // least one) of batches of kLuRowBatch rows, so that the device loads a
  // batch with no per-row selection; lf_hdr[k] = {first L position, first
  // update record, L rows | update rows << 16, diagonal position} (two empty
  // pivots past the end).
  if (P.lu_nnz + 64 > 65535) return "LU fill exceeds 65471 entries per cell";
  P.lf_pos.clear();
  P.up_rec.clear();
  P.lf_hdr.assign((size_t)4 * (nv + 2), 0);
  for (int k = 0; k < nv; ++k) {
    int* h = &P.lf_hdr[(size_t)4 * k];
    h[0] = (int)P.lf_pos.size();
    h[1] = (int)P.up_rec.size();
    int nl = P.l_ptr[k + 1] - P.l_ptr[k], nu = P.up_ptr[k + 1] - P.up_ptr[k];
    int rl = std::max(1, (nl + 63) / 64);
    int ru = std::max(kLuRowBatch, (nu + 63) / 64);
    ru = (ru + kLuRowBatch - 1) / kLuRowBatch * kLuRowBatch;
    for (int q = 0; q < 64 * rl; ++q)
      P.lf_pos.push_back(q < nl ? P.l_pos[P.l_ptr[k] + q] : P.lu_nnz + q % 64);
    for (int q = 0; q < 64 * ru; ++q) {
      if (q < nu) {
        size_t e = (size_t)P.up_ptr[k] + q;
        P.up_rec.push_back((uint64_t)P.up_dst[e] | ((uint64_t)P.up_l[e] << 16) |
                           ((uint64_t)P.up_u[e] << 32));
      } else {
        P.up_rec.push_back((uint64_t)(P.lu_nnz + q % 64));
      }
    }
    h[2] = rl | (ru << 16);
    h[3] = P.diag_pos[k];
  }
  initialized = true;
  return std::string();
}

// ---------------------------------------------------------------------------
// Multi-phase kernel tables: aerosol phase instances, ZSR items, SIMPOL and
// aqueous-equilibrium instances with their f-contribution slots, the special
// Jacobian records of those instances per reaction-pattern element, and the
// J_rxn * J_params map per solver element (camp_solver.c:1100-1106).
// ---------------------------------------------------------------------------
std::string Model::compile_multiphase(std::vector<std::vector<int>>& f_slots) {
  Program& P = prog;
  P.mp = true;
  P.rnnz = rjac.nnz();
  P.pnnz = pjac.nnz();
  P.pm_ptr.assign(1, 0);
  for (int p = 0; p < n_phase; ++p) {
    PhaseView v{&ph_int[ph_int_idx[p]], &ph_flt[ph_flt_idx[p]]};
    for (int i = 0; i < v.n(); ++i) {
      P.pm_counted.push_back(v.counted(i) ? 1 : 0);
      P.pm_mw.push_back(v.mw(i));
      P.pm_dens.push_back(v.density(i));
    }
    P.pm_ptr.push_back((int)P.pm_counted.size());
  }
  // phase instances in the representation's phase-index order
  std::vector<int> rep_pi_base(n_rep, 0);
  std::vector<double> pi_radius;
  for (int r = 0; r < n_rep; ++r) {
    if (rep_int[rep_int_idx[r]] != kAeroRepModalBinnedMass)
      return "unsupported aerosol representation type";
    rep_pi_base[r] = (int)P.pi_start.size();
    BinnedRepView v = rep_view(*this, r);
    for (int s = 0; s < v.nsec(); ++s) {
      if (v.type(s) != kBinned) return "MODAL sections are not supported by the multi-phase kernels";
      if (v.nph(s) != 1) return "bins holding more than one phase are not supported by the multi-phase kernels";
      for (int b = 0; b < v.nbin(s); ++b) {
        P.pi_start.push_back(v.phase_state(s, 0, b));
        P.pi_model.push_back(v.phase_model(s, 0, b));
        P.pi_dp.push_back(v.bin_dp(s, b));
        pi_radius.push_back(v.radius(s, b));
      }
    }
  }
  // ZSR items
  P.zi_pair_ptr.assign(1, 0);
  for (int k = 0; k < n_sub; ++k) {
    ZsrView z{&sub_int[sub_int_idx[k]] + 1, &sub_flt[sub_flt_idx[k]]};
    for (int p = 0; p < z.nph(); ++p) {
      P.zi_water.push_back(z.phase(p));
      P.zi_gw.push_back(z.gas_water());
      P.zi_sub.push_back(k);
      for (int x = 0; x < z.npair(); ++x) {
        if (z.ny(x) > 8) return "ZSR molality polynomials of more than 8 terms";
        P.zp_cat.push_back(z.phase(p) + z.cation(x));
        P.zp_an.push_back(z.phase(p) + z.anion(x));
        P.zp_ncat.push_back(z.n_cation(x));
        P.zp_nan.push_back(z.n_anion(x));
        P.zp_ny.push_back(z.ny(x));
        P.zp_jgw.push_back(z.gw_jac(p, x));
        P.zp_jcat.push_back(z.cation_jac(p, x));
        P.zp_jan.push_back(z.anion_jac(p, x));
        P.zp_catmw.push_back(z.cation_mw(x));
        P.zp_anmw.push_back(z.anion_mw(x));
        P.zp_lowrh.push_back(z.low_rh(x));
        for (int q = 0; q < 8; ++q) P.zp_y.push_back(q < z.ny(x) ? z.y(x, q) : 0.0);
      }
      P.zi_pair_ptr.push_back((int)P.zp_cat.size());
    }
  }
  P.n_sub_env = n_sub;

  std::vector<std::vector<int>> ja(rjac.nnz());
  std::vector<std::vector<double>> jc(rjac.nnz());
  auto addj = [&](int e, int kind, int inst, int p1, int p2, double c) {
    if (e < 0) return;
    ja[e].insert(ja[e].end(), {kind, inst, p1, p2});
    jc[e].push_back(c);
  };
  int slot = 0;
  for (int r = 0; r < n_rxn; ++r) {
    int* id = rint(r);
    int t = id[0];
    if (t == kRxnSimpolPhaseTransfer) {
      SimpolView v{id + 1, rflt(r)};
      for (int i = 0; i < v.n(); ++i) {
        if (v.act(i) > -1) return "SIMPOL activity coefficients are not supported by the multi-phase kernels";
        if (v.rep(i) < 0 || v.rep(i) >= n_rep) return "SIMPOL: bad aerosol representation id";
        int inst = (int)P.sp_gas.size();
        int pi = rep_pi_base[v.rep(i)] + v.phase(i);
        P.sp_gas.push_back(v.gas());
        P.sp_aero.push_back(v.aero(i));
        P.sp_dgas.push_back(v.deriv_id(0));
        P.sp_daero.push_back(v.deriv_id(1 + i));
        P.sp_pi.push_back(pi);
        P.sp_env.push_back(rxn_env_idx[r]);
        P.sp_flt.push_back(rxn_float_idx[r]);
        P.sp_radius.push_back(pi_radius[pi]);
        P.sp_slot.push_back(slot);
        // rxn_SIMPOL_phase_transfer.c:355-399 (total particle mass): gas
        // +N evap, -N cond; aerosol -N evap/kp, +N cond/kp
        if (v.deriv_id(0) >= 0) {
          f_slots[v.deriv_id(0)].push_back(slot);
          f_slots[v.deriv_id(0)].push_back(slot + 1);
        }
        if (v.deriv_id(1 + i) >= 0) {
          f_slots[v.deriv_id(1 + i)].push_back(slot + 2);
          f_slots[v.deriv_id(1 + i)].push_back(slot + 3);
        }
        slot += 4;
        addj(v.jac_id(0), kJsSimpolGG, inst, 0, 0, 0.0);
        addj(v.jac_id(1 + 3 * i), kJsSimpolAG, inst, 0, 0, 0.0);
        addj(v.jac_id(1 + 3 * i + 1), kJsSimpolGA, inst, 0, 0, 0.0);
        addj(v.jac_id(1 + 3 * i + 2), kJsSimpolAA, inst, 0, 0, 0.0);
        // phase element k = k-th counted species of the bin's phase
        int pm = P.pi_model[pi], k = 0;
        for (int q = P.pm_ptr[pm]; q < P.pm_ptr[pm + 1]; ++q) {
          if (!P.pm_counted[q]) continue;
          if (k < v.njac(i)) {
            if (v.phase_jac(i, 0, k) > 0) addj(v.phase_jac(i, 0, k), kJsSimpolGP, inst, q, 0, 0.0);
            if (v.phase_jac(i, 1, k) > 0) addj(v.phase_jac(i, 1, k), kJsSimpolAP, inst, q, 0, 0.0);
          }
          ++k;
        }
      }
    } else if (t == kRxnAqueousEquilibrium) {
      AqEqView v{id + 1, rflt(r)};
      int nr = v.nr(), np = v.np();
      if (nr > 3 || np > 3) return "aqueous equilibria with more than 3 reactants or products";
      for (int p = 0; p < v.nph(); ++p) {
        if (v.act(p) >= 0) return "aqueous-equilibrium activity coefficients are not supported by the multi-phase kernels";
        int inst = (int)P.aq_nr.size();
        P.aq_nr.push_back(nr);
        P.aq_np.push_back(np);
        for (int q = 0; q < 3; ++q) P.aq_sp.push_back(q < nr ? v.react(p * nr + q) : -1);
        for (int q = 0; q < 3; ++q) P.aq_sp.push_back(q < np ? v.prod(p * np + q) : -1);
        P.aq_water.push_back(v.water(p));
        P.aq_env.push_back(rxn_env_idx[r]);
        P.aq_flt.push_back(rxn_float_idx[r]);
        P.aq_small_water.push_back(v.small_water(p));
        P.aq_slot.push_back(slot);
        int d0 = p * (nr + np);
        for (int q = 0; q < 6; ++q) P.aq_deriv.push_back(q < nr + np ? v.deriv_id(d0 + q) : -1);
        // rxn_aqueous_equilibrium.c:362-386: -rf/m, +rr/m per reactant;
        // +rf/m, -rr/m per product
        for (int q = 0; q < nr + np; ++q)
          if (v.deriv_id(d0 + q) >= 0) {
            f_slots[v.deriv_id(d0 + q)].push_back(slot + 2 * q);
            f_slots[v.deriv_id(d0 + q)].push_back(slot + 2 * q + 1);
          }
        slot += 2 * (nr + np);
        // Jacobian ids in the reference's order (rxn_aqueous_equilibrium.c:418-545)
        int j = p * (nr + np) * (nr + np + 2);
        for (int ri = 0; ri < nr; ++ri) {
          for (int rd = 0; rd < nr; ++rd) addj(v.jac_id(j++), kJsAqRR, inst, ri, rd, v.mftm(rd));
          for (int pd = 0; pd < np; ++pd)
            addj(v.jac_id(j++), kJsAqRP, inst, ri, nr + pd, v.mftm(nr + pd));
        }
        for (int pj = 0; pj < np; ++pj) {
          for (int rd = 0; rd < nr; ++rd) addj(v.jac_id(j++), kJsAqPR, inst, 3 + pj, rd, v.mftm(rd));
          for (int pd = 0; pd < np; ++pd)
            addj(v.jac_id(j++), kJsAqPP, inst, 3 + pj, nr + pd, v.mftm(nr + pd));
        }
        for (int rd = 0; rd < nr; ++rd) addj(v.jac_id(j++), kJsAqRW, inst, 0, rd, v.mftm(rd));
        for (int pd = 0; pd < np; ++pd)
          addj(v.jac_id(j++), kJsAqPW, inst, 0, nr + pd, v.mftm(nr + pd));
      }
    }
  }
  P.n_slots = slot;
  P.n_pi = (int)P.pi_start.size();
  P.n_zi = (int)P.zi_water.size();
  P.n_sp = (int)P.sp_gas.size();
  P.n_aq = (int)P.aq_nr.size();
  P.js_ptr.assign(1, 0);
  for (int e = 0; e < rjac.nnz(); ++e) {
    P.js_a.insert(P.js_a.end(), ja[e].begin(), ja[e].end());
    P.js_c.insert(P.js_c.end(), jc[e].begin(), jc[e].end());
    P.js_ptr.push_back((int)P.js_c.size());
  }
  if (P.js_c.empty()) {
    P.js_a.assign(4, 0);
    P.js_c.push_back(0.0);
  }
  // solver element <- (reaction element, parameter element) in map order
  std::vector<std::vector<int>> mr(P.nnz), mpar(P.nnz);
  for (size_t m = 0; m < map_solver.size(); ++m) {
    mr[map_solver[m]].push_back(map_rxn[m]);
    mpar[map_solver[m]].push_back(map_param[m]);
  }
  P.jm_ptr.assign(1, 0);
  for (int e = 0; e < P.nnz; ++e) {
    P.jm_rxn.insert(P.jm_rxn.end(), mr[e].begin(), mr[e].end());
    P.jm_param.insert(P.jm_param.end(), mpar[e].begin(), mpar[e].end());
    P.jm_ptr.push_back((int)P.jm_rxn.size());
  }
  for (auto* v : {&P.pi_start, &P.pi_model, &P.zi_water, &P.zi_gw, &P.zi_sub, &P.zp_cat, &P.zp_an,
                  &P.zp_ncat, &P.zp_nan, &P.zp_ny, &P.zp_jgw, &P.zp_jcat, &P.zp_jan, &P.sp_gas,
                  &P.sp_aero, &P.sp_dgas, &P.sp_daero, &P.sp_pi, &P.sp_env, &P.sp_flt, &P.sp_slot,
                  &P.aq_nr, &P.aq_np, &P.aq_sp, &P.aq_water, &P.aq_env, &P.aq_flt, &P.aq_deriv,
                  &P.aq_slot, &P.jm_rxn, &P.jm_param, &P.pm_counted})
    if (v->empty()) v->push_back(0);
  for (auto* v : {&P.pi_dp, &P.pm_mw, &P.pm_dens, &P.zp_catmw, &P.zp_anmw, &P.zp_lowrh, &P.zp_y,
                  &P.sp_radius, &P.aq_small_water})
    if (v->empty()) v->push_back(0.0);
  return std::string();
}

}  // namespace camp
