// C ABI of the MI355X-native CAMP solver and its host-side drivers.
//
//  * solver_new / rxn_add_condensed_data / solver_initialize build the host
//    Model (camp_model.cpp) and upload its kernel program once;
//  * solver_run moves the host state to HBM and runs either
//      - the per-cell path (default): the persistent k_cell_bdf kernel,
//        every cell with its own BDF control, or
//      - the block path (the reference's semantics, camp_solver.c:544-693):
//        one BDF over all cells driven from the host, one kernel launch per
//        f(y), J(y), LU factorisation and triangular solve;
//  * camp_solver_run_device runs on arrays already resident in HBM.
//
// Error convention follows the reference: setup errors print a message and
// exit(EXIT_FAILURE) (camp_solver.c:93-96); solver_run returns
// CAMP_SOLVER_SUCCESS (0) or CAMP_SOLVER_FAIL (1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "camp_solver_impl.hpp"

namespace camp {

static inline int nblk(long n, int b = 256) { return (int)((n + b - 1) / b); }

void Solver::setup_device() {
  const Program& G = m.prog;
  HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&ev0));
  HIP_OK(hipEventCreate(&ev1));
  P.ns = G.ns;
  P.nv = G.nv;
  P.nnz = G.nnz;
  P.n_env = G.n_env;
  P.n_terms = G.n_terms;
  P.lu_nnz = G.lu_nnz;
  P.n_rxn = m.n_rxn;
  P.dep_state = dev_upload(G.dep_state, owned);
  P.j_ptr = dev_upload(G.j_ptr, owned);
  P.colptr = dev_upload(G.colptr, owned);
  P.rowval = dev_upload(G.rowval, owned);
  P.n_k = G.n_k;
  for (int w = 0; w < 3; ++w) {
    P.g_f_rows[w] = G.g_f_rows[w];
    P.g_j_rows[w] = G.g_j_rows[w];
    P.gf_rec[w] = G.gf_rec[w].empty()
                      ? nullptr
                      : reinterpret_cast<const int4*>(dev_upload(G.gf_rec[w], owned));
    P.gj_rec[w] = G.gj_rec[w].empty()
                      ? nullptr
                      : reinterpret_cast<const int4*>(dev_upload(G.gj_rec[w], owned));
  }
  P.cj_rows = G.cj_rows;
  P.cj_rec = G.cj_rec.empty() ? nullptr
                              : reinterpret_cast<const int4*>(dev_upload(G.cj_rec, owned));
  P.k_slot = dev_upload(G.k_slot, owned);
  P.f_count = dev_upload(G.f_count, owned);
  P.n_fzero = (int)G.f_zero.size();
  P.f_zero = dev_upload(G.f_zero, owned);
  P.r_ptr = dev_upload(G.r_ptr, owned);
  P.r_elem = dev_upload(G.r_elem, owned);
  P.r_col = dev_upload(G.r_col, owned);
  {
    std::vector<char> dep(G.ns, 0);
    for (int d : G.dep_state) dep[d] = 1;
    std::vector<int> cs;
    for (int i = 0; i < G.ns; ++i)
      if (!dep[i]) cs.push_back(i);
    P.n_cs = (int)cs.size();
    P.cs_ids = dev_upload(cs, owned);
  }
  P.perm = dev_upload(G.perm, owned);
  P.m_to_lu = dev_upload(G.m_to_lu, owned);
  P.diag_pos = dev_upload(G.diag_pos, owned);
  P.e_diag = dev_upload(G.e_diag, owned);
  // (the column solve's padded copies; the host keeps the plain columns)
  P.l_pos = dev_upload(G.cs_lpos, owned);
  P.l_row = dev_upload(G.cs_lrow, owned);
  P.u_pos = dev_upload(G.cs_upos, owned);
  P.u_row = dev_upload(G.cs_urow, owned);
  P.lu_hdr = reinterpret_cast<const int4*>(dev_upload(G.lu_hdr, owned));
  P.up_rec = dev_upload(G.up_rec, owned);
  P.lf_pos = dev_upload(G.lf_pos, owned);
  P.lf_hdr = reinterpret_cast<const int4*>(dev_upload(G.lf_hdr, owned));
  P.fs_species = dev_upload(G.fs_species, owned);
  P.t_rec = reinterpret_cast<const int4*>(dev_upload(G.t_rec, owned));
  P.fl_hdr = reinterpret_cast<const int4*>(dev_upload(G.fl_hdr, owned));
  P.re_ell = reinterpret_cast<const int2*>(dev_upload(G.re_ell, owned));
  P.fb_rows = G.fb_rows;
  P.fb_nparts = G.fb_nparts;
  P.fb_rec = reinterpret_cast<const int4*>(dev_upload(G.fb_rec, owned));
  P.fb_fin = reinterpret_cast<const int2*>(dev_upload(G.fb_fin, owned));
  P.je_a = reinterpret_cast<const int4*>(dev_upload(G.je_a, owned));
  P.je_b = dev_upload(G.je_b, owned);
  P.je_c = dev_upload(G.je_c, owned);
  P.ts_tbl = G.ts_tbl.empty() ? nullptr
                              : reinterpret_cast<const int4*>(dev_upload(G.ts_tbl, owned));
  std::vector<int> rx_type(m.n_rxn), rx_fidx(m.n_rxn), rx_eidx(m.n_rxn), rx_nreact(m.n_rxn);
  for (int r = 0; r < m.n_rxn; ++r) {
    int* id = m.rint(r);
    rx_type[r] = id[0];
    rx_fidx[r] = m.rxn_float_idx[r];
    rx_eidx[r] = m.rxn_env_idx[r];
    rx_nreact[r] = (id[0] == kRxnArrhenius || id[0] == kRxnTroe || id[0] == kRxnCmaqH2O2 ||
                    id[0] == kRxnCmaqOhHno3) ? id[1] : 0;
  }
  P.rx_type = dev_upload(rx_type, owned);
  P.rx_fidx = dev_upload(rx_fidx, owned);
  P.rx_eidx = dev_upload(rx_eidx, owned);
  P.rx_nreact = dev_upload(rx_nreact, owned);
  P.rx_float = dev_upload(m.rxn_float, owned);
  std::vector<double> atol_dep(G.nv);
  for (int d = 0; d < G.nv; ++d) atol_dep[d] = m.abs_tol[G.dep_state[d]];
  P.abs_tol_dep = dev_upload(atol_dep, owned);
  P.mp = G.mp ? 1 : 0;
  if (G.mp) {
    P.rnnz = G.rnnz;
    P.pnnz = G.pnnz;
    P.n_pi = G.n_pi;
    P.n_zi = G.n_zi;
    P.n_sp = G.n_sp;
    P.n_aq = G.n_aq;
    P.n_slots = G.n_slots;
    P.n_sub_env = G.n_sub_env;
    // aero_phase_solver.c:19-24 minimum mass, moles and volume, as
    // double-double renderings of the long double constants
    long double mm = 1.0e-25L, mo = 1.0e-25L / 0.1L, mv = 1.0e-25L / 1800.0L;
    P.min_mass_hi = (double)mm;
    P.min_mass_lo = (double)(mm - (long double)P.min_mass_hi);
    P.min_moles_hi = (double)mo;
    P.min_moles_lo = (double)(mo - (long double)P.min_moles_hi);
    P.min_volume = (double)mv;
    P.pi_start = dev_upload(G.pi_start, owned);
    P.pi_model = dev_upload(G.pi_model, owned);
    P.pi_dp = dev_upload(G.pi_dp, owned);
    P.pm_ptr = dev_upload(G.pm_ptr, owned);
    P.pm_counted = dev_upload(G.pm_counted, owned);
    P.pm_mw = dev_upload(G.pm_mw, owned);
    P.pm_dens = dev_upload(G.pm_dens, owned);
    P.zi_water = dev_upload(G.zi_water, owned);
    P.zi_gw = dev_upload(G.zi_gw, owned);
    P.zi_sub = dev_upload(G.zi_sub, owned);
    P.zi_pair_ptr = dev_upload(G.zi_pair_ptr, owned);
    P.zp_cat = dev_upload(G.zp_cat, owned);
    P.zp_an = dev_upload(G.zp_an, owned);
    P.zp_ncat = dev_upload(G.zp_ncat, owned);
    P.zp_nan = dev_upload(G.zp_nan, owned);
    P.zp_ny = dev_upload(G.zp_ny, owned);
    P.zp_jgw = dev_upload(G.zp_jgw, owned);
    P.zp_jcat = dev_upload(G.zp_jcat, owned);
    P.zp_jan = dev_upload(G.zp_jan, owned);
    P.zp_catmw = dev_upload(G.zp_catmw, owned);
    P.zp_anmw = dev_upload(G.zp_anmw, owned);
    P.zp_lowrh = dev_upload(G.zp_lowrh, owned);
    P.zp_y = dev_upload(G.zp_y, owned);
    P.sp_gas = dev_upload(G.sp_gas, owned);
    P.sp_aero = dev_upload(G.sp_aero, owned);
    P.sp_dgas = dev_upload(G.sp_dgas, owned);
    P.sp_daero = dev_upload(G.sp_daero, owned);
    P.sp_pi = dev_upload(G.sp_pi, owned);
    P.sp_env = dev_upload(G.sp_env, owned);
    P.sp_flt = dev_upload(G.sp_flt, owned);
    P.sp_slot = dev_upload(G.sp_slot, owned);
    P.sp_radius = dev_upload(G.sp_radius, owned);
    P.aq_nr = dev_upload(G.aq_nr, owned);
    P.aq_np = dev_upload(G.aq_np, owned);
    P.aq_sp = dev_upload(G.aq_sp, owned);
    P.aq_water = dev_upload(G.aq_water, owned);
    P.aq_env = dev_upload(G.aq_env, owned);
    P.aq_flt = dev_upload(G.aq_flt, owned);
    P.aq_slot = dev_upload(G.aq_slot, owned);
    P.aq_small_water = dev_upload(G.aq_small_water, owned);
    P.js_ptr = dev_upload(G.js_ptr, owned);
    P.js_a = reinterpret_cast<const int4*>(dev_upload(G.js_a, owned));
    P.js_c = dev_upload(G.js_c, owned);
    P.mj_rows = G.mj_rows;
    P.mj_elist = dev_upload(G.mj_elist, owned);
    P.mj_rec = reinterpret_cast<const int4*>(dev_upload(G.mj_rec, owned));
    P.mj_rhdr = reinterpret_cast<const int2*>(dev_upload(G.mj_rhdr, owned));
    P.jm_ptr = dev_upload(G.jm_ptr, owned);
    P.jm_rxn = dev_upload(G.jm_rxn, owned);
    P.jm_param = dev_upload(G.jm_param, owned);
  }

  size_t nc = m.nc;
  d_state = dev_alloc<double>(nc * G.ns, owned);
  d_env = dev_alloc<double>(nc * 2, owned);
  d_renv = dev_alloc<double>(nc * G.n_env, owned);
  d_Js = dev_alloc<double>(nc * G.nnz, owned);
  d_Jy = dev_alloc<double>(nc * G.nv, owned);
  d_Jf = dev_alloc<double>(nc * G.nv, owned);
  d_stats = dev_alloc<double>(nc * kCellStats, owned);
  d_counter = dev_alloc<int>(1, owned);
  d_fail = dev_alloc<int>(1, owned);
  d_var_type = dev_upload(m.var_type, owned);
  h_stats.assign(nc * kCellStats, 0.0);
  h_order.reserve(nc);
}

void Solver::upload_renv() {
  if (!renv_dirty) return;
  HIP_OK(hipMemcpyAsync(d_renv, m.rxn_env.data(), m.rxn_env.size() * sizeof(double),
                        hipMemcpyHostToDevice, stream));
  renv_dirty = false;
}

void check_jac_device(Solver* s, const double* d_y, const double* d_state, const double* d_env,
                      const double* d_J, double time_step, int* d_fails);

void Solver::refresh_kval() {
  if (P.mp || P.n_k == 0) return;
  hipLaunchKernelGGL(k_kval, dim3(nblk((long)m.nc * P.n_k)), dim3(256), 0, stream, P, m.nc, d_renv,
                     kval());
}

void Solver::launch_deriv(const double* y, const double* st, const double* env, double* out,
                          double h, int use_est, int* fail) {
  int nc = m.nc;
  if (!P.mp) {
    launch_group_deriv(f_lanes, P, nc, y, st, kval(), d_Js, d_Jy, d_Jf, out, h, nullptr, use_est,
                       fail, stream);
    return;
  }
  int wpb = fj_waves_per_block(P);
  size_t lds = wpb * fj_lds_doubles(P) * sizeof(double);
  hipLaunchKernelGGL(k_deriv, dim3(std::min(nblk(nc, wpb), 4096)), dim3(64 * wpb), lds, stream, P,
                     nc, y, st, env, d_renv, d_Js, d_Jy, d_Jf, out, h, use_est, fail);
}

void Solver::launch_jac(const double* y, const double* st, const double* env, double* J, double h,
                        double* jr) {
  int nc = m.nc;
  if (!P.mp) {
    launch_group_jac(j_lanes, P, nc, y, st, kval(), J, h, stream);
    return;
  }
  int wpb = fj_waves_per_block(P);
  size_t lds = wpb * fj_lds_doubles(P) * sizeof(double);
  hipLaunchKernelGGL(k_jac, dim3(std::min(nblk(nc, wpb), 4096)), dim3(64 * wpb), lds, stream, P,
                     nc, y, st, env, d_renv, J, h, jr);
}

int Solver::run(double* st, double* env, double t0, double t1) {
  upload_renv();
  if (m.n_rxn == 0) return CAMP_SOLVER_SUCCESS;  // no_solve, camp_solver.c:240
  if (mode == CAMP_MODE_BLOCK && P.mp)
    die("multi-phase mechanisms run in the per-cell mode (CAMP_MODE_PER_CELL)");
  int rc = mode == CAMP_MODE_BLOCK ? run_block(st, env, t0, t1) : run_cells(st, env, t0, t1);
  if (eval_jac) {
    // solver_set_eval_jac: each cell's last Jacobian (its J history) against
    // difference quotients of f at the state it was evaluated at; a cell
    // that fails counts as one Jacobian evaluation failure
    if (!d_jac_fails) d_jac_fails = dev_alloc<int>((size_t)m.nc, owned);
    check_jac_device(this, d_Jy, st, env, d_Js, t1 - t0, d_jac_fails);
    std::vector<int> f(m.nc);
    HIP_OK(hipMemcpyAsync(f.data(), d_jac_fails, (size_t)m.nc * sizeof(int),
                          hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    for (int c = 0; c < m.nc; ++c) jac_eval_fails += f[c] != 0;
  }
  return rc;
}

CellBdfArgs Solver::cell_args() const {
  CellBdfArgs a;
  a.P = P;
  a.var_type = d_var_type;
  a.rtol = m.rel_tol;
  a.max_steps = m.max_steps;
  a.max_conv_fails = m.max_conv_fails;
  a.guess_on = guess_on;
  a.place = cell_placement(P, place);
  return a;
}

void Solver::cell_launch_shape(long nq, int* grid, int* block, size_t* lds, int at_place) const {
  int cpb = cells_per_block;
  const int place = at_place < 0 ? this->place : at_place;
  *block = kWave * cpb;
  *lds = (size_t)cpb * cell_bdf_lds_doubles(P, place) * sizeof(double);
  int dev = 0, ncu = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int per_cu = cell_bdf_blocks_per_cu(P, place, *block, *lds);
  if (per_cu <= 0) die("per-cell kernel does not fit on a compute unit (LDS)");
  *grid = (int)std::max<long>(1, std::min<long>((long)per_cu * ncu, (nq + cpb - 1) / cpb));
}

int Solver::run_cells(double* st, double* env, double t0, double t1, double* noalloc,
                      size_t noalloc_doubles) {
  CellBdfArgs a = cell_args();
  a.nc = m.nc;
  a.state = st;
  a.env = env;
  a.rxn_env = d_renv;
  a.Js = d_Js;
  a.Jy = d_Jy;
  a.Jf = d_Jf;
  a.t0 = t0;
  a.t1 = t1;
  a.counter = d_counter;
  a.stats = d_stats;
  if (order_valid) a.cell_list = d_order;
  a.poison = poison;
  if (profiling && !noalloc) {
    if (!d_prof) d_prof = dev_alloc<double>((size_t)m.nc * kCellProf, owned);
    a.prof = d_prof;
  }
  int grid = 0, block = 0;
  size_t lds = 0;
  cell_launch_shape(m.nc, &grid, &block, &lds);
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (noalloc) {
    // the queue counter in the scratch's first 256 bytes, the wavefronts'
    // scratch after it; as many workgroups as the scratch holds
    const size_t per_wave = cell_bdf_wave_scratch(P, place);
    if (per_wave > 0) {
      long waves_fit = (long)((noalloc_doubles - kNoAllocHeaderDoubles) / per_wave);
      grid = (int)std::max<long>(1, std::min<long>(grid, waves_fit / (block / kWave)));
    }
    a.counter = reinterpret_cast<int*>(noalloc);
    a.wave_scratch = per_wave > 0 ? noalloc + kNoAllocHeaderDoubles : nullptr;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, stream, 1L, 0.0, noalloc);
  } else {
    a.wave_scratch = wave_scratch((size_t)grid * (block / kWave), place);
    HIP_OK(hipMemsetAsync(d_counter, 0, sizeof(int), stream));
  }
  HIP_OK(hipEventRecord(ev0, stream));
  launch_cell_bdf(a, grid, block, lds, stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(ev1, stream));
  HIP_OK(hipStreamSynchronize(stream));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
  last_kernel_ms = ms;
  std::vector<double>& stats = h_stats;
  HIP_OK(hipMemcpy(stats.data(), d_stats, (size_t)m.nc * kCellStats * sizeof(double),
                   hipMemcpyDeviceToHost));
  // One integrator per cell: the counters are sums over the cells, the
  // flag is the lowest-numbered failing cell's, and the step sizes are the
  // smallest over the cells (the stiffest cell's, the step a shared
  // integrator over all cells would be held to).
  BlockStats s;
  int fail = 0;
  for (int c = 0; c < m.nc; ++c) {
    const double* x = &stats[(size_t)c * kCellStats];
    s.nst += (long)x[0]; s.nfe += (long)x[1]; s.nsetups += (long)x[2]; s.netf += (long)x[3];
    s.nni += (long)x[4]; s.ncfn += (long)x[5]; s.nje += (long)x[6];
    if (x[7] < 0 && !fail) {
      fail = 1;
      s.flag = (int)x[7];
    }
    s.hu = c == 0 ? x[8] : std::min(s.hu, x[8]);
    s.next_h = c == 0 ? x[9] : std::min(s.next_h, x[9]);
  }
  bstats = s;
  stats_from_cells = true;
  if (!explicit_order && (d_order || !noalloc)) {
    // Next call's queue order: the cells by this call's BDF step count,
    // most first (a counting sort on quarter octaves), so that a launch's
    // tail holds the cheap cells.  A cell's result does not depend on the
    // order or on which wavefront takes it.
    constexpr int kKeys = 128;
    auto key = [&](int c) {
      double n = stats[(size_t)c * kCellStats];
      int k = n > 0.0 ? (int)(4.0 * std::log2(1.0 + n)) : 0;
      return kKeys - 1 - std::min(kKeys - 1, std::max(0, k));
    };
    int start[kKeys + 1] = {0};
    for (int c = 0; c < m.nc; ++c) start[key(c) + 1]++;
    for (int k = 0; k < kKeys; ++k) start[k + 1] += start[k];
    h_order.resize(m.nc);
    for (int c = 0; c < m.nc; ++c) h_order[start[key(c)]++] = c;
    if (!d_order) d_order = dev_alloc<int>((size_t)m.nc, owned);
    HIP_OK(hipMemcpy(d_order, h_order.data(), (size_t)m.nc * sizeof(int),
                     hipMemcpyHostToDevice));
    order_valid = true;
  }
  if (profiling && !noalloc) {
    std::vector<double> pr((size_t)m.nc * kCellProf);
    HIP_OK(hipMemcpy(pr.data(), d_prof, pr.size() * sizeof(double), hipMemcpyDeviceToHost));
    int khz = 0;
    HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev));
    double cf = 0.0, cj = 0.0;
    for (int c = 0; c < m.nc; ++c) {
      cf += pr[(size_t)c * kCellProf];
      cj += pr[(size_t)c * kCellProf + 1];
    }
    rhs_time_s = khz > 0 ? cf / (khz * 1.0e3) : 0.0;
    jac_time_s = khz > 0 ? cj / (khz * 1.0e3) : 0.0;
  }
  return fail ? CAMP_SOLVER_FAIL : CAMP_SOLVER_SUCCESS;
}

// -------------------------------------------------------------------------
// Block path: host-driven BDF over all cells (CVODE algorithm, same
// constants as the per-cell kernel), one launch per vector operation,
// f(y), J(y), LU and solve.  Norms are reduced on the device and finished
// on the host.
// -------------------------------------------------------------------------
namespace {
constexpr int QMAX = 5;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2;
constexpr double ETAMIN = 0.1, ETACF = 0.25, ADDON = 0.000001;
constexpr double BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5;
constexpr int SMALL_NST = 10, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr double CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0, NLSCOEF = 0.1;
constexpr int MSBP = 20, MAXCOR = 3, MSBJ = 50;
constexpr double DGMAX_JBAD = 0.2;
constexpr int GUESS_MAX_ITER = 5;
constexpr int kRedBlocks = 512;
enum { CV_SUCCESS = 0, CV_TOO_MUCH_WORK = -1, CV_TOO_MUCH_ACC = -2, CV_ERR_FAILURE = -3,
       CV_CONV_FAILURE = -4, CV_FIRST_RHSFUNC_ERR = -9, CV_REPTD_RHSFUNC_ERR = -10,
       CV_UNREC_RHSFUNC_ERR = -11 };
enum { FIRST_CALL = 1, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN, CONV_FAIL, TRY_AGAIN, RHSFUNC_RECVR };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

double rpowerr(double b, double e) { return b <= 0.0 ? 0.0 : std::pow(b, e); }
double rpoweri(double b, int e) {
  double p = 1.0;
  for (int i = 1; i <= std::abs(e); ++i) p *= b;
  return e < 0 ? 1.0 / p : p;
}

struct BlockCv {
  Solver& S;
  long n;
  int nc;
  double* st;
  hipStream_t s;
  double *zn[6], *ewt, *acor, *y, *tempv, *ftemp, *last, *acinit, *gh1, *gh2;
  double tn = 0, h = 0, hscale = 0, hprime = 0, next_h = 0, eta = 1, etamax = ETAMX1, etaq = 0,
         etaqm1 = 0, etaqp1 = 0, hu = 0, init_h = 0;
  double l[6] = {0}, tq[6] = {0}, tau[8] = {0};
  double rl1 = 0, gamma = 0, gammap = 1, gamrat = 1, crate = 1, acnrm = 0, saved_tq5 = 0;
  int q = 1, qprime = 1, L = 2, qwait = 2, jcur = 0, use_acinit = 0, projected = 0;
  long nst = 0, nfe = 0, nsetups = 0, nni = 0, ncfn = 0, netf = 0, nje = 0, nstlp = 0,
       nstlj = 0;

  BlockCv(Solver& S_, double* st_) : S(S_), st(st_) {
    nc = S.m.nc;
    n = (long)S.P.nv * nc;
    s = S.stream;
    for (int j = 0; j < 6; ++j) zn[j] = S.d_zn + j * n;
    ewt = S.d_ewt; acor = S.d_acor; y = S.d_y; tempv = S.d_tempv; ftemp = S.d_ftemp;
    last = S.d_last; acinit = S.d_acinit; gh1 = S.d_gh1; gh2 = S.d_gh2;
  }

  void copy(double* dst, const double* src) {
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  void lin(double a, const double* x, double b, const double* yv, double* z) {
    hipLaunchKernelGGL(k_vec_lin, dim3(nblk(n)), dim3(256), 0, s, n, a, x, b, yv, z);
  }
  double wrms(const double* x) {
    hipLaunchKernelGGL(k_vec_wsq, dim3(kRedBlocks), dim3(256), 0, s, n, x, ewt, S.d_red);
    HIP_OK(hipMemcpyAsync(S.h_red.data(), S.d_red, kRedBlocks * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double acc = 0.0;
    for (int i = 0; i < kRedBlocks; ++i) acc += S.h_red[i];
    return std::sqrt(acc / (double)n);
  }
  double vmin(const double* x) {
    hipLaunchKernelGGL(k_vec_min, dim3(kRedBlocks), dim3(256), 0, s, n, x, S.d_red);
    HIP_OK(hipMemcpyAsync(S.h_red.data(), S.d_red, kRedBlocks * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double v = S.h_red[0];
    for (int i = 1; i < kRedBlocks; ++i) v = std::min(v, S.h_red[i]);
    return v;
  }
  // max over i with y_i < -SMALL of -y_i * ewt_i, or -1 when there is none
  double vnegmax(const double* x) {
    hipLaunchKernelGGL(k_vec_negmax, dim3(kRedBlocks), dim3(256), 0, s, n, x, ewt, S.d_red);
    HIP_OK(hipMemcpyAsync(S.h_red.data(), S.d_red, kRedBlocks * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double v = S.h_red[0];
    for (int i = 1; i < kRedBlocks; ++i) v = std::max(v, S.h_red[i]);
    return v;
  }
  void ewt_set(const double* ycur) {
    hipLaunchKernelGGL(k_vec_ewt, dim3(nblk(n)), dim3(256), 0, s, n, S.P.nv, S.m.rel_tol,
                       S.P.abs_tol_dep, ycur, ewt);
  }
  int check_fail() {
    int f = 0;
    HIP_OK(hipMemcpyAsync(&f, S.d_fail, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return f;
  }
  int deriv(const double* yv, double* out, int use_est) {
    double ts = next_h > 0.0 ? next_h : init_h;
    HIP_OK(hipMemsetAsync(S.d_fail, 0, sizeof(int), s));
    S.launch_deriv(yv, st, nullptr, out, ts, use_est, S.d_fail);
    return check_fail();
  }
  int f(const double* yv, double* out) {
    nfe++;
    return deriv(yv, out, 1);
  }
  int jac(const double* yv, double* fy) {
    if (deriv(yv, fy, 0)) return 1;
    double ts = next_h > 0.0 ? next_h : init_h;
    S.launch_jac(yv, st, nullptr, S.d_Js, ts, nullptr);
    copy(S.d_Jy, yv);
    copy(S.d_Jf, fy);
    return 0;
  }
  void rescale() {
    hipLaunchKernelGGL(k_vec_rescale, dim3(nblk(n)), dim3(256), 0, s, n, q, eta, S.d_zn);
    h = hscale * eta;
    next_h = h;
    hscale = h;
  }
  void predict() {
    tn += h;
    hipLaunchKernelGGL(k_vec_pascal, dim3(nblk(n)), dim3(256), 0, s, n, q, S.d_zn, 1);
  }
  void restore(double saved_t) {
    tn = saved_t;
    hipLaunchKernelGGL(k_vec_pascal, dim3(nblk(n)), dim3(256), 0, s, n, q, S.d_zn, -1);
  }
  void set_tq(double hsum, double alpha0, double alpha0_hat, double xi_inv, double xistar_inv) {
    double A1 = 1.0 - alpha0_hat + alpha0;
    double A2 = 1.0 + q * A1;
    tq[2] = std::fabs(A1 / (alpha0 * A2));
    tq[5] = std::fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        double C = xistar_inv / l[q];
        double A3 = alpha0 + 1.0 / q;
        double A4 = alpha0_hat + xi_inv;
        double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = std::fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      double A5 = alpha0 - (1.0 / (q + 1));
      double A6 = alpha0_hat - xi_inv;
      double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = std::fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
  }
  void set() {
    double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
    l[0] = l[1] = xi_inv = xistar_inv = 1.0;
    for (int z = 2; z <= q; ++z) l[z] = 0.0;
    alpha0 = alpha0_hat = -1.0;
    hsum = h;
    if (q > 1) {
      for (int j = 2; j < q; ++j) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        alpha0 -= 1.0 / j;
        for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
      for (int i = q; i >= 1; --i) l[i] += l[i - 1] * xistar_inv;
    }
    set_tq(hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }
  void increase_bdf() {
    double alpha0, alpha1, prod, xi, xiold, hsum, A1;
    for (int j = 0; j <= QMAX; ++j) l[j] = 0.0;
    l[2] = alpha1 = prod = xiold = 1.0;
    alpha0 = -1.0;
    hsum = hscale;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += tau[j + 1];
        xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    A1 = (-alpha0 - alpha1) / prod;
    lin(A1, zn[QMAX], 0.0, zn[QMAX], zn[L]);
    for (int j = 2; j <= q; ++j) lin(l[j], zn[L], 1.0, zn[j], zn[j]);
  }
  void decrease_bdf() {
    double hsum, xi;
    for (int j = 0; j <= QMAX; ++j) l[j] = 0.0;
    l[2] = 1.0;
    hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; ++j) lin(-l[j], zn[q], 1.0, zn[j], zn[j]);
  }
  void adjust_params() {
    if (qprime != q) {
      int dq = qprime - q;
      if (dq == 1) increase_bdf();
      else if (dq == -1) decrease_bdf();
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    rescale();
  }
  // guess_helper (camp_solver.c:1321-1408) over the whole block vector
  int guess_helper(double t_n, double h_n, const double* y_n, const double* y_n1, const double* hf,
                   double* tmp1, double* corr) {
    if (vmin(y_n) > -1.0e-30) return 0;
    copy(tmp1, y_n1);
    lin(1.0 / h_n, hf, 0.0, hf, corr);
    double t_0 = t_n - h_n, t_j = 0.0;
    for (int iter = 0; iter < GUESS_MAX_ITER && t_0 + t_j < t_n; iter++) {
      double h_j = t_n - (t_0 + t_j);
      hipLaunchKernelGGL(k_vec_tstar, dim3(kRedBlocks), dim3(256), 0, s, n, tmp1, corr, h_j,
                         S.d_red);
      std::vector<double> red(2 * kRedBlocks);
      HIP_OK(hipMemcpyAsync(red.data(), S.d_red, red.size() * sizeof(double),
                            hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      bool found = false;
      double best = h_j;
      for (int b = 0; b < kRedBlocks; ++b)
        if (red[2 * b + 1] >= 0.0) {
          found = true;
          best = std::min(best, red[2 * b]);
        }
      h_j = best;
      if (found) h_j *= 0.95 + 0.1 * iter / (double)GUESS_MAX_ITER;
      h_j = t_n < t_0 + t_j + h_j ? t_n - (t_0 + t_j) : h_j;
      lin(1.0, tmp1, h_j, corr, tmp1);
      t_j += h_j;
      if (f(tmp1, corr)) return -1;
    }
    lin(1.0, tmp1, -1.0, y_n, corr);
    hipLaunchKernelGGL(k_vec_scale, dim3(nblk(n)), dim3(256), 0, s, n, 0.999, corr);
    return 1;
  }
  int lsetup(int convfail, const double* ypred, double* fpred) {
    double dgamma = std::fabs(gamma / gammap - 1.0);
    int jbad = (nst == 0) || (nst > nstlj + MSBJ) ||
               ((convfail == CV_FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) || (convfail == CV_FAIL_OTHER);
    if (!jbad) {
      jcur = 0;
    } else {
      nje++;
      nstlj = nst;
      jcur = 1;
      if (jac(ypred, fpred)) return 1;
    }
    HIP_OK(hipMemsetAsync(S.d_fail, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_lu_factor, dim3(std::min(nc, 16384)), dim3(64),
                       ((size_t)S.P.lu_nnz + kWave) * sizeof(double), s, S.P, nc, S.d_Js, gamma, S.d_LU,
                       S.d_fail);
    return check_fail();
  }
  double lsolve_scale() const { return gamrat != 1.0 ? 2.0 / (1.0 + gamrat) : 1.0; }
  void lsolve(double* b, const double* sp = nullptr) {
    hipLaunchKernelGGL(k_lu_solve, dim3(std::min(nc, 16384)), dim3(64),
                       ((size_t)S.P.lu_nnz + kWave + 1 + 2 * (size_t)S.P.nv) * sizeof(double), s,
                       S.P, nc, S.d_LU, b, lsolve_scale(), sp);
  }
  // Non-negativity safeguard, the per-cell kernel's rule: components of y
  // below -SMALL by at most their tolerance go to TINY through corr.
  // Returns -1 (nothing below -SMALL), 0 (projected) or 1 (f() would fail).
  int project(double* corr) {
    double worst = vnegmax(y);
    if (worst > 1.0) return 1;
    if (worst < 0.0) return -1;
    hipLaunchKernelGGL(k_vec_project, dim3(nblk(n)), dim3(256), 0, s, n, zn[0], corr, y);
    return 0;
  }
  // ---- HIP graphs of the Newton iteration ---------------------------------
  // One Newton iteration's device work up to the host's convergence decision
  // is captured once per solver as graph A: the per-iteration scalars (rl1,
  // gamma, the solve's scale, the time step) copied from pinned host memory,
  // the residual, the batched LU solve, the WRMS partial sums of the
  // correction, the update of acor and y, and the partial maxima of the
  // negative-concentration test of y -- with both partial arrays copied to
  // pinned host memory.  Graph B is f(y) into ftemp with its failure flag
  // (captured per state pointer and f kernel width).  A Newton iteration then costs one graph
  // launch and one synchronisation, plus graph B when it continues, instead
  // of seven launches and three synchronisations; the kernels and the host's
  // reductions are the same, so the results are identical.  No memset is
  // recorded in a graph (flags are cleared with a fill kernel).
  void graph_capture(hipGraphExec_t* out, bool b) {
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (!b) {
      HIP_OK(hipMemcpyAsync(S.d_ctl, S.h_ctl, 4 * sizeof(double), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_vec_resid, dim3(nblk(n)), dim3(256), 0, s, n, 0.0, zn[1], acor, 0.0,
                         ftemp, tempv, (const double*)S.d_ctl);
      lsolve(tempv, S.d_ctl + 2);
      hipLaunchKernelGGL(k_vec_wsq, dim3(kRedBlocks), dim3(256), 0, s, n, tempv, ewt, S.d_red);
      hipLaunchKernelGGL(k_vec_newton_update, dim3(nblk(n)), dim3(256), 0, s, n, tempv, acor,
                         zn[0], y);
      hipLaunchKernelGGL(k_vec_negmax, dim3(kRedBlocks), dim3(256), 0, s, n, y, ewt,
                         S.d_red + kRedBlocks);
      HIP_OK(hipMemcpyAsync(S.h_red_pin, S.d_red, 2 * kRedBlocks * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    } else {
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, 1L, 0, S.d_fail);
      launch_group_deriv(S.f_lanes, S.P, nc, y, st, S.kval(), S.d_Js, S.d_Jy, S.d_Jf, ftemp,
                         0.0, S.d_ctl + 3, 1, S.d_fail, s);
      HIP_OK(hipMemcpyAsync(S.h_fail_pin, S.d_fail, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    HIP_OK(hipStreamEndCapture(s, &g));
    HIP_OK(hipGraphInstantiate(out, g, nullptr, nullptr, 0));
    HIP_OK(hipGraphDestroy(g));
  }
  void graphs_ready() {
    if (!S.h_ctl) {
      HIP_OK(hipHostMalloc((void**)&S.h_ctl, 4 * sizeof(double), hipHostMallocDefault));
      HIP_OK(hipHostMalloc((void**)&S.h_red_pin, 2 * kRedBlocks * sizeof(double),
                           hipHostMallocDefault));
      HIP_OK(hipHostMalloc((void**)&S.h_fail_pin, sizeof(int), hipHostMallocDefault));
      S.d_ctl = dev_alloc<double>(4, S.owned);
    }
    if (!S.graph_a) graph_capture(&S.graph_a, false);
    if (!S.graph_b || S.graph_b_state != st || S.graph_b_lanes != S.f_lanes) {
      if (S.graph_b) HIP_OK(hipGraphExecDestroy(S.graph_b));
      graph_capture(&S.graph_b, true);
      S.graph_b_state = st;
      S.graph_b_lanes = S.f_lanes;
    }
  }
  int newton_iteration_graph() {
    graphs_ready();
    int m = 0;
    double del = 0.0, delp = 0.0;
    for (;;) {
      S.h_ctl[0] = rl1;
      S.h_ctl[1] = gamma;
      S.h_ctl[2] = lsolve_scale();
      S.h_ctl[3] = next_h > 0.0 ? next_h : init_h;
      HIP_OK(hipGraphLaunch(S.graph_a, s));
      HIP_OK(hipStreamSynchronize(s));
      nni++;
      double acc = 0.0, worst = S.h_red_pin[kRedBlocks];
      for (int i = 0; i < kRedBlocks; ++i) acc += S.h_red_pin[i];
      for (int i = 1; i < kRedBlocks; ++i) worst = std::max(worst, S.h_red_pin[kRedBlocks + i]);
      del = std::sqrt(acc / (double)n);
      if (m > 0) crate = std::fmax(CRDOWN * crate, del / delp);
      double dcon = del * std::fmin(1.0, crate) / tq[4];
      if (dcon <= 1.0) {
        int pr = project_with(acor, worst);
        if (pr > 0) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
        projected = pr == 0;
        acnrm = (m == 0 && pr < 0) ? del : wrms(acor);
        jcur = 0;
        return CV_SUCCESS;
      }
      m++;
      if ((m == MAXCOR) || ((m >= 2) && (del > RDIV * delp))) return jcur ? CONV_FAIL : TRY_AGAIN;
      delp = del;
      if (project_with(acor, worst) > 0) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
      nfe++;
      HIP_OK(hipGraphLaunch(S.graph_b, s));
      HIP_OK(hipStreamSynchronize(s));
      if (*S.h_fail_pin) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
    }
  }
  // project() with the worst excursion already known
  int project_with(double* corr, double worst) {
    if (worst > 1.0) return 1;
    if (worst < 0.0) return -1;
    hipLaunchKernelGGL(k_vec_project, dim3(nblk(n)), dim3(256), 0, s, n, zn[0], corr, y);
    return 0;
  }

  int newton_iteration() {
    if (S.use_graphs && !S.P.mp) return newton_iteration_graph();
    int m = 0;
    double del = 0.0, delp = 0.0;
    for (;;) {
      hipLaunchKernelGGL(k_vec_resid, dim3(nblk(n)), dim3(256), 0, s, n, rl1, zn[1], acor, gamma,
                         ftemp, tempv, (const double*)nullptr);
      lsolve(tempv);
      nni++;
      del = wrms(tempv);
      hipLaunchKernelGGL(k_vec_newton_update, dim3(nblk(n)), dim3(256), 0, s, n, tempv, acor,
                         zn[0], y);
      if (m > 0) crate = std::fmax(CRDOWN * crate, del / delp);
      double dcon = del * std::fmin(1.0, crate) / tq[4];
      if (dcon <= 1.0) {
        int pr = project(acor);
        if (pr > 0) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
        projected = pr == 0;
        acnrm = (m == 0 && pr < 0) ? del : wrms(acor);
        jcur = 0;
        return CV_SUCCESS;
      }
      m++;
      if ((m == MAXCOR) || ((m >= 2) && (del > RDIV * delp))) return jcur ? CONV_FAIL : TRY_AGAIN;
      delp = del;
      if (project(acor) > 0) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
      if (f(y, ftemp)) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
    }
  }
  int nls(int nflag) {
    int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
    int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) ||
                    (nst >= nstlp + MSBP) || (std::fabs(gamrat - 1.0) > DGMAX);
    use_acinit = 0;
    if (S.guess_on) {
      lin(1.0, zn[0], -1.0, last, acinit);
      if (guess_helper(tn, h, zn[0], last, acinit, gh1, gh2) == 1) {
        copy(acinit, gh2);
        use_acinit = 1;
      }
    }
    for (;;) {
      if (use_acinit) lin(1.0, zn[0], 1.0, acinit, y);
      else copy(y, zn[0]);
      if (vmin(y) < -1.0e-30) {
        if (!use_acinit) {
          HIP_OK(hipMemsetAsync(acinit, 0, n * sizeof(double), s));
          use_acinit = 1;
        }
        project(acinit);
      }
      if (f(y, ftemp)) return RHSFUNC_RECVR;
      if (callSetup) {
        int ier = lsetup(convfail, y, ftemp);
        nsetups++;
        callSetup = 0;
        gamrat = crate = 1.0;
        gammap = gamma;
        nstlp = nst;
        if (ier > 0) return CONV_FAIL;
      }
      if (use_acinit) copy(acor, acinit);
      else HIP_OK(hipMemsetAsync(acor, 0, n * sizeof(double), s));
      int ier = newton_iteration();
      if (ier != TRY_AGAIN) return ier;
      callSetup = 1;
      convfail = CV_FAIL_BAD_J;
    }
  }
  int handle_nflag(int* nflagPtr, double saved_t, int* ncfPtr) {
    int nflag = *nflagPtr;
    if (nflag == CV_SUCCESS) return DO_ERROR_TEST;
    ncfn++;
    restore(saved_t);
    (*ncfPtr)++;
    etamax = 1.0;
    if (*ncfPtr == S.m.max_conv_fails) {
      if (nflag == CONV_FAIL) return CV_CONV_FAILURE;
      if (nflag == RHSFUNC_RECVR) return CV_REPTD_RHSFUNC_ERR;
    }
    eta = ETACF;
    *nflagPtr = PREV_CONV_FAIL;
    rescale();
    return PREDICT_AGAIN;
  }
  int do_error_test(int* nflagPtr, double saved_t, int* nefPtr, double* dsmPtr) {
    double dsm = acnrm * tq[2];
    *dsmPtr = dsm;
    if (dsm <= 1.0) return CV_SUCCESS;
    (*nefPtr)++;
    netf++;
    *nflagPtr = PREV_ERR_FAIL;
    restore(saved_t);
    if (*nefPtr == S.m.max_conv_fails) return CV_ERR_FAILURE;
    etamax = 1.0;
    if (*nefPtr <= MXNEF1) {
      eta = 1.0 / (rpowerr(BIAS2 * dsm, 1.0 / L) + ADDON);
      eta = std::fmax(ETAMIN, std::fmax(eta, 0.0));
      if (*nefPtr >= SMALL_NEF) eta = std::fmin(eta, ETAMXF);
      rescale();
      return TRY_AGAIN;
    }
    if (q > 1) {
      eta = ETAMIN;
      decrease_bdf();
      L = q;
      q--;
      qwait = L;
      rescale();
      return TRY_AGAIN;
    }
    eta = ETAMIN;
    h *= eta;
    next_h = h;
    hscale = h;
    qwait = LONG_WAIT;
    if (f(zn[0], tempv)) return CV_UNREC_RHSFUNC_ERR;
    lin(h, tempv, 0.0, tempv, zn[1]);
    return TRY_AGAIN;
  }
  void complete_step() {
    nst++;
    hu = h;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if ((q == 1) && (nst > 1)) tau[2] = tau[1];
    tau[1] = h;
    qwait--;
    int save = (qwait == 1) && (q != QMAX);
    Coef6 c6;
    for (int j = 0; j < 6; ++j) c6.v[j] = l[j];
    hipLaunchKernelGGL(k_vec_complete, dim3(nblk(n)), dim3(256), 0, s, n, q, c6, acor, S.d_zn,
                       save, last, projected, y);
    if (save) saved_tq5 = tq[5];
  }
  void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = std::fmin(eta, etamax);
      hprime = h * eta;
    }
  }
  void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = std::max(qwait, 2);
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (rpowerr(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    etaqm1 = 0.0;
    if (q > 1) {
      double ddn = wrms(zn[q]) * tq[1];
      etaqm1 = 1.0 / (rpowerr(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    etaqp1 = 0.0;
    if (q != QMAX && saved_tq5 != 0.0) {
      double cquot = (tq[5] / saved_tq5) * rpoweri(h / tau[2], L);
      lin(-cquot, zn[QMAX], 1.0, acor, tempv);
      double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (rpowerr(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    double etam = std::fmax(etaqm1, std::fmax(etaq, etaqp1));
    if (etam < THRESH) {
      eta = 1.0;
      qprime = q;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
      copy(zn[QMAX], acor);
    }
    set_eta();
  }
  int step() {
    double saved_t = tn, dsm = 0.0;
    int ncf = 0, nef = 0, nflag = FIRST_CALL;
    if ((nst > 0) && (hprime != h)) adjust_params();
    for (;;) {
      predict();
      set();
      nflag = nls(nflag);
      int kflag = handle_nflag(&nflag, saved_t, &ncf);
      if (kflag == PREDICT_AGAIN) continue;
      if (kflag != DO_ERROR_TEST) return kflag;
      int eflag = do_error_test(&nflag, saved_t, &nef, &dsm);
      if (eflag == TRY_AGAIN) continue;
      if (eflag != CV_SUCCESS) return eflag;
      break;
    }
    complete_step();
    prepare_next_step(dsm);
    etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    hipLaunchKernelGGL(k_vec_scale, dim3(nblk(n)), dim3(256), 0, s, n, tq[2], acor);
    return CV_SUCCESS;
  }
  int solve(double t0, double tout) {
    copy(last, zn[0]);
    tn = t0;
    init_h = tout - t0;
    next_h = 0.0;
    ewt_set(zn[0]);
    if (f(zn[0], zn[1])) return CV_FIRST_RHSFUNC_ERR;
    h = init_h;
    hipLaunchKernelGGL(k_vec_scale, dim3(nblk(n)), dim3(256), 0, s, n, h, zn[1]);
    hscale = h;
    hprime = h;
    long nstloc = 0;
    for (;;) {
      next_h = h;
      if (nst > 0) ewt_set(zn[0]);
      if (nstloc >= S.m.max_steps) return CV_TOO_MUCH_WORK;
      double nrm = wrms(zn[0]);
      if (2.220446049250313e-16 * nrm > 1.0) return CV_TOO_MUCH_ACC;
      int kflag = step();
      if (kflag != CV_SUCCESS) return kflag;
      nstloc++;
      if ((tn - tout) * h >= 0.0) {
        double sv = (tout - tn) / h;
        hipLaunchKernelGGL(k_vec_interp, dim3(nblk(n)), dim3(256), 0, s, n, q, sv, S.d_zn, y);
        next_h = hprime;
        return CV_SUCCESS;
      }
    }
  }
};
}  // namespace

int Solver::run_block(double* st, double* env, double t0, double t1) {
  if (!d_zn) {
    long n = (long)P.nv * m.nc;
    d_zn = dev_alloc<double>(6 * n, owned);
    d_ewt = dev_alloc<double>(n, owned);
    d_acor = dev_alloc<double>(n, owned);
    d_y = dev_alloc<double>(n, owned);
    d_tempv = dev_alloc<double>(n, owned);
    d_ftemp = dev_alloc<double>(n, owned);
    d_last = dev_alloc<double>(n, owned);
    d_acinit = dev_alloc<double>(n, owned);
    d_gh1 = dev_alloc<double>(n, owned);
    d_gh2 = dev_alloc<double>(n, owned);
    d_LU = dev_alloc<double>((size_t)m.nc * P.lu_nnz, owned);
    d_red = dev_alloc<double>(2 * kRedBlocks, owned);
    h_red.assign(2 * kRedBlocks, 0.0);
  }
  int nc = m.nc;
  hipLaunchKernelGGL(k_clamp_constants, dim3(nblk((long)nc * P.ns)), dim3(256), 0, stream, P, nc,
                     d_var_type, st);
  hipLaunchKernelGGL(k_rates, dim3(nblk((long)nc * m.n_rxn)), dim3(256), 0, stream, P, nc, env,
                     d_renv);
  refresh_kval();
  hipLaunchKernelGGL(k_state_gather, dim3(nblk((long)nc * P.nv)), dim3(256), 0, stream, P, nc, st,
                     d_zn);
  BlockCv cv(*this, st);
  HIP_OK(hipEventRecord(ev0, stream));
  int flag = cv.solve(t0, t1);
  HIP_OK(hipEventRecord(ev1, stream));
  if (flag >= 0)
    hipLaunchKernelGGL(k_state_scatter, dim3(nblk((long)nc * P.nv)), dim3(256), 0, stream, P, nc,
                       d_y, st);
  HIP_OK(hipStreamSynchronize(stream));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
  last_kernel_ms = ms;
  BlockStats b;
  b.nst = cv.nst; b.nfe = cv.nfe; b.nsetups = cv.nsetups; b.netf = cv.netf;
  b.nni = cv.nni; b.ncfn = cv.ncfn; b.nje = cv.nje; b.flag = flag;
  b.hu = cv.hu; b.next_h = cv.next_h;
  bstats = b;
  stats_from_cells = false;
  return flag < 0 ? CAMP_SOLVER_FAIL : CAMP_SOLVER_SUCCESS;
}

static Solver* SD(void* p) { return static_cast<Solver*>(p); }

}  // namespace camp

using camp::Solver;
using camp::SD;

// ---------------------------------------------------------------------------
// Drop-in boundary
// ---------------------------------------------------------------------------
extern "C" void* solver_new(int n_state_var, int n_cells, int* var_type, int n_rxn,
                            int n_rxn_int_param, int n_rxn_float_param, int n_rxn_env_param,
                            int n_aero_phase, int n_aero_phase_int_param,
                            int n_aero_phase_float_param, int n_aero_rep, int n_aero_rep_int_param,
                            int n_aero_rep_float_param, int n_aero_rep_env_param,
                            int n_sub_model, int n_sub_model_int_param,
                            int n_sub_model_float_param, int n_sub_model_env_param) {
  if (n_state_var <= 0 || n_cells <= 0 || n_rxn < 0 || n_aero_phase < 0 || n_aero_rep < 0 ||
      n_sub_model < 0)
    camp::die("invalid solver_new dimensions");
  (void)n_aero_rep_env_param;
  (void)n_sub_model_env_param;
  Solver* s = new Solver();
  camp::Model& m = s->m;
  m.ns = n_state_var;
  m.nc = n_cells;
  m.var_type.assign(var_type, var_type + n_state_var);
  m.nv = 0;
  for (int i = 0; i < n_state_var; ++i) m.nv += var_type[i] == camp::kVariable;
  m.n_rxn = n_rxn;
  m.rxn_int.assign((size_t)n_rxn_int_param + n_rxn, 0);
  m.rxn_float.assign(std::max(n_rxn_float_param, 0), 0.0);
  m.rxn_env.assign((size_t)n_cells * std::max(n_rxn_env_param, 0), 0.0);
  m.rxn_int_idx.assign(n_rxn + 1, 0);
  m.rxn_float_idx.assign(n_rxn + 1, 0);
  m.rxn_env_idx.assign(n_rxn + 1, 0);
  m.size_aero(n_aero_phase, n_aero_phase_int_param, n_aero_phase_float_param, n_aero_rep,
              n_aero_rep_int_param, n_aero_rep_float_param, n_sub_model, n_sub_model_int_param,
              n_sub_model_float_param);
  return s;
}

extern "C" void rxn_add_condensed_data(int rxn_type, int n_int_param, int n_float_param,
                                       int n_env_param, int* int_param, double* float_param,
                                       void* solver_data) {
  Solver* s = SD(solver_data);
  if (s->m.n_added >= s->m.n_rxn) camp::die("more reactions added than declared in solver_new");
  s->m.add_rxn(rxn_type, n_int_param, n_float_param, n_env_param, int_param, float_param);
}

extern "C" void solver_initialize(void* solver_data, double* abs_tol, double rel_tol,
                                  int max_steps, int max_conv_fails) {
  Solver* s = SD(solver_data);
  camp::Model& m = s->m;
  if ((size_t)m.nc * m.n_env != m.rxn_env.size()) m.rxn_env.assign((size_t)m.nc * m.n_env, 0.0);
  m.abs_tol.assign(abs_tol, abs_tol + m.ns);
  m.rel_tol = rel_tol;
  m.max_steps = max_steps;
  m.max_conv_fails = max_conv_fails;
  std::string err = m.initialize();
  if (!err.empty()) camp::die(err.c_str());
  s->renv_dirty = true;
}

extern "C" int solver_run(void* solver_data, double* state, double* env, double t_initial,
                          double t_final) {
  Solver* s = SD(solver_data);
  if (!s->m.initialized) return CAMP_SOLVER_FAIL;
  s->ensure_device();
  size_t nst = (size_t)s->m.nc * s->m.ns;
  HIP_OK(hipMemcpyAsync(s->d_state, state, nst * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemcpyAsync(s->d_env, env, (size_t)s->m.nc * 2 * sizeof(double),
                        hipMemcpyHostToDevice, s->stream));
  int rc = s->run(s->d_state, s->d_env, t_initial, t_final);
  HIP_OK(hipMemcpyAsync(state, s->d_state, nst * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  return rc;
}

extern "C" void solver_get_statistics(void* solver_data, int* solver_flag, int* num_steps,
                                      int* RHS_evals, int* LS_setups, int* error_test_fails,
                                      int* NLS_iters, int* NLS_convergence_fails,
                                      int* DLS_Jac_evals, int* DLS_RHS_evals,
                                      double* last_time_step__s, double* next_time_step__s,
                                      int* Jac_eval_fails, int* RHS_evals_total,
                                      int* Jac_evals_total, double* RHS_time__s,
                                      double* Jac_time__s, double* max_loss_precision) {
  const camp::BlockStats& b = SD(solver_data)->bstats;
  // counters summed over many cells can pass INT_MAX: saturate
  auto sat = [](long v) { return (int)std::min<long>(v, std::numeric_limits<int>::max()); };
  *solver_flag = b.flag;
  *num_steps = sat(b.nst);
  *RHS_evals = sat(b.nfe);
  *LS_setups = sat(b.nsetups);
  *error_test_fails = sat(b.netf);
  *NLS_iters = sat(b.nni);
  *NLS_convergence_fails = sat(b.ncfn);
  *DLS_Jac_evals = sat(b.nje);
  // the Jacobian is analytic: no difference-quotient f() calls
  *DLS_RHS_evals = 0;
  *last_time_step__s = b.hu;
  *next_time_step__s = b.next_h;
  // Jac() results checked against difference quotients and found wrong
  // (solver_set_eval_jac); the remaining fields are those of a build
  // without CAMP_DEBUG (camp_solver.c:784-789), except that the f()/Jac()
  // times are reported when phase timing is on
  *Jac_eval_fails = sat(SD(solver_data)->jac_eval_fails);
  *RHS_evals_total = -1;
  *Jac_evals_total = -1;
  *RHS_time__s = SD(solver_data)->rhs_time_s;
  *Jac_time__s = SD(solver_data)->jac_time_s;
  *max_loss_precision = 0.0;
}

extern "C" void solver_reset_timers(void* solver_data) {
  SD(solver_data)->rhs_time_s = 0.0;
  SD(solver_data)->jac_time_s = 0.0;
}

extern "C" void solver_free(void* solver_data) { delete SD(solver_data); }

extern "C" void rxn_update_data(int cell_id, int* rxn_id, int update_rxn_type, void* update_data,
                                void* solver_data) {
  Solver* s = SD(solver_data);
  if (cell_id < 0 || cell_id >= s->m.nc) return;
  int uid;
  double rate;
  std::memcpy(&uid, update_data, sizeof(int));
  std::memcpy(&rate, static_cast<char*>(update_data) + sizeof(int), sizeof(double));
  if (s->m.update_data(cell_id, rxn_id, update_rxn_type, uid, rate)) s->renv_dirty = true;
}

extern "C" void rxn_print_data(void* solver_data) {
  Solver* s = SD(solver_data);
  std::printf("\n\nReaction data\n\nnumber of reactions: %d\n\n", s->m.n_rxn);
  for (int r = 0; r < s->m.n_rxn; ++r) std::printf("reaction %d type %d\n", r, s->m.rint(r)[0]);
  std::fflush(stdout);
}

extern "C" void rxn_free_update_data(void* update_data) { std::free(update_data); }

static void* new_rate_payload() {
  void* p = std::malloc(sizeof(int) + sizeof(double));
  if (!p) camp::die("allocating space for rate update data");
  return p;
}
static void set_rate_payload(void* p, int id, double rate) {
  std::memcpy(p, &id, sizeof(int));
  std::memcpy(static_cast<char*>(p) + sizeof(int), &rate, sizeof(double));
}
extern "C" void* rxn_photolysis_create_rate_update_data(void) { return new_rate_payload(); }
extern "C" void rxn_photolysis_set_rate_update_data(void* u, int photo_id, double base_rate) {
  set_rate_payload(u, photo_id, base_rate);
}
extern "C" void* rxn_emission_create_rate_update_data(void) { return new_rate_payload(); }
extern "C" void rxn_emission_set_rate_update_data(void* u, int rxn_id, double base_rate) {
  set_rate_payload(u, rxn_id, base_rate);
}
extern "C" void* rxn_first_order_loss_create_rate_update_data(void) { return new_rate_payload(); }
extern "C" void rxn_first_order_loss_set_rate_update_data(void* u, int rxn_id, double base_rate) {
  set_rate_payload(u, rxn_id, base_rate);
}

// aero_phase_solver.c:165-194, aero_rep_solver.c:443-482, sub_model_solver.c:413-453
extern "C" void aero_phase_add_condensed_data(int n_int_param, int n_float_param, int* int_param,
                                              double* float_param, void* solver_data) {
  SD(solver_data)->m.add_phase(n_int_param, n_float_param, int_param, float_param);
}
extern "C" void aero_rep_add_condensed_data(int aero_rep_type, int n_int_param, int n_float_param,
                                            int n_env_param, int* int_param, double* float_param,
                                            void* solver_data) {
  SD(solver_data)->m.add_rep(aero_rep_type, n_int_param, n_float_param, n_env_param, int_param,
                             float_param);
}
extern "C" void sub_model_add_condensed_data(int sub_model_type, int n_int_param,
                                             int n_float_param, int n_env_param, int* int_param,
                                             double* float_param, void* solver_data) {
  SD(solver_data)->m.add_sub(sub_model_type, n_int_param, n_float_param, n_env_param, int_param,
                             float_param);
}

// sub_model_solver.c:466-504: sub models of the requested type are scanned
// from *sub_model_id on, advancing it; no supported sub-model type takes
// update data (ZSR_AEROSOL_WATER has none), so the scan only advances the id.
extern "C" void sub_model_update_data(int cell_id, int* sub_model_id, int update_sub_model_type,
                                      void* update_data, void* solver_data) {
  (void)cell_id;
  (void)update_sub_model_type;
  (void)update_data;
  camp::Model& m = SD(solver_data)->m;
  for (; *sub_model_id < m.n_added_sub; ++(*sub_model_id)) {
  }
}

// sub_model_solver.c:509-541
extern "C" void sub_model_print_data(void* solver_data) {
  camp::Model& m = SD(solver_data)->m;
  for (int i = 0; i < m.n_added_sub; ++i) {
    const int* d = &m.sub_int[m.sub_int_idx[i]];
    std::printf("\n\nSub model %d: type %d\n", i, d[0]);
  }
  std::fflush(stdout);
}

// aero_phase_solver.c:199-218
extern "C" void aero_phase_print_data(void* solver_data) {
  camp::Model& m = SD(solver_data)->m;
  for (int i = 0; i < m.n_added_phase; ++i) std::printf("\n\nAerosol Phase %d\n\n", i);
  std::fflush(stdout);
}

// aero_rep_solver.c:492-541 with aero_rep_modal_binned_mass.c:594-642: the
// representations of the requested type are scanned from *aero_rep_id on
// (advancing it); a MODAL_BINNED_MASS one whose unique id (int slot 3 after
// the type) matches a non-zero payload id takes the update.  The payload is
// {int aero_rep_id, int update_type (0 GMD, 1 GSD), int section_id, double
// value} as the *_create/set_*_update_data helpers below pack it.  GMD and
// GSD describe MODAL sections; this solver accepts BINNED sections only, so
// a matching update is the reference's error for a non-modal section.
extern "C" void aero_rep_update_data(int cell_id, int* aero_rep_id, int update_aero_rep_type,
                                     void* update_data, void* solver_data) {
  (void)cell_id;
  camp::Model& m = SD(solver_data)->m;
  int uid = 0, utype = 0, section = 0;
  std::memcpy(&uid, update_data, sizeof(int));
  std::memcpy(&utype, static_cast<char*>(update_data) + sizeof(int), sizeof(int));
  std::memcpy(&section, static_cast<char*>(update_data) + 2 * sizeof(int), sizeof(int));
  for (; *aero_rep_id < m.n_added_rep; ++(*aero_rep_id)) {
    const int* d = &m.rep_int[m.rep_int_idx[*aero_rep_id]];
    if (d[0] != update_aero_rep_type || d[0] != camp::kAeroRepModalBinnedMass) continue;
    const int* id = d + 1;
    if (uid != id[3] || id[3] == 0) continue;
    if (utype == 0) {
      std::printf("\n\nERROR Trying to set geometric mean diameter for non-modal"
                  " aerosol section.");
    } else {
      std::printf("\n\nERROR Trying to set geometric standard deviation for non-modal"
                  " aerosol section.");
    }
    std::fflush(stdout);
    std::exit(1);
  }
}

// aero_rep_solver.c:546-582
extern "C" void aero_rep_print_data(void* solver_data) {
  camp::Model& m = SD(solver_data)->m;
  std::printf("\n\nAerosol representation data\n\nnumber of aerosol representations: %d\n\n",
              m.n_added_rep);
  for (int i = 0; i < m.n_added_rep; ++i) {
    const int* d = &m.rep_int[m.rep_int_idx[i]];
    if (d[0] == camp::kAeroRepModalBinnedMass)
      std::printf("\n\nModal/binned mass-only aerosol representation\n");
  }
  std::fflush(stdout);
}

// aero_rep_solver.c:589
extern "C" void aero_rep_free_update_data(void* update_data) { std::free(update_data); }

// Payload helpers of reaction and representation types this solver does not
// integrate (a mechanism that uses them stops at solver_initialize with the
// setup error), exported so that every Fortran module of the reference
// links: rxn_wet_deposition.c:218-239 and rxn_condensed_phase_photolysis.c:
// 464-485 ({int rxn_id; double rate}, as photolysis), aero_rep_single_particle
// .c:405-433 ({int aero_rep_id; int update_type (0: number); int
// particle_id; double number_conc}).  rxn_update_data and
// aero_rep_update_data find no matching reaction or representation for them.
extern "C" void* rxn_wet_deposition_create_rate_update_data(void) { return new_rate_payload(); }
extern "C" void rxn_wet_deposition_set_rate_update_data(void* u, int rxn_id, double base_rate) {
  set_rate_payload(u, rxn_id, base_rate);
}
extern "C" void* rxn_condensed_phase_photolysis_create_rate_update_data(void) {
  return new_rate_payload();
}
extern "C" void rxn_condensed_phase_photolysis_set_rate_update_data(void* u, int photo_id,
                                                                    double base_rate) {
  set_rate_payload(u, photo_id, base_rate);
}
extern "C" void* aero_rep_single_particle_create_number_update_data(void) {
  void* p = std::malloc(3 * sizeof(int) + sizeof(double));
  if (!p) camp::die("allocating space for number update data");
  return p;
}
extern "C" void aero_rep_single_particle_set_number_update_data__n_m3(void* update_data,
                                                                     int aero_rep_id,
                                                                     int particle_id,
                                                                     double number_conc) {
  int head[3] = {aero_rep_id, 0, particle_id};
  std::memcpy(update_data, head, sizeof(head));
  std::memcpy(static_cast<char*>(update_data) + sizeof(head), &number_conc, sizeof(double));
}

// aero_rep_modal_binned_mass.c:677-738: payload {int aero_rep_id, int
// update_type, int section_id, double value} (3 ints then a double, packed)
static void* new_section_payload() {
  void* p = std::malloc(3 * sizeof(int) + sizeof(double));
  if (!p) camp::die("allocating space for aerosol representation update data");
  return p;
}
static void set_section_payload(void* p, int rep_id, int type, int section, double v) {
  std::memcpy(p, &rep_id, sizeof(int));
  std::memcpy(static_cast<char*>(p) + sizeof(int), &type, sizeof(int));
  std::memcpy(static_cast<char*>(p) + 2 * sizeof(int), &section, sizeof(int));
  std::memcpy(static_cast<char*>(p) + 3 * sizeof(int), &v, sizeof(double));
}
extern "C" void* aero_rep_modal_binned_mass_create_gmd_update_data(void) {
  return new_section_payload();
}
extern "C" void aero_rep_modal_binned_mass_set_gmd_update_data(void* update_data, int aero_rep_id,
                                                               int section_id, double gmd) {
  set_section_payload(update_data, aero_rep_id, 0, section_id, gmd);
}
extern "C" void* aero_rep_modal_binned_mass_create_gsd_update_data(void) {
  return new_section_payload();
}
extern "C" void aero_rep_modal_binned_mass_set_gsd_update_data(void* update_data, int aero_rep_id,
                                                               int section_id, double gsd) {
  set_section_payload(update_data, aero_rep_id, 1, section_id, gsd);
}

// camp_solver.c:502-533 (CAMP_DEBUG builds): the debug-output flag is
// accepted and ignored (this solver prints no per-step trace); the
// Jacobian-check flag turns on the finite-difference check of every
// Jacobian the solver evaluates (Jac_eval_fails counts cells that fail).
extern "C" int solver_set_debug_out(void* solver_data, bool do_output) {
  (void)solver_data;
  (void)do_output;
  return CAMP_SOLVER_SUCCESS;
}
extern "C" int solver_set_eval_jac(void* solver_data, bool eval_jac) {
  SD(solver_data)->eval_jac = eval_jac;
  return CAMP_SOLVER_SUCCESS;
}

// ---------------------------------------------------------------------------
// Extensions
// ---------------------------------------------------------------------------
extern "C" int camp_solver_set_mode(void* solver_data, int mode) {
  Solver* s = SD(solver_data);
  int old = s->mode;
  if (mode == CAMP_MODE_PER_CELL || mode == CAMP_MODE_BLOCK) s->mode = mode;
  return old;
}

extern "C" int camp_solver_set_cells_per_block(void* solver_data, int cpb) {
  using namespace camp;
  Solver* s = SD(solver_data);
  if (cpb == 0) cpb = 1;
  if (cpb == 1 || cpb == 2 || cpb == 4 || cpb == 8 || cpb == 16 || cpb == 32) {
    size_t lds = (size_t)cpb * cell_bdf_lds_doubles(s->P, s->place) * sizeof(double);
    if (!s->m.initialized || !s->stream ||
        cell_bdf_blocks_per_cu(s->P, s->place, kWave * cpb, lds) > 0)
      s->cells_per_block = cpb;
  }
  return s->cells_per_block;
}

extern "C" int camp_solver_set_eval_kernels(void* solver_data, int group) {
  camp::Solver* s = SD(solver_data);
  int old = s->eval_group ? 1 : 0;
  s->eval_group = group != 0;
  return old;
}

extern "C" long camp_solver_cell_lds_bytes(void* solver_data, int place) {
  using namespace camp;
  Solver* s = SD(solver_data);
  if (!s->m.initialized || place < 0 || place >= kNumPlacements) return -1;
  s->ensure_device();
  return (long)(cell_bdf_lds_doubles(s->P, cell_placement(s->P, place)) * sizeof(double));
}

extern "C" int camp_solver_set_placement(void* solver_data, int place) {
  using namespace camp;
  Solver* s = SD(solver_data);
  if (place >= 0 && place < kNumPlacements) s->place = place;
  return cell_placement(s->P, s->place);
}

// Autotuner for the persistent per-cell kernel.  Candidates: the cell
// placement (history in LDS: fewer cells per CU; history in global memory:
// more cells per CU, two wavefronts per SIMD; history and LU factors in
// global memory: three wavefronts per SIMD) times cells per workgroup
// {1, 2, 4, 8, 16, 32}; every candidate that fits a compute unit solves the
// same sample of cells (every (nc / n_sample)-th cell, copied with its rate
// data and J history, so the solver's own cells are untouched) over
// [t0, t1], and the fastest by HIP events on the solver's stream is kept.
extern "C" int camp_solver_autotune(void* solver_data, const double* d_state,
                                    const double* d_env, double t0, double t1, int n_sample,
                                    double* ms_out) {
  using namespace camp;
  Solver* s = SD(solver_data);
  if (!s->m.initialized) return 0;
  s->ensure_device();
  s->upload_renv();
  const Model& m = s->m;
  const DevProgram& P = s->P;
  long nc = m.nc;
  long ns = n_sample > 0 ? std::min<long>(n_sample, nc) : std::min<long>(nc, 8192);
  long stride = std::max<long>(1, nc / ns);
  std::vector<void*> tmp;
  auto take = [&](long n) {
    double* p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<long>(n, 1) * sizeof(double)));
    tmp.push_back(p);
    return p;
  };
  double* st = take(ns * P.ns);
  double* env = take(ns * 2);
  double* renv = take(ns * P.n_env);
  double* js = take(ns * P.nnz);
  double* jy = take(ns * P.nv);
  double* jf = take(ns * P.nv);
  double* stats = take(ns * kCellStats);
  auto gather = [&](const double* src, int len, double* dst) {
    long n = ns * len;
    hipLaunchKernelGGL(k_gather_rows, dim3(nblk(n)), dim3(256), 0, s->stream, ns, len, stride,
                       src, dst);
  };
  // The sample's queue order: longest first by the last call's BDF step
  // counts, as a solve orders its cells (else index order).
  int* d_sample_order = nullptr;
  if (s->stats_from_cells && s->h_stats.size() >= (size_t)nc * kCellStats) {
    std::vector<std::pair<double, int>> by(ns);
    for (long i = 0; i < ns; ++i) by[i] = {-s->h_stats[(size_t)(i * stride) * kCellStats], (int)i};
    std::stable_sort(by.begin(), by.end());
    std::vector<int> ord(ns);
    for (long i = 0; i < ns; ++i) ord[i] = by[i].second;
    std::vector<void*> o;
    d_sample_order = dev_upload(ord, o);
    tmp.push_back(d_sample_order);
  }
  const int cand[6] = {1, 2, 4, 8, 16, 32};
  const int keep_cpb = s->cells_per_block, keep_place = s->place;
  int best = keep_cpb, best_place = cell_placement(P, keep_place);
  double best_ms = -1.0;
  for (int pl = 0; pl < kNumPlacements; ++pl) {
    const int place = pl;
    for (int k = 0; k < 6; ++k) {
      double* slot = ms_out ? &ms_out[pl * 6 + k] : nullptr;
      if (slot) *slot = -1.0;
      // multi-phase programs have one placement
      if (cell_placement(P, place) != place) continue;
      int cpb = cand[k];
      size_t lds = (size_t)cpb * cell_bdf_lds_doubles(P, place) * sizeof(double);
      if (cell_bdf_blocks_per_cu(P, place, kWave * cpb, lds) <= 0) continue;
      gather(d_state, P.ns, st);
      gather(d_env, 2, env);
      gather(s->d_renv, P.n_env, renv);
      gather(s->d_Js, P.nnz, js);
      gather(s->d_Jy, P.nv, jy);
      gather(s->d_Jf, P.nv, jf);
      s->cells_per_block = cpb;
      s->place = place;
      CellBdfArgs a = s->cell_args();
      a.nc = (int)ns;
      a.state = st;
      a.env = env;
      a.rxn_env = renv;
      a.Js = js;
      a.Jy = jy;
      a.Jf = jf;
      a.t0 = t0;
      a.t1 = t1;
      a.counter = s->d_counter;
      a.stats = stats;
      a.cell_list = d_sample_order;
      int grid = 0, block = 0;
      s->cell_launch_shape(ns, &grid, &block, &lds);
      a.wave_scratch = s->wave_scratch((size_t)grid * (block / kWave), place);
      s->cells_per_block = keep_cpb;
      s->place = keep_place;
      HIP_OK(hipMemsetAsync(s->d_counter, 0, sizeof(int), s->stream));
      HIP_OK(hipEventRecord(s->ev0, s->stream));
      launch_cell_bdf(a, grid, block, lds, s->stream);
      HIP_OK(hipGetLastError());
      HIP_OK(hipEventRecord(s->ev1, s->stream));
      HIP_OK(hipStreamSynchronize(s->stream));
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
      if (slot) *slot = ms;
      if (best_ms < 0.0 || ms < best_ms) {
        best_ms = ms;
        best = cpb;
        best_place = place;
      }
    }
  }
  for (void* p : tmp) HIP_OK(hipFree(p));
  s->cells_per_block = best;
  s->place = best_place;
  return best + 100 * best_place;
}

// Phase measurement of the per-cell kernel (gas-phase programs, placements
// 0 and 1): every wavefront stages a cell of a sample (copies, as the
// autotuner takes them) and times f(y), J(y), the LU factorisation and one
// triangular solve at its state, reps times each; out4 receives the mean
// cycles over the sample.  Returns 0, or 1 when the program or placement is
// not measurable this way.
extern "C" int camp_solver_phase_cycles(void* solver_data, const double* d_state,
                                        const double* d_env, int n_sample, int reps,
                                        double* out4) {
  using namespace camp;
  Solver* s = SD(solver_data);
  if (!s->m.initialized || reps < 1) return 1;
  s->ensure_device();
  if (cell_placement(s->P, s->place) == kPlaceGlobalLu) return 1;
  s->upload_renv();
  const DevProgram& P = s->P;
  long nc = s->m.nc;
  long ns = n_sample > 0 ? std::min<long>(n_sample, nc) : std::min<long>(nc, 4096);
  long stride = std::max<long>(1, nc / ns);
  std::vector<void*> tmp;
  auto take = [&](long n) { return dev_alloc<double>((size_t)std::max<long>(n, 1), tmp); };
  double* st = take(ns * P.ns);
  double* env = take(ns * 2);
  double* renv = take(ns * P.n_env);
  double* js = take(ns * P.nnz);
  double* jy = take(ns * P.nv);
  double* jf = take(ns * P.nv);
  double* stats = take(ns * kCellStats);
  double* prof = take(ns * kCellProf);
  auto gather = [&](const double* src, int len, double* dst) {
    long n = ns * len;
    hipLaunchKernelGGL(k_gather_rows, dim3(nblk(n)), dim3(256), 0, s->stream, ns, len, stride,
                       src, dst);
  };
  gather(d_state, P.ns, st);
  gather(d_env, 2, env);
  gather(s->d_renv, P.n_env, renv);
  gather(s->d_Js, P.nnz, js);
  gather(s->d_Jy, P.nv, jy);
  gather(s->d_Jf, P.nv, jf);
  CellBdfArgs a = s->cell_args();
  a.nc = (int)ns;
  a.state = st;
  a.env = env;
  a.rxn_env = renv;
  a.Js = js;
  a.Jy = jy;
  a.Jf = jf;
  a.counter = s->d_counter;
  a.stats = stats;
  a.prof = prof;
  a.bench_reps = reps;
  int grid = 0, block = 0;
  size_t lds = 0;
  s->cell_launch_shape(ns, &grid, &block, &lds);
  a.wave_scratch = s->wave_scratch((size_t)grid * (block / kWave), s->place);
  HIP_OK(hipMemsetAsync(s->d_counter, 0, sizeof(int), s->stream));
  launch_cell_bdf_bench(a, grid, block, lds, s->stream);
  HIP_OK(hipGetLastError());
  std::vector<double> h((size_t)ns * kCellProf);
  HIP_OK(hipMemcpyAsync(h.data(), prof, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 4; ++k) {
    double acc = 0.0;
    for (long c = 0; c < ns; ++c) acc += h[(size_t)c * kCellProf + k];
    out4[k] = acc / (double)ns;
  }
  for (void* p : tmp) HIP_OK(hipFree(p));
  return 0;
}

extern "C" void camp_solver_set_guess_helper(void* solver_data, int on) { SD(solver_data)->guess_on = on != 0; }

extern "C" int camp_solver_set_cell_order(void* solver_data, const int* order, int n) {
  using namespace camp;
  Solver* s = SD(solver_data);
  if (!order || n == 0) {
    s->h_order.clear();
    s->explicit_order = false;
    s->order_valid = false;
    return 0;
  }
  if (n != s->m.nc) return 1;
  std::vector<char> seen((size_t)n, 0);
  for (int i = 0; i < n; ++i) {
    if (order[i] < 0 || order[i] >= n || seen[order[i]]) return 1;
    seen[order[i]] = 1;
  }
  s->ensure_device();
  if (!s->d_order) s->d_order = dev_alloc<int>((size_t)n, s->owned);
  s->h_order.assign(order, order + n);
  s->explicit_order = true;
  s->order_valid = true;
  HIP_OK(hipMemcpy(s->d_order, s->h_order.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

extern "C" void camp_solver_set_poison(void* solver_data, int on) { SD(solver_data)->poison = on != 0; }

extern "C" int camp_solver_run_device(void* solver_data, double* d_state, double* d_env,
                                      double t_initial, double t_final) {
  Solver* s = SD(solver_data);
  if (!s->m.initialized) return CAMP_SOLVER_FAIL;
  s->ensure_device();
  int rc = s->run(d_state, d_env, t_initial, t_final);
  HIP_OK(hipStreamSynchronize(s->stream));
  return rc;
}

extern "C" int camp_solver_set_base_rates(void* solver_data, int rxn_type, int unique_id,
                                          const double* rates) {
  Solver* s = SD(solver_data);
  int found = 0;
  for (int c = 0; c < s->m.nc; ++c) {
    int rid = 0;
    found += s->m.update_data(c, &rid, rxn_type, unique_id, rates[c]);
  }
  s->renv_dirty = true;
  return found == s->m.nc ? 0 : 1;
}

extern "C" int camp_solver_dims(void* solver_data, int* n_dep, int* nnz, int* lu_nnz,
                                int* nnz_rxn) {
  Solver* s = SD(solver_data);
  *n_dep = s->m.prog.nv;
  *nnz = s->m.prog.nnz;
  *lu_nnz = s->m.prog.lu_nnz;
  *nnz_rxn = s->m.rjac.nnz();
  return 0;
}

extern "C" void camp_solver_get_pattern(void* solver_data, int* rxn_colptr, int* rxn_rowval,
                                        int* colptr, int* rowval, int* deriv_ids,
                                        int* rxn_int_data) {
  camp::Model& m = SD(solver_data)->m;
  std::copy(m.rjac.colptr.begin(), m.rjac.colptr.end(), rxn_colptr);
  std::copy(m.rjac.rowval.begin(), m.rjac.rowval.end(), rxn_rowval);
  std::copy(m.prog.colptr.begin(), m.prog.colptr.end(), colptr);
  std::copy(m.prog.rowval.begin(), m.prog.rowval.end(), rowval);
  std::copy(m.deriv_ids.begin(), m.deriv_ids.end(), deriv_ids);
  std::copy(m.rxn_int.begin(), m.rxn_int.end(), rxn_int_data);
}

static void eval_prepare(Solver* s, const double* state, const double* env, double* d_y) {
  int nc = s->m.nc;
  const camp::DevProgram& P = s->P;
  HIP_OK(hipMemcpyAsync(s->d_state, state, (size_t)nc * P.ns * sizeof(double),
                        hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemcpyAsync(s->d_env, env, (size_t)nc * 2 * sizeof(double), hipMemcpyHostToDevice,
                        s->stream));
  s->upload_renv();
  hipLaunchKernelGGL(camp::k_rates, dim3(camp::nblk((long)nc * s->m.n_rxn)), dim3(256), 0,
                     s->stream, P, nc, s->d_env, s->d_renv);
  // the solver variables, taken from the state as given (no clamping)
  std::vector<double> y((size_t)nc * P.nv);
  for (int c = 0; c < nc; ++c)
    for (int d = 0; d < P.nv; ++d)
      y[(size_t)c * P.nv + d] = state[(size_t)c * P.ns + s->m.prog.dep_state[d]];
  HIP_OK(hipMemcpyAsync(d_y, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
}

extern "C" int camp_eval_deriv(void* solver_data, const double* state, const double* env,
                               double time_step, int use_deriv_est, double* deriv) {
  Solver* s = SD(solver_data);
  s->ensure_device();
  int nc = s->m.nc;
  const camp::DevProgram& P = s->P;
  std::vector<void*> tmp;
  double* d_y = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  double* d_f = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  eval_prepare(s, state, env, d_y);
  HIP_OK(hipMemsetAsync(s->d_fail, 0, sizeof(int), s->stream));
  if (s->eval_group && !P.mp) {
    s->refresh_kval();
    s->launch_deriv(d_y, s->d_state, s->d_env, d_f, time_step, use_deriv_est, s->d_fail);
  } else {
    int wpb = camp::fj_waves_per_block(P);
    size_t lds = wpb * camp::fj_lds_doubles(P) * sizeof(double);
    hipLaunchKernelGGL(camp::k_deriv, dim3(std::min(camp::nblk(nc, wpb), 4096)), dim3(64 * wpb),
                       lds, s->stream, P, nc, d_y, s->d_state, s->d_env, s->d_renv, s->d_Js,
                       s->d_Jy, s->d_Jf, d_f, time_step, use_deriv_est, s->d_fail);
  }
  int fail = 0;
  HIP_OK(hipMemcpyAsync(&fail, s->d_fail, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(deriv, d_f, (size_t)nc * P.nv * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  for (void* p : tmp) (void)hipFree(p);
  return fail;
}

extern "C" int camp_eval_jac(void* solver_data, const double* state, const double* env,
                             double time_step, double* jac) {
  Solver* s = SD(solver_data);
  s->ensure_device();
  int nc = s->m.nc;
  const camp::DevProgram& P = s->P;
  std::vector<void*> tmp;
  double* d_y = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  eval_prepare(s, state, env, d_y);
  HIP_OK(hipMemsetAsync(s->d_fail, 0, sizeof(int), s->stream));
  int wpb = camp::fj_waves_per_block(P);
  size_t lds = wpb * camp::fj_lds_doubles(P) * sizeof(double);
  double* d_jr = P.mp ? camp::dev_alloc<double>((size_t)nc * P.rnnz, tmp) : nullptr;
  hipLaunchKernelGGL(camp::k_deriv, dim3(std::min(camp::nblk(nc, wpb), 4096)), dim3(64 * wpb), lds,
                     s->stream, P, nc, d_y, s->d_state, s->d_env, s->d_renv, s->d_Js, s->d_Jy,
                     s->d_Jf, s->d_Jf, time_step, 0, s->d_fail);
  if (s->eval_group && !P.mp) {
    s->refresh_kval();
    s->launch_jac(d_y, s->d_state, s->d_env, s->d_Js, time_step, nullptr);
  } else {
    hipLaunchKernelGGL(camp::k_jac, dim3(std::min(camp::nblk(nc, wpb), 4096)), dim3(64 * wpb),
                       lds, s->stream, P, nc, d_y, s->d_state, s->d_env, s->d_renv, s->d_Js,
                       time_step, d_jr);
  }
  HIP_OK(hipMemcpyAsync(s->d_Jy, d_y, (size_t)nc * P.nv * sizeof(double), hipMemcpyDeviceToDevice,
                        s->stream));
  int fail = 0;
  HIP_OK(hipMemcpyAsync(&fail, s->d_fail, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(jac, s->d_Js, (size_t)nc * P.nnz * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  for (void* p : tmp) (void)hipFree(p);
  return fail;
}

namespace camp {
// Finite-difference check of per-cell Jacobians on the device (k_check_jac):
// fails[c] = failing elements of cell c (-1: f() failed at its state).
void check_jac_device(Solver* s, const double* d_y, const double* d_state, const double* d_env,
                      const double* d_J, double time_step, int* d_fails) {
  const DevProgram& P = s->P;
  size_t lds = check_jac_lds_doubles(P) * sizeof(double);
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_check_jac,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_check_jac, dim3(std::min(s->m.nc, 16384)), dim3(64), lds, s->stream, P,
                     s->m.nc, d_y, d_state, d_env, s->d_renv, d_J, time_step, d_fails);
  HIP_OK(hipGetLastError());
}
}  // namespace camp

extern "C" int camp_solver_check_jac(void* solver_data, const double* state, const double* env,
                                     double time_step, const double* jac, int* fails) {
  Solver* s = SD(solver_data);
  if (!s->m.initialized) return -1;
  s->ensure_device();
  int nc = s->m.nc;
  const camp::DevProgram& P = s->P;
  std::vector<void*> tmp;
  double* d_y = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  int* d_fails = camp::dev_alloc<int>((size_t)nc, tmp);
  double* d_J = s->d_Js;
  if (jac) {
    // the caller's Jacobian: rate constants and state staged as for J(y)
    eval_prepare(s, state, env, d_y);
    d_J = camp::dev_alloc<double>((size_t)nc * P.nnz, tmp);
    HIP_OK(hipMemcpyAsync(d_J, jac, (size_t)nc * P.nnz * sizeof(double), hipMemcpyHostToDevice,
                          s->stream));
  } else {
    std::vector<double> J((size_t)nc * P.nnz);
    if (camp_eval_jac(solver_data, state, env, time_step, J.data()) != 0) {
      for (void* p : tmp) (void)hipFree(p);
      return -1;
    }
    HIP_OK(hipMemcpyAsync(d_y, s->d_Jy, (size_t)nc * P.nv * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
  }
  camp::check_jac_device(s, d_y, s->d_state, s->d_env, d_J, time_step, d_fails);
  std::vector<int> f(nc);
  HIP_OK(hipMemcpyAsync(f.data(), d_fails, (size_t)nc * sizeof(int), hipMemcpyDeviceToHost,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  for (void* p : tmp) (void)hipFree(p);
  long total = 0;
  bool fbad = false;
  for (int c = 0; c < nc; ++c) {
    if (fails) fails[c] = f[c];
    if (f[c] < 0) fbad = true;
    else total += f[c];
  }
  return fbad ? -1 : (int)std::min<long>(total, std::numeric_limits<int>::max());
}

extern "C" int camp_eval_lu_solve(void* solver_data, const double* jac, double gamma, double* b) {
  Solver* s = SD(solver_data);
  s->ensure_device();
  int nc = s->m.nc;
  const camp::DevProgram& P = s->P;
  std::vector<void*> tmp;
  double* d_J = camp::dev_alloc<double>((size_t)nc * P.nnz, tmp);
  double* d_LU = camp::dev_alloc<double>((size_t)nc * P.lu_nnz, tmp);
  double* d_b = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  HIP_OK(hipMemcpyAsync(d_J, jac, (size_t)nc * P.nnz * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemcpyAsync(d_b, b, (size_t)nc * P.nv * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemsetAsync(s->d_fail, 0, sizeof(int), s->stream));
  hipLaunchKernelGGL(camp::k_lu_factor, dim3(std::min(nc, 16384)), dim3(64),
                     ((size_t)P.lu_nnz + camp::kWave) * sizeof(double), s->stream, P, nc, d_J,
                     gamma, d_LU, s->d_fail);
  hipLaunchKernelGGL(camp::k_lu_solve, dim3(std::min(nc, 16384)), dim3(64),
                     ((size_t)P.lu_nnz + camp::kWave + 1 + 2 * (size_t)P.nv) * sizeof(double),
                     s->stream, P, nc, d_LU, d_b, 1.0, (const double*)nullptr);
  int fail = 0;
  HIP_OK(hipMemcpyAsync(&fail, s->d_fail, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(b, d_b, (size_t)nc * P.nv * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  for (void* p : tmp) (void)hipFree(p);
  return fail;
}

extern "C" void camp_solver_get_cell_stats(void* solver_data, double* stats) {
  Solver* s = SD(solver_data);
  s->ensure_device();
  HIP_OK(hipMemcpy(stats, s->d_stats, (size_t)s->m.nc * camp::kCellStats * sizeof(double),
                   hipMemcpyDeviceToHost));
}

extern "C" double camp_solver_last_kernel_ms(void* solver_data) { return SD(solver_data)->last_kernel_ms; }

extern "C" void camp_solver_set_profiling(void* solver_data, int on) {
  SD(solver_data)->profiling = on != 0;
}

extern "C" int camp_solver_get_cell_profile(void* solver_data, double* out) {
  Solver* s = SD(solver_data);
  if (!s->d_prof) return 1;
  HIP_OK(hipMemcpy(out, s->d_prof, (size_t)s->m.nc * camp::kCellProf * sizeof(double),
                   hipMemcpyDeviceToHost));
  return 0;
}

extern "C" void camp_solver_work_model(void* solver_data, double* out8) {
  const camp::Program& G = SD(solver_data)->m.prog;
  double f = 0.0, j = 0.0;
  // mass-action records: the rate product and the signed sum; slot records
  // (multi-phase contributions, counted with their instances below): a sum
  for (int i = 0; i < G.nv; ++i)
    for (int q = G.f_ptr[i]; q < G.f_ptr[i + 1]; ++q)
      f += G.f_term[q] >= 0 ? G.t_nreact[G.f_term[q]] + 2 : 2;
  f += 4.0 * G.nv + 2.0 * G.nnz + G.nv;  // blend + Jacobian-estimate matvec
  int jn = G.mp ? G.rnnz : G.nnz;
  for (int e = 0; e < jn; ++e)
    for (int q = G.j_ptr[e]; q < G.j_ptr[e + 1]; ++q) j += G.t_nreact[G.j_term[q]] + 1;
  if (G.mp) {
    // per-instance arithmetic of the multi-phase types (approximate counts
    // of their double-double operations) and the J_rxn * J_params map
    f += 40.0 * G.n_pi + 60.0 * G.n_zi + 80.0 * G.n_sp + 60.0 * G.n_aq;
    j += 12.0 * (G.js_ptr.back()) + 2.0 * G.jm_rxn.size() + 120.0 * G.n_sp + 40.0 * G.n_aq;
  }
  double lu = 2.0 * G.nnz + (double)G.l_pos.size() + 2.0 * G.up_dst.size();
  double sv = 2.0 * (G.lu_nnz - G.nv) + G.nv;
  double nr = G.n_terms;
  out8[0] = f;
  out8[1] = j;
  out8[2] = lu;
  out8[3] = sv;
  out8[4] = 8.0 * (2.0 * G.nv + nr);          // f(y):  read y, k; write f
  out8[5] = 8.0 * (G.nv + nr + G.nnz);        // J(y):  read y, k; write J
  out8[6] = 8.0 * (G.nnz + G.lu_nnz);         // LU:    read J; write LU
  out8[7] = 8.0 * (G.lu_nnz + 2.0 * G.nv);    // solve: read LU, b; write x
}

// f(y) and J(y) of the block path timed on device-resident data (HIP
// events on the solver's stream, `reps` launches after one untimed).  With
// tune, every lanes-per-cell width (8, 16, 32) of the lane-group kernels is
// timed and the fastest kept for f and for J; ms_out then holds
// [f at 8, 16, 32, J at 8, 16, 32].
static int time_fj_impl(Solver* s, const double* d_state, const double* d_env, int reps,
                        double* ms_deriv, double* ms_jac, bool tune, double* ms_out) {
  s->ensure_device();
  int nc = s->m.nc;
  const camp::DevProgram& P = s->P;
  std::vector<void*> tmp;
  double* d_y = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  double* d_f = camp::dev_alloc<double>((size_t)nc * P.nv, tmp);
  double* d_J = camp::dev_alloc<double>((size_t)nc * P.nnz, tmp);
  s->upload_renv();
  hipLaunchKernelGGL(camp::k_rates, dim3(camp::nblk((long)nc * s->m.n_rxn)), dim3(256), 0,
                     s->stream, P, nc, d_env, s->d_renv);
  s->refresh_kval();
  hipLaunchKernelGGL(camp::k_state_gather, dim3(camp::nblk((long)nc * P.nv)), dim3(256), 0,
                     s->stream, P, nc, d_state, d_y);
  double* d_jr = P.mp ? camp::dev_alloc<double>((size_t)nc * P.rnnz, tmp) : nullptr;
  auto time = [&](auto&& launch) {
    float ms = 0.f;
    launch();  // one untimed launch
    HIP_OK(hipEventRecord(s->ev0, s->stream));
    for (int r = 0; r < reps; ++r) launch();
    HIP_OK(hipEventRecord(s->ev1, s->stream));
    HIP_OK(hipEventSynchronize(s->ev1));
    HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    return (double)ms / reps;
  };
  auto run_f = [&] { s->launch_deriv(d_y, d_state, d_env, d_f, 1.0, 0, s->d_fail); };
  auto run_j = [&] { s->launch_jac(d_y, d_state, d_env, d_J, 1.0, d_jr); };
  if (tune && !P.mp) {
    const int widths[3] = {8, 16, 32};
    double bf = -1.0, bj = -1.0;
    int wf = s->f_lanes, wj = s->j_lanes;
    for (int w = 0; w < 3; ++w) {
      s->f_lanes = s->j_lanes = widths[w];
      double mf = time(run_f), mj = time(run_j);
      if (ms_out) {
        ms_out[w] = mf;
        ms_out[3 + w] = mj;
      }
      if (bf < 0.0 || mf < bf) { bf = mf; wf = widths[w]; }
      if (bj < 0.0 || mj < bj) { bj = mj; wj = widths[w]; }
    }
    s->f_lanes = wf;
    s->j_lanes = wj;
  }
  ms_deriv[0] = time(run_f);
  ms_jac[0] = time(run_j);
  HIP_OK(hipStreamSynchronize(s->stream));
  for (void* p : tmp) (void)hipFree(p);
  return 0;
}

extern "C" int camp_solver_time_fj(void* solver_data, const double* d_state, const double* d_env,
                                   int reps, double* ms_deriv, double* ms_jac) {
  return time_fj_impl(SD(solver_data), d_state, d_env, reps, ms_deriv, ms_jac, false, nullptr);
}

extern "C" int camp_solver_set_graphs(void* solver_data, int on) {
  camp::Solver* s = SD(solver_data);
  int old = s->use_graphs ? 1 : 0;
  s->use_graphs = on != 0;
  return old;
}

extern "C" int camp_solver_set_fj_lanes(void* solver_data, int f_lanes, int j_lanes) {
  camp::Solver* s = SD(solver_data);
  auto ok = [](int w) { return w == 8 || w == 16 || w == 32; };
  if (ok(f_lanes)) s->f_lanes = f_lanes;
  if (ok(j_lanes)) s->j_lanes = j_lanes;
  return s->f_lanes + 100 * s->j_lanes;
}

extern "C" int camp_solver_tune_fj(void* solver_data, const double* d_state, const double* d_env,
                                   int reps, double* ms_out) {
  camp::Solver* s = SD(solver_data);
  double a = 0.0, b = 0.0;
  time_fj_impl(s, d_state, d_env, reps, &a, &b, true, ms_out);
  return s->f_lanes + 100 * s->j_lanes;
}

extern "C" void camp_solver_get_rxn_env(void* solver_data, double* rxn_env) {
  Solver* s = SD(solver_data);
  s->ensure_device();
  HIP_OK(hipMemcpy(rxn_env, s->d_renv, s->m.rxn_env.size() * sizeof(double),
                   hipMemcpyDeviceToHost));
}

extern "C" void camp_cr_math(int fn, int n, const double* x, const double* y, double* out) {
  for (int i = 0; i < n; ++i) {
    switch (fn) {
      case 0: out[i] = camp::cr_exp(x[i]); break;
      case 1: out[i] = camp::cr_log(x[i]); break;
      case 2: out[i] = camp::cr_log10(x[i]); break;
      default: out[i] = camp::cr_pow(x[i], y[i]); break;
    }
  }
}
