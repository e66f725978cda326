// Device-side building blocks shared by the plain-path kernels and the
// persistent per-cell BDF kernel.
//
// Execution model: one wavefront (64 lanes) owns one grid cell at a time.
// The cell's state, rate constants, BDF vectors and LU factors live in LDS;
// lanes split the work of each sparse phase (rate terms and species in
// f(y), Jacobian elements in J(y), updates of one pivot in the LU, the
// column of one pivot in the triangular solves).  Every decision that
// controls the cell's integration is made on wave-uniform values (lane 0's
// reduction result is broadcast), so a wavefront never diverges between
// cells and never needs a workgroup barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "camp_dd.hpp"
#include "camp_model.hpp"

namespace camp {

constexpr int kWave = 64;

struct DevProgram {
  int ns, nv, nnz, n_env, n_terms, lu_nnz, n_rxn;
  const int* dep_state;
  const int* j_ptr;
  const int *colptr, *rowval;  // solver pattern (dep ids, CSC)
  // lane-group f/J programs (gas-phase programs; Program::gf_rec), one per
  // group width (8, 16, 32 lanes per cell)
  int g_f_rows[3], g_j_rows[3], n_k, n_fzero;
  const int4 *gf_rec[3], *gj_rec[3];   // 2 int4 per record
  const int4* cj_rec;                  // per-cell J (Program::cj_rec), 2 int4 per record
  int cj_rows;
  const int *k_slot, *f_count, *f_zero;
  const int *r_ptr, *r_elem, *r_col;  // row view of the solver pattern
  int n_cs;                            // state variables that are not solver variables
  const int* cs_ids;
  const int *perm, *m_to_lu, *diag_pos, *e_diag;
  const int *l_pos, *l_row, *u_pos, *u_row;  // column solve: padded columns (Program::cs_*)
  const uint64_t* up_rec;  // LU update records: dst | l << 16 | u << 32, rows of 64
  const int4* lu_hdr;       // per pivot: {l range, update range}, {diag position, u range}
  const int* lf_pos;        // L positions the factorisation scales, rows of 64
  const int4* lf_hdr;       // per pivot: {lf_pos row start, up_rec row start,
                            //  L rows | update rows << 16, diag position}
  const int4* ts_tbl;        // register solve: per block of 8 pivots and lane, 2 int4
                             //  of pos(row t) | pos(row t + 64) << 16; nullptr when nv > 128
  const int4* je_a;  // packed J records (see Program in camp_model.hpp)
  const int* je_b;
  const double* je_c;
  const int* fs_species;             // lane slot -> species of the f gather
  const int4* t_rec;                 // per term: {k index | flags << 28, reactants x3}
  const int4* fl_hdr;                // per pass: -, -, re_ell offset, width
  const int2* re_ell;                // row-view entries in slot-major order
  int fb_rows, fb_nparts;            // balanced f gather: rows per lane, partials
  const int4* fb_rec;                // {index | target << 16, guard | flags << 24, coef}
  const int2* fb_fin;                // per species: {first partial, count}
  // rate-constant program (rxn_update_env_state)
  const int *rx_type, *rx_fidx, *rx_eidx, *rx_nreact;
  const double* rx_float;
  const double* abs_tol_dep;  // [nv]
  // multi-phase tables (mp == 0 for gas-phase programs); see Program in
  // camp_model.hpp for their meaning
  int mp, rnnz, pnnz, n_pi, n_zi, n_sp, n_aq, n_slots, n_sub_env;
  double min_mass_hi, min_mass_lo, min_moles_hi, min_moles_lo, min_volume;
  const int *pi_start, *pi_model, *pm_ptr, *pm_counted;
  const double *pi_dp, *pm_mw, *pm_dens;
  const int *zi_water, *zi_gw, *zi_sub, *zi_pair_ptr, *zp_cat, *zp_an, *zp_ncat, *zp_nan, *zp_ny,
      *zp_jgw, *zp_jcat, *zp_jan;
  const double *zp_catmw, *zp_anmw, *zp_lowrh, *zp_y;
  const int *sp_gas, *sp_aero, *sp_dgas, *sp_daero, *sp_pi, *sp_env, *sp_flt, *sp_slot;
  const double* sp_radius;
  const int *aq_nr, *aq_np, *aq_sp, *aq_water, *aq_env, *aq_flt, *aq_slot;
  const double* aq_small_water;
  const int* js_ptr;
  const int4* js_a;
  const double* js_c;
  const int *jm_ptr, *jm_rxn, *jm_param;
  const int* mj_elist;  // J_rxn assembly order (Program::mj_elist)
  int mj_rows;
  const int4* mj_rec;   // the rows' record streams (Program::mj_rec), 2 int4 per record
  const int2* mj_rhdr;  // per row: first record row, records
};


// ---------------------------------------------------------------------
// Wavefront helpers
// ---------------------------------------------------------------------
// Orders the LDS (and global) accesses of one wavefront across lanes.  A
// wavefront's memory instructions are processed in program order, so a
// wavefront-scope fence needs no s_waitcnt: it only stops the compiler from
// moving memory operations across it, and outstanding global prefetches
// stay in flight.
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// Wait for every outstanding global load and store of the wavefront
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left open).  Placed in uniform
// control flow before a loop that keeps loads in flight across iterations:
// the compiler counts those loads exactly only when no store is pending on
// some path into the loop; with one pending (stores and loads complete in no
// fixed order) it waits for all loads at the top of every iteration, and
// the look-ahead hides nothing.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// The lane index made opaque to the compiler at the start of a phase (f, J,
// LU, solve): per-lane program addresses derived from it are then computed
// inside the phase instead of being hoisted to the kernel's start and held
// in registers through the whole solve (they raised the per-cell kernel's
// register use past the two-wavefront bound).
__host__ __device__ __forceinline__ int phase_lane(int lane) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(lane));
#endif
  return lane;
}
// Read-only program data at a wave-uniform index through the scalar cache
// (the program is never written while a kernel runs): the value lands in
// scalar registers without a vector load and a wait on it.
// (Host builds of the shared routines read it directly.)
template <typename T>
__host__ __device__ __forceinline__ T ld_uniform(const T* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(4))) T*)p)[i];
#else
  return p[i];
#endif
}
// Wavefront reductions with DPP lane moves (no LDS round trip): quad
// butterflies, half-row and row mirrors give every lane its row's value,
// then row broadcasts of lanes 15 and 31 carry the rows into lane 63, whose
// value is read into a scalar register.  Rows a broadcast does not write
// keep the identity.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v, int id) {
  return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v, double id) {
  int lo = dpp_i<CTRL, ROWS>(__double2loint(v), __double2loint(id));
  int hi = dpp_i<CTRL, ROWS>(__double2hiint(v), __double2hiint(id));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane63_d(double v) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), kWave - 1);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), kWave - 1);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0xB1, 0xF>(v, 0.0);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E, 0xF>(v, 0.0);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141, 0xF>(v, 0.0);  // row_half_mirror
  v += dpp_d<0x140, 0xF>(v, 0.0);  // row_mirror
  v += dpp_d<0x142, 0xA>(v, 0.0);  // row_bcast:15 into rows 1, 3
  v += dpp_d<0x143, 0xC>(v, 0.0);  // row_bcast:31 into rows 2, 3
  return lane63_d(v);
}
__device__ __forceinline__ double wave_min(double v) {
  const double inf = __builtin_huge_val();
  v = fmin(v, dpp_d<0xB1, 0xF>(v, inf));
  v = fmin(v, dpp_d<0x4E, 0xF>(v, inf));
  v = fmin(v, dpp_d<0x141, 0xF>(v, inf));
  v = fmin(v, dpp_d<0x140, 0xF>(v, inf));
  v = fmin(v, dpp_d<0x142, 0xA>(v, inf));
  v = fmin(v, dpp_d<0x143, 0xC>(v, inf));
  return lane63_d(v);
}
__device__ __forceinline__ int wave_or(int v) {
  v |= dpp_i<0xB1, 0xF>(v, 0);
  v |= dpp_i<0x4E, 0xF>(v, 0);
  v |= dpp_i<0x141, 0xF>(v, 0);
  v |= dpp_i<0x140, 0xF>(v, 0);
  v |= dpp_i<0x142, 0xA>(v, 0);
  v |= dpp_i<0x143, 0xC>(v, 0);
  return __builtin_amdgcn_readlane(v, kWave - 1);
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kGasConst = 8.314472;  // util.h:17

// --- environment: SIMPOL and aqueous equilibrium (rxn env slots) ----------
// rxn_SIMPOL_phase_transfer_update_env_state (:220-266):
// e = [MFP_M, ALPHA, EQUIL_CONST, KGM3_TO_PPM]
__host__ __device__ __forceinline__ void mp_simpol_env(const double* f, double* e, double T, double Pa) {
  double alpha = 0.1;
  if (f[0] != 0.0 || f[1] != 0.0) {
    double del_g = f[0] - T * f[1];
    alpha = cr_exp(-del_g / (kGasConst * T));
    alpha = alpha / (1.0 + alpha);
  }
  e[1] = alpha;
  e[0] = 3.0 * f[2] / sqrt(8.0 * kGasConst * T / (kPi * f[9]));
  double vp = f[4] / T + f[5] + f[6] * T + f[7] * cr_log(T);
  vp = 101325.0 * cr_pow(10.0, vp);
  e[3] = f[8] * T / Pa;
  e[2] = vp / Pa / f[9] * 1.0e6;
}
// rxn_aqueous_equilibrium_update_env_state (:244-265): e = [k_forward]
__host__ __device__ __forceinline__ void mp_aqeq_env(const double* f, double* e, double T) {
  double k = f[1] == 0.0 ? f[0] : f[0] * cr_exp(f[1] * (1.0 / T - 1.0 / 298.0));
  e[0] = k * f[2];
}
// sub_model_ZSR_aerosol_water_update_env_state (:189-207): PPM_TO_RH
__host__ __device__ __forceinline__ double mp_ppm_to_rh(double T, double Pa) {
  double a = 1.0 - 373.15 / T;
  a = (((-0.1299 * a - 0.6445) * a - 1.976) * a + 13.3185) * a;
  return Pa / (101325.0 * cr_exp(a)) / 1.0e6;
}

// ---------------------------------------------------------------------
// Rate constants of one cell: rxn_update_env_state (rxn_solver.c:224-313)
// with rxn_arrhenius.c:103-118, rxn_troe.c:106-130, rxn_CMAQ_H2O2.c:
// 106-124, rxn_CMAQ_OH_HNO3.c:110-133, rxn_photolysis.c:137-147,
// rxn_emission.c:109-119, rxn_first_order_loss.c:119-129.
// env_cell points at the cell's rxn_env block (k slots written in place).
// ---------------------------------------------------------------------
__host__ __device__ __forceinline__ void rate_constant(const DevProgram& P, int r, double T, double Pa,
                                              double* env_cell) {
  const double* f = P.rx_float + P.rx_fidx[r];
  double* e = env_cell + P.rx_eidx[r];
  int t = P.rx_type[r];
  int nr = P.rx_nreact[r];
  if (t == 1) {  // ARRHENIUS
    e[0] = f[0] * cr_exp(f[2] / T) * (f[1] == 0.0 ? 1.0 : cr_pow(T / f[3], f[1])) *
           (f[4] == 0.0 ? 1.0 : (1.0 + f[4] * Pa)) * cr_pow(f[5] * Pa / T, (double)(nr - 1));
  } else if (t == 2) {  // TROE
    double conv = f[9] * Pa / T;
    double k0 = f[0] * (f[2] == 0.0 ? 1.0 : cr_exp(f[2] / T)) *
                (f[1] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[1])) * conv;
    double kinf = k0 / (f[3] * (f[5] == 0.0 ? 1.0 : cr_exp(f[5] / T)) *
                        (f[4] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[4])));
    double lg = cr_log10(kinf) / f[7];
    e[0] = (k0 / (1.0 + kinf)) * cr_pow(f[6], (1.0 / (1.0 + cr_pow(lg, 2.0)))) *
           cr_pow(conv, (double)(nr - 1)) * f[8];
  } else if (t == 3) {  // CMAQ_H2O2: (k1 + k2 [M]) conv^(nr-1), [M] in k2_A
    double conv = f[6] * Pa / T;
    e[0] = (f[0] * (f[2] == 0.0 ? 1.0 : cr_exp(f[2] / T)) *
                (f[1] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[1])) +
            f[3] * (f[5] == 0.0 ? 1.0 : cr_exp(f[5] / T)) *
                (f[4] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[4])) * conv) *
           cr_pow(conv, (double)(nr - 1));
  } else if (t == 4) {  // CMAQ_OH_HNO3: (k0 + k3 / (1 + k3 / k2)) conv^(nr-1) SCALING
    double conv = f[10] * Pa / T;
    double k2 = f[3] * (f[5] == 0.0 ? 1.0 : cr_exp(f[5] / T)) *
                (f[4] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[4]));
    double k3 = f[6] * (f[8] == 0.0 ? 1.0 : cr_exp(f[8] / T)) *
                (f[7] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[7])) * conv;
    e[0] = (f[0] * (f[2] == 0.0 ? 1.0 : cr_exp(f[2] / T)) *
                (f[1] == 0.0 ? 1.0 : cr_pow(T / 300.0, f[1])) +
            k3 / (1.0 + k3 / k2)) *
           cr_pow(conv, (double)(nr - 1)) * f[9];
  } else if (t == 10) {  // SIMPOL_PHASE_TRANSFER
    mp_simpol_env(f, e, T, Pa);
  } else if (t == 7) {  // AQUEOUS_EQUILIBRIUM
    mp_aqeq_env(f, e, T);
  } else {  // PHOTOLYSIS, EMISSION, FIRST_ORDER_LOSS: SCALING * BASE_RATE
    e[0] = f[0] * e[1];
  }
}

// All rate constants of one cell, lanes over reactions.
__device__ __forceinline__ void cell_rate_constants(const DevProgram& P, int lane, double T, double Pa,
                                                 double* s_k) {
  for (int r = lane; r < P.n_rxn; r += kWave) rate_constant(P, r, T, Pa, s_k);
}

// ---------------------------------------------------------------------
// f(y) of one cell: per-type calc_deriv_contrib + time_derivative_output
// (camp_solver.c:901-954, time_derivative.c:46-97).  s_state holds the
// cell state with the solver variables already substituted, s_k the
// cell's rxn_env block.  Lane i handles species i, i+64, ...: it gathers
// its signed contributions into a double-double D (= production - loss)
// and their magnitudes into S (= production + loss).
// est != nullptr selects the Jacobian-extrapolated blend; est is then
// computed from the cell's J history (J_deriv + J_solver (y - J_state)).
// ---------------------------------------------------------------------
// Phase 1 of f(y): the rate of every term (reaction) once, into s_rate
// (n_terms double-doubles): k * prod(reactants) as the per-type
// calc_deriv_contrib computes it (rxn_arrhenius.c:143-147 and alike);
// first-order loss rounds k * y to double (rxn_first_order_loss.c:150).
__host__ __device__ __forceinline__ void cell_rates(const DevProgram& P, int lane, const double* s_state,
                                           const double* s_k, double* s_rate) {
  lane = phase_lane(lane);
  // a lane's terms in groups of four whose program loads are issued together
  constexpr int TG = 4;
  for (int t0 = lane; t0 < P.n_terms; t0 += TG * kWave) {
    int4 tr[TG];
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      int t = t0 + g * kWave;
      tr[g] = t < P.n_terms ? P.t_rec[t] : make_int4(0, -1, -1, -1);
    }
#pragma unroll
    for (int g = 0; g < TG; ++g) {
      int t = t0 + g * kWave;
      if (t >= P.n_terms) break;
      int fl = tr[g].x >> 28;
      double k = s_k[tr[g].x & 0x0FFFFFFF];
      double y0 = s_state[tr[g].y >= 0 ? tr[g].y : 0], y1 = s_state[tr[g].z >= 0 ? tr[g].z : 0],
             y2 = s_state[tr[g].w >= 0 ? tr[g].w : 0];
      dd rate = {k, 0.0};
      if (fl & 2) {
        rate = {k * y0, 0.0};
      } else {
        rate = tr[g].y >= 0 ? dd_mul_d(rate, y0) : rate;
        rate = tr[g].z >= 0 ? dd_mul_d(rate, y1) : rate;
        rate = tr[g].w >= 0 ? dd_mul_d(rate, y2) : rate;
      }
      s_rate[2 * t] = rate.hi;
      s_rate[2 * t + 1] = rate.lo;
    }
  }
}

// The double whose high and low 32-bit words are hi and lo.
__host__ __device__ __forceinline__ double hilo_to_double(int hi, int lo) {
  unsigned long long b = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  return __builtin_bit_cast(double, b);
}

// LDS scratch of the f gather: per species value and magnitude sum, the
// partials of split species (hi, lo, magnitudes) and a spare partial that
// padding records write to.
__host__ __device__ inline size_t fb_lds_doubles(const DevProgram& P) {
  return 2 * (size_t)P.nv + 3 * ((size_t)P.fb_nparts + 1);
}
// The per-cell kernel places the f gather's scratch right after the LU
// values, where it also serves as the factorisation's 64 sink slots and
// the triangular solve's zero slot (the 65th).
__host__ __device__ inline size_t fb_sink_doubles(const DevProgram& P) {
  size_t n = fb_lds_doubles(P);
  return n > (size_t)kWave + 1 ? n : (size_t)kWave + 1;
}

// One record of the f gather, without branches: the term's rate (or a
// multi-phase slot value, whose coefficient is 1) times the yield (or -1),
// unless the record fails the negative-yield guard -rate * yield * h <=
// y[guard] (rxn_arrhenius.c:157 and alike), in which case its coefficient
// is zero.
// The product of the double-double rate and the coefficient is p + e
// (exact up to the rounding of lo * c); p joins the running sum Dh with an
// error-free TwoSum whose error, with e, goes to Dl, so Dh + Dl is the
// record sum to about 2^-100 of the magnitudes S (the reference's long
// double production/loss sums keep 2^-64).  The running sums are stored at
// the piece's target after every record -- the last store of a piece is its
// value -- and a piece's last record (flag 16) restarts them from zero.
struct FOps {
  double hi, lo, sg;
};
__host__ __device__ __forceinline__ FOps f_operands(int4 rec, const double* s_state,
                                           const double* s_rate, const double* s_slot) {
  int flags = (int)((unsigned)rec.y >> 24);
  int g = (rec.y << 8) >> 8;  // sign-extended 24-bit guard (-1 = none)
  const double* src = (flags & 4) ? s_slot : s_rate;
  int idx = rec.x & 0xFFFF;
  FOps o;
  o.hi = src[2 * idx];
  o.lo = src[2 * idx + 1];
  o.sg = s_state[g >= 0 ? g : 0];
  return o;
}
__host__ __device__ __forceinline__ void f_record(int4 rec, FOps o, double h, int nv, double* s_fb,
                                         double& Dh, double& Dl, double& S) {
  int flags = (int)((unsigned)rec.y >> 24);
  int g = (rec.y << 8) >> 8;
  double c = hilo_to_double(rec.w, rec.z);
  // (padding has coefficient 0 and no guard; a zero-rate term flagged 1
  // adds zero, as the reference's skip does)
  bool keep = g < 0 || -o.hi * c * h <= o.sg;
  c = keep ? c : 0.0;
  double p = o.hi * c;
  double e = fma(o.hi, c, -p);
  e = fma(o.lo, c, e);
  double s = Dh + p;
  double bb = s - Dh;
  double err = (Dh - (s - bb)) + (p - bb);
  Dh = s;
  Dl = Dl + (err + e);
  S += fabs(p);
  // whole species: value Dh + Dl at [t], S at [nv + t]; partial: (Dh, Dl,
  // S) at [t], [t + 1], [t + 2] (the lo store first, so that for a whole
  // species the value overwrites it)
  int t = (int)((unsigned)rec.x >> 16);
  bool part = flags & 32;
  s_fb[part ? t + 1 : t] = Dl;
  s_fb[t] = part ? Dh : Dh + Dl;
  s_fb[part ? t + 2 : nv + t] = S;
  if (flags & 16) {
    Dh = 0.0;
    Dl = 0.0;
    S = 0.0;
  }
}

// f(y) after cell_rates and a wave sync.  Phase 2: every lane walks its bin
// of fb_rows records (pieces of species' record lists, packed so that the
// lanes' bins are balanced) accumulating Dh + Dl (= production - loss,
// replacing CAMP's long double production/loss sums) and S (the
// magnitudes), with the next four records in flight while the current four
// are processed and each batch's LDS operands read before the batch is
// accumulated.  Phase 3: lane per species, the pieces of a split species
// are added in order, the Jacobian-based estimate J_hist (y - y_hist) +
// f_hist is formed (ELL over the row view), and the precision-loss blend of
// time_derivative.c:58-64 gives f.  s_fb is LDS scratch of fb_lds_doubles.
__host__ __device__ __forceinline__ void cell_deriv(const DevProgram& P, int lane, const double* s_state,
                                           const double* s_rate, const double* y,
                                           const double* Js, const double* Jy,
                                           const double* Jf, bool use_est, double h,
                                           double* out, double* s_fb,
                                           const double* s_slot = nullptr) {
  lane = phase_lane(lane);
  const int nv = P.nv, L = P.fb_rows;
  const double* slot_or_rate = s_slot ? s_slot : s_rate;
  {
    const int4* fr = P.fb_rec + lane;
    // fb_rec has padding rows up to a multiple of 4 plus 4 beyond fb_rows
    int4 a0 = fr[0], a1 = fr[kWave], a2 = fr[2 * kWave], a3 = fr[3 * kWave];
    double Dh = 0.0, Dl = 0.0, S = 0.0;
    for (int j = 0; j < L; j += 4) {
      int4 c0 = a0, c1 = a1, c2 = a2, c3 = a3;
      a0 = fr[(j + 4) * kWave];
      a1 = fr[(j + 5) * kWave];
      a2 = fr[(j + 6) * kWave];
      a3 = fr[(j + 7) * kWave];
      FOps o0 = f_operands(c0, s_state, s_rate, slot_or_rate),
           o1 = f_operands(c1, s_state, s_rate, slot_or_rate),
           o2 = f_operands(c2, s_state, s_rate, slot_or_rate),
           o3 = f_operands(c3, s_state, s_rate, slot_or_rate);
      f_record(c0, o0, h, nv, s_fb, Dh, Dl, S);
      f_record(c1, o1, h, nv, s_fb, Dh, Dl, S);
      f_record(c2, o2, h, nv, s_fb, Dh, Dl, S);
      f_record(c3, o3, h, nv, s_fb, Dh, Dl, S);
    }
  }
  wsync();
  const double* part = s_fb + 2 * nv;
  for (int pass = 0; pass * kWave < nv; ++pass) {
    int slot = pass * kWave + lane;
    int i = slot < nv ? P.fs_species[slot] : -1;
    int4 hd = ld_uniform(P.fl_hdr, pass);
    const int2* er = P.re_ell + hd.z + lane;
    double acc = 0.0;
    if (use_est) {
      // re_ell has padding rows up to a multiple of kEstBatch plus kEstBatch
      // beyond hd.w; a batch's J-history loads are all issued before the
      // first is used
      constexpr int EB = kEstBatch;
      int2 b[EB];
#pragma unroll
      for (int u = 0; u < EB; ++u) b[u] = er[u * kWave];
      for (int j = 0; j < hd.w; j += EB) {
        int2 c[EB];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
          c[u] = b[u];
          b[u] = er[(j + EB + u) * kWave];
        }
        double jv[EB], yv[EB], jyv[EB];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
          int e = c[u].x >= 0 ? c[u].x : 0, col = c[u].x >= 0 ? c[u].y : 0;
          jv[u] = Js[e];
          yv[u] = y[col];
          jyv[u] = Jy[col];
        }
#pragma unroll
        for (int u = 0; u < EB; ++u)
          if (c[u].x >= 0) acc += jv[u] * (yv[u] - jyv[u]);
      }
    }
    if (i < 0) continue;
    int2 fin = P.fb_fin[i];
    double Dhi = 0.0, S = 0.0;
    if (fin.y == 0) {
      Dhi = s_fb[i];
      S = s_fb[nv + i];
    } else if (fin.y > 0) {
      // the partials' (hi, lo) pairs are not normalised: TwoSum each first
      dd D = two_sum(part[3 * fin.x], part[3 * fin.x + 1]);
      S = part[3 * fin.x + 2];
      for (int k = 1; k < fin.y; ++k) {
        const double* q = part + 3 * (fin.x + k);
        D = dd_add(D, two_sum(q[0], q[1]));
        S += q[2];
      }
      Dhi = D.hi;
    }
    double v = 0.0;
    if (S != 0.0) {
      if (use_est) {
        double est = Jf[i] + acc;
        double inv = 1.0 / S;
        double sf = inv / (inv + 1.0e-14 / fabs(Dhi));
        v = sf * Dhi + (1.0 - sf) * est;
      } else {
        v = Dhi;
      }
    }
    out[i] = v;
  }
}

// camp_solver_update_model_state (camp_solver.c:805-833) for one cell:
// returns 1 (wave-uniform) if any y < -SMALL, else copies y into s_state.
__device__ __forceinline__ int cell_set_state(const DevProgram& P, int lane, const double* y,
                                              double* s_state) {
  lane = phase_lane(lane);
  int bad = 0;
  for (int i = lane; i < P.nv; i += kWave) bad |= (y[i] < -1.0e-30);
  bad = wave_or(bad);
  if (bad) return 1;
  for (int i = lane; i < P.nv; i += kWave) {
    double v = y[i];
    s_state[P.dep_state[i]] = v > -1.0e-30 ? v : 1.0e-60;
  }
  wsync();
  return 0;
}

// ---------------------------------------------------------------------
// J(y) of one cell into the solver Jacobian (CSC element order): per-type
// calc_jac_contrib + jacobian_output + jac_map (camp_solver.c:1052-1106,
// Jacobian.c:223-241).
// ---------------------------------------------------------------------
// Mass-action contributions to Jacobian element e of the gather space
// (solver elements, or reaction-pattern elements for multi-phase programs).
__host__ __device__ __forceinline__ dd jac_elem_mass_action(const DevProgram& P, int e, const double* s_state,
                                                   const double* s_k, double h) {
  dd D = {0.0, 0.0};
  for (int q = P.j_ptr[e]; q < P.j_ptr[e + 1]; ++q) {
    int4 a = P.je_a[q];
    double c = P.je_c[q];
    double v = s_k[a.x];
    if (a.y >= 0) v *= s_state[a.y];
    if (a.z >= 0) v *= s_state[a.z];
    if (c == -1.0) {
      D = dd_add(D, dd{-v, 0.0});
    } else {
      int g = P.je_b[q];
      if (g >= 0 && !(-v * s_state[a.w] * c * h <= s_state[g])) continue;
      D = dd_add(D, dd{c * v, 0.0});
    }
  }
  return D;
}

// Operands of one J record.
struct GjOps {
  int e, last, g, h1, h2;
  double coef, k, y1, y2, yw, sg;
};
__device__ __forceinline__ GjOps gj_ops(int4 x, int4 w, const double* sy, const double* sk) {
  GjOps o;
  o.e = x.x & 0x3FFFFFFF;
  o.last = (x.x >> 30) & 1;
  o.g = x.y;
  o.coef = hilo_to_double(x.w, x.z);
  o.h1 = w.y >= 0;
  o.h2 = w.z >= 0;
  o.k = sk[w.x];
  o.y1 = sy[o.h1 ? w.y : 0];
  o.y2 = sy[o.h2 ? w.z : 0];
  o.yw = sy[w.w >= 0 ? w.w : 0];
  o.sg = sy[o.g >= 0 ? o.g : 0];
  return o;
}
// One J record: k * the other reactants; loss records (coefficient -1)
// unguarded, others guarded (rxn_arrhenius.c:186-207 and alike); the
// element's double-double sum stored after its last record.  A guarded
// record that fails adds an exact zero, which leaves the sum as skipping it
// would.
__device__ __forceinline__ void gj_acc(const GjOps& o, double h, int nnz, double* Jc, dd& D) {
  double v = o.k;
  if (o.h1) v *= o.y1;
  if (o.h2) v *= o.y2;
  double add;
  if (o.coef == -1.0) {
    add = -v;
  } else {
    bool keep = o.g < 0 || -v * o.yw * o.coef * h <= o.sg;
    add = keep ? o.coef * v : 0.0;
  }
  D = dd_add(D, dd{add, 0.0});
  if (o.last) {
    if (o.e < nnz) Jc[o.e] = D.hi;
    D = dd{0.0, 0.0};
  }
}

// J(y) of one gas-phase cell: lane j walks its run of whole elements in the
// per-cell J program (runs of near-equal record counts), four records in
// flight while the current four are accumulated, and writes each element
// after its last record.
__device__ __forceinline__ void cell_jac(const DevProgram& P, int lane, const double* s_state,
                                         const double* s_k, double h, double* J) {
  lane = phase_lane(lane);
  vm_drain();
  const int nnz = P.nnz;
  constexpr int RS = 2 * kWave;
  const int4* rec = P.cj_rec + 2 * lane;
  dd D = {0.0, 0.0};
  // the program has 4 rows of padding past cj_rows for the look-ahead
  int4 a0 = rec[0], b0 = rec[1], a1 = rec[RS], b1 = rec[RS + 1];
  int4 a2 = rec[2 * RS], b2 = rec[2 * RS + 1], a3 = rec[3 * RS], b3 = rec[3 * RS + 1];
  for (int r = 0; r < P.cj_rows; r += 4) {
    int4 x0 = a0, w0 = b0, x1 = a1, w1 = b1, x2 = a2, w2 = b2, x3 = a3, w3 = b3;
    const int4* nx = rec + (size_t)(r + 4) * RS;
    a0 = nx[0]; b0 = nx[1]; a1 = nx[RS]; b1 = nx[RS + 1];
    a2 = nx[2 * RS]; b2 = nx[2 * RS + 1]; a3 = nx[3 * RS]; b3 = nx[3 * RS + 1];
    GjOps o0 = gj_ops(x0, w0, s_state, s_k), o1 = gj_ops(x1, w1, s_state, s_k),
          o2 = gj_ops(x2, w2, s_state, s_k), o3 = gj_ops(x3, w3, s_state, s_k);
    gj_acc(o0, h, nnz, J, D);
    gj_acc(o1, h, nnz, J, D);
    gj_acc(o2, h, nnz, J, D);
    gj_acc(o3, h, nnz, J, D);
  }
}

// ---------------------------------------------------------------------
// Sparse LU of M = I - gamma*J (SUNMatScaleAddI) with the static pivot
// order of the program; right-looking, one pivot at a time.  Returns 1 on
// a zero or non-finite pivot (a recoverable linear-setup failure).  lu
// holds lu_nnz values and 64 sink slots after them (scratch).
// ---------------------------------------------------------------------
// One pivot of the factorisation's program: its header {lf_pos row start,
// up_rec row start, L rows | update rows << 16, diagonal position}, the
// lane's first L position and its first kLuRowBatch update records.
struct LuPivot {
  int4 h;
  int l_p;
  uint64_t rec[kLuRowBatch];
};
__device__ __forceinline__ void lu_fetch_pivot(const DevProgram& P, int lane, int k, LuPivot& S) {
  // lf_hdr has two empty pivots past the end; later ones are never used
  S.h = ld_uniform(P.lf_hdr, k < P.nv + 2 ? k : P.nv + 1);
  S.l_p = P.lf_pos[S.h.x + lane];
#pragma unroll
  for (int r = 0; r < kLuRowBatch; ++r) S.rec[r] = P.up_rec[S.h.y + r * kWave + lane];
}
// Eliminate one pivot; returns true on a zero or non-finite pivot.
template <int R>
__device__ __forceinline__ bool lu_pivot(const DevProgram& P, int lane, const LuPivot& S,
                                         double* lu) {
  const int nb = S.h.z >> 16;
  double piv = lu[S.h.w];
  double lval = lu[S.l_p];
  double od[R], oa[R], ob[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    od[r] = lu[(int)(S.rec[r] & 0xFFFF)];
    oa[r] = lu[(int)((S.rec[r] >> 16) & 0xFFFF)];
    ob[r] = lu[(int)((S.rec[r] >> 32) & 0xFFFF)];
  }
  // a zero or non-finite pivot fails the factorisation; the check is
  // folded into a flag tested after the loop, so it does not make the
  // pivot's load a wait ahead of the others (the values computed after a
  // bad pivot are discarded with the failure)
  bool bad = !(piv != 0.0 && isfinite(piv));
  double rinv = 1.0 / piv;
  lu[S.l_p] = lval * rinv;
  for (int r = 1; r < (S.h.z & 0xFFFF); ++r) {
    int p = P.lf_pos[S.h.x + r * kWave + lane];
    lu[p] = lu[p] * rinv;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) lu[(int)(S.rec[r] & 0xFFFF)] = od[r] - (oa[r] * rinv) * ob[r];
  if (nb > R) {
    // the remaining rows (pivots of the dense trailing block) read the
    // scaled L values; they go in batches of R rows whose loads are all
    // issued before the first is used (the program pads them to whole
    // batches: a per-row selection on a loaded record made the compiler
    // wait for each load before issuing the next)
    wsync();
    for (int b = R; b < nb; b += R) {
      uint64_t x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = P.up_rec[S.h.y + (b + r) * kWave + lane];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        od[r] = lu[(int)(x[r] & 0xFFFF)];
        oa[r] = lu[(int)((x[r] >> 16) & 0xFFFF)];
        ob[r] = lu[(int)((x[r] >> 32) & 0xFFFF)];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) lu[(int)(x[r] & 0xFFFF)] = od[r] - oa[r] * ob[r];
    }
  }
  wsync();
  return bad;
}

__device__ __forceinline__ int cell_lu_factor(const DevProgram& P, int lane, const double* J,
                                              double gamma, double* lu) {
  lane = phase_lane(lane);
  vm_drain();
  for (int p = lane; p < P.lu_nnz; p += kWave) lu[p] = 0.0;
  wsync();
  for (int e = lane; e < P.nnz; e += kWave) {
    double v = -gamma * J[e];
    if (P.e_diag[e]) v = v + 1.0;
    lu[P.m_to_lu[e]] = v;
  }
  wsync();
  // Software pipeline over pivots, unrolled by two over two register sets:
  // a pivot's header (scalar), first row of L positions and first R rows of
  // update records are fetched two pivots before it is eliminated, so their
  // latency hides behind the pivot in between (a rotation through copies
  // would make every pivot wait for the newest loads).  The program pads
  // every pivot to at least R rows of update records and its L positions to
  // whole rows of 64, with entries that write the lane's own sink slot
  // lu[lu_nnz + lane], so every load and store is unconditional.  Within one
  // pivot every update has a distinct destination (neither the pivot's L
  // column nor its U row).  A wavefront's LDS accesses execute in program
  // order, so the pivot, the L values and every update operand are read
  // before any of them is written: the updates use the unscaled L value
  // times the pivot's reciprocal -- the same product that is stored as the
  // scaled L value -- and the pivot needs a single wavefront sync.
  constexpr int R = kLuRowBatch;
  bool bad = false;
  LuPivot A, B;
  lu_fetch_pivot(P, lane, 0, A);
  lu_fetch_pivot(P, lane, 1, B);
  for (int k = 0; k < P.nv; k += 2) {
    bad |= lu_pivot<R>(P, lane, A, lu);
    lu_fetch_pivot(P, lane, k + 2, A);
    if (k + 1 < P.nv) bad |= lu_pivot<R>(P, lane, B, lu);
    lu_fetch_pivot(P, lane, k + 3, B);
  }
  return bad ? 1 : 0;
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// One block of 8 pivots of a triangular sweep whose pivot values sit in
// register X of their owner lanes (lane k - k_base, k_base = 64 when X1).
// The lane's two program words per pivot give the LU positions of its rows'
// entries in the pivot's column (the zero slot where a row has none or the
// pivot is outside the sweep), so the block's 16 weights are LDS loads that
// depend on nothing but the program; then the chain, one readlane pair and
// one FMA per row and pivot.  DIR is +1 (forward, L) or -1 (backward, U,
// weights pre-scaled by the row's inverse diagonal).
template <int DIR, bool X1>
__device__ __forceinline__ void tri_block(int kb, int k_base, int4 t0, int4 t1, const double* lu,
                                          double s0, double s1, double& x0, double& x1) {
  constexpr int B = 8;
  const int tw[B] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
  double w0[B], w1[B];
#pragma unroll
  for (int u = 0; u < B; ++u) {
    w0[u] = lu[tw[u] & 0xFFFF] * s0;
    w1[u] = lu[(unsigned)tw[u] >> 16] * s1;
  }
#pragma unroll
  for (int u = 0; u < B; ++u) {
    int k = kb + DIR * u;
    double xk = readlane_d(X1 ? x1 : x0, (k - k_base) & (kWave - 1));
    x0 = fma(-w0[u], xk, x0);
    x1 = fma(-w1[u], xk, x1);
  }
}

// Triangular solves with the solution vector in registers (nv <= 128):
// lane t owns permuted rows t and t+64.  Pivot k's value is broadcast with
// v_readlane (no LDS round trip, no barrier), so the dependency chain per
// pivot is one readlane pair and one FMA.  The LU values each lane needs
// come from the program's position table (the next block's words are
// fetched while a block runs), so they are LDS loads issued ahead of the
// chain; every block runs its 8 steps, pivots outside the sweep having
// zero weights.  The sweeps are split at pivot 64 so each loop reads one
// register, and the backward sweep solves D^-1 U x = D^-1 y, so no division
// sits on the chain.
__device__ __forceinline__ void cell_lu_solve_reg(const DevProgram& P, int lane, double* lu,
                                                  double* b) {
  lane = phase_lane(lane);
  const int nv = P.nv;
  const int r0 = lane, r1 = lane + kWave;
  const bool has0 = r0 < nv, has1 = r1 < nv;
  double x0 = has0 ? b[P.perm[r0]] : 0.0;
  double x1 = has1 ? b[P.perm[r1]] : 0.0;
  // the zero slot, after the LU values and the factorisation's sinks
  if (lane == 0) lu[P.lu_nnz + kWave] = 0.0;
  wsync();
  constexpr int B = 8;
  const int4* tb = P.ts_tbl + 2 * lane;
  int blk = 0;
  int4 t0 = tb[0], t1 = tb[1];
  int4 u0 = tb[2 * kWave], u1 = tb[2 * kWave + 1];
  const int kf = nv - 1;  // forward pivots 0 .. nv-2
  const int split = kf < kWave ? kf : kWave;
  // blocks are fetched two ahead (a block's work is shorter than the
  // memory latency); the table has two blocks of padding past the last, so
  // the look-ahead loads never leave it
  for (int kb = 0; kb < split; kb += B) {
    int4 n0 = tb[2 * kWave * (blk + 2)], n1 = tb[2 * kWave * (blk + 2) + 1];
    tri_block<1, false>(kb, 0, t0, t1, lu, 1.0, 1.0, x0, x1);
    t0 = u0; t1 = u1; u0 = n0; u1 = n1; ++blk;
  }
  for (int kb = split; kb < kf; kb += B) {
    int4 n0 = tb[2 * kWave * (blk + 2)], n1 = tb[2 * kWave * (blk + 2) + 1];
    tri_block<1, true>(kb, kWave, t0, t1, lu, 1.0, 1.0, x0, x1);
    t0 = u0; t1 = u1; u0 = n0; u1 = n1; ++blk;
  }
  double d0 = has0 ? 1.0 / lu[P.diag_pos[r0]] : 0.0;
  double d1 = has1 ? 1.0 / lu[P.diag_pos[r1]] : 0.0;
  x0 *= d0;
  x1 *= d1;
  for (int kb = nv - 1; kb >= kWave; kb -= B) {
    int4 n0 = tb[2 * kWave * (blk + 2)], n1 = tb[2 * kWave * (blk + 2) + 1];
    tri_block<-1, true>(kb, kWave, t0, t1, lu, d0, d1, x0, x1);
    t0 = u0; t1 = u1; u0 = n0; u1 = n1; ++blk;
  }
  for (int kb = (nv - 1 < kWave - 1 ? nv - 1 : kWave - 1); kb >= 0; kb -= B) {
    int4 n0 = tb[2 * kWave * (blk + 2)], n1 = tb[2 * kWave * (blk + 2) + 1];
    tri_block<-1, false>(kb, 0, t0, t1, lu, d0, d1, x0, x1);
    t0 = u0; t1 = u1; u0 = n0; u1 = n1; ++blk;
  }
  wsync();
  if (has0) b[P.perm[r0]] = x0;
  if (has1) b[P.perm[r1]] = x1;
  wsync();
}

// Column sweeps of the triangular solves (nv > 128): one stage per pivot k
// -- its column range, the rows and LU positions of its first 64 entries
// (one per lane), their LU values and, backward, the pivot's diagonal -- and
// the range of the pivot the stage takes next (rn).
struct ColStage {
  int4 r, rn;
  int row, pos;
  double val, dg;
};
// {column start, column end, diagonal position} of pivot k; zero outside
// 0..nv-1
template <bool FWD>
__device__ __forceinline__ int4 col_range(const DevProgram& P, int k) {
  if (k < 0 || k >= P.nv) return int4{0, 0, 0, 0};
  int4 h = ld_uniform(P.lu_hdr, 2 * k);
  int4 h2 = ld_uniform(P.lu_hdr, 2 * k + 1);
  return FWD ? int4{h.x, h.y, h2.x, 0} : int4{h2.y, h2.z, h2.x, 0};
}
// the stage's next pivot: its range (fetched a step earlier, in rn), rows
// and positions
template <bool FWD>
__device__ __forceinline__ void col_fetch_rows(const DevProgram& P, int lane, ColStage& S) {
  const int* pos_a = FWD ? P.l_pos : P.u_pos;
  const int* row_a = FWD ? P.l_row : P.u_row;
  S.r = S.rn;
  // (columns are padded to whole rows of 64 with row -1, position 0)
  const int q = S.r.x + lane;
  S.row = row_a[q];
  S.pos = pos_a[q];
}
template <bool FWD>
__device__ __forceinline__ void col_fetch_vals(const DevProgram& P, int k, const double* lu,
                                               ColStage& S) {
  S.val = lu[S.pos];  // (position 0 for padding: read, never used)
  S.dg = (!FWD && k >= 0 && k < P.nv) ? lu[S.r.z] : 1.0;
}
// Pivot k of a sweep: x_k (divided by the diagonal, backward) updates the
// rows of its column.
template <bool FWD>
__device__ __forceinline__ void col_pivot(const DevProgram& P, int lane, int k, const ColStage& S,
                                          const double* lu, double* w) {
  const int* pos_a = FWD ? P.l_pos : P.u_pos;
  const int* row_a = FWD ? P.l_row : P.u_row;
  double xk = w[k];
  if (!FWD) {
    xk = xk / S.dg;
    wsync();
    if (lane == 0) w[k] = xk;
  }
  if (S.row >= 0) w[S.row] = w[S.row] - S.val * xk;
  for (int q = S.r.x + kWave + lane; q < S.r.y; q += kWave) {
    int i = row_a[q];
    w[i] = w[i] - lu[pos_a[q]] * xk;
  }
  wsync();
}
// One step of the pipelined sweep: the LU values of pivot k + 2 (stage C,
// whose rows were fetched a step earlier), pivot k (stage A), the rows of
// pivot k + 3 into A (its range fetched a step earlier) and the range of
// pivot k + 4 into B, the stage of the next step.
template <bool FWD>
__device__ __forceinline__ void col_step(const DevProgram& P, int lane, int n, const double* lu,
                                         double* w, ColStage& A, ColStage& B, ColStage& C) {
  constexpr int st = FWD ? 1 : -1;
  const int k = FWD ? n : P.nv - 1 - n;
  col_fetch_vals<FWD>(P, k + 2 * st, lu, C);
  if (n < P.nv) col_pivot<FWD>(P, lane, k, A, lu, w);
  col_fetch_rows<FWD>(P, lane, A);
  B.rn = col_range<FWD>(P, k + 4 * st);
}
// The sweep, unrolled by three over three stages so that each stage keeps
// its registers: LU values are fetched two pivots ahead, rows three ahead
// and ranges four ahead, with no copy between stages that would wait for
// the newest load.
template <bool FWD>
__device__ __forceinline__ void col_sweep(const DevProgram& P, int lane, const double* lu,
                                          double* w) {
  constexpr int st = FWD ? 1 : -1;
  const int k0 = FWD ? 0 : P.nv - 1;
  ColStage S0, S1, S2;
  S0.rn = col_range<FWD>(P, k0);
  S1.rn = col_range<FWD>(P, k0 + st);
  S2.rn = col_range<FWD>(P, k0 + 2 * st);
  col_fetch_rows<FWD>(P, lane, S0);
  col_fetch_rows<FWD>(P, lane, S1);
  col_fetch_rows<FWD>(P, lane, S2);
  S0.rn = col_range<FWD>(P, k0 + 3 * st);
  col_fetch_vals<FWD>(P, k0, lu, S0);
  col_fetch_vals<FWD>(P, k0 + st, lu, S1);
  vm_drain();
  for (int n = 0; n < P.nv; n += 3) {
    col_step<FWD>(P, lane, n, lu, w, S0, S1, S2);
    col_step<FWD>(P, lane, n + 1, lu, w, S1, S2, S0);
    col_step<FWD>(P, lane, n + 2, lu, w, S2, S0, S1);
  }
}

// Solve (LU) x = b in place; b in dep order, w a scratch vector (nv).
__device__ __forceinline__ void cell_lu_solve(const DevProgram& P, int lane, double* lu,
                                              double* b, double* w) {
  lane = phase_lane(lane);
  if (P.ts_tbl) {
    cell_lu_solve_reg(P, lane, lu, b);
    return;
  }
  // Column sweeps with the solution in LDS (nv > 128).  Pivot k's L (or U)
  // column: the first 64 entries go one per lane, the rest inline; the
  // program and LU reads run ahead of the pivots (col_sweep), so a pivot
  // waits on LDS only (the LU of multi-phase cells lives in global memory).
  for (int k = lane; k < P.nv; k += kWave) w[k] = b[P.perm[k]];
  wsync();
  col_sweep<true>(P, lane, lu, w);
  col_sweep<false>(P, lane, lu, w);
  for (int k = lane; k < P.nv; k += kWave) b[P.perm[k]] = w[k];
  wsync();
}

}  // namespace camp
