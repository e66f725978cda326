// Multi-phase device functions: the per-cell aerosol state (phase mass,
// average molecular weight, number concentration of every bin), ZSR water
// uptake and its parameter Jacobian, and the SIMPOL phase-transfer and
// aqueous-equilibrium contributions to f(y) and J(y).
//
// Arithmetic follows the reference operation by operation: where it
// computes in double the same double operations are used, and where it
// uses long double (accumulators and rate products) double-double is used,
// which is at least as precise.  Sources restated:
//   aero_phase_solver.c:78-150           phase mass, average MW, volume
//   aero_rep_modal_binned_mass.c:183-266 bin number concentration (BINNED)
//   sub_model_ZSR_aerosol_water.c:216-440  water (JACOBSON pairs), Jacobian
//   rxn_SIMPOL_phase_transfer.c:220-740  environment, f, J
//   rxn_aqueous_equilibrium.c:244-548    environment, f, J
//   util.h:19-68                          transition-regime rate constants
#pragma once

#include "camp_aero.hpp"
#include "camp_device.hpp"

namespace camp {

constexpr int kAeroSlots = 8;           // per phase instance in LDS
constexpr int kSpScalars = 13;          // per SIMPOL instance in LDS
constexpr int kAqScalars = 6;           // per aqueous-equilibrium instance

// --- transition-regime condensation (util.h:19-68), all double -------------
__host__ __device__ __forceinline__ double mp_transition_corr(double mfp, double r, double alpha) {
  double kn = mfp / r;
  return (0.75 * alpha * (1.0 + kn)) / (kn * kn + (1.0 + 0.283 * alpha) * kn + 0.75 * alpha);
}
__host__ __device__ __forceinline__ double mp_cond_rate(double D, double mfp, double r, double alpha) {
  return 4.0 * kPi * r * D * mp_transition_corr(mfp, r, alpha);
}
__host__ __device__ __forceinline__ double mp_dcond_rate_dr(double D, double mfp, double r, double alpha) {
  double kn = mfp / r;
  double dcorr = (0.75 * alpha * mfp * (kn * kn + 2.0 * kn + 1.0 + (0.283 - 0.75) * alpha)) /
                 cr_powi(r * (kn * kn + (1.0 + 0.283 * alpha) * kn + 0.75 * alpha), 2);
  return 4.0 * kPi * D * (mp_transition_corr(mfp, r, alpha) + r * dcorr);
}

// --- aerosol state of one cell (aero_rep_update_state) ---------------------
// s_aero[i*8]: mass, average MW, number concentration (cached, as the
// representation caches them), then the mass and mole sums (dd) of the
// current state, which the MW partial derivatives use
__host__ __device__ __forceinline__ void mp_phase_sums(const DevProgram& P, int i, const double* s_state,
                                              dd* mass_out, dd* moles_out, double* vol_out) {
  int st0 = P.pi_start[i], pm = P.pi_model[i];
  dd mass = {P.min_mass_hi, P.min_mass_lo};
  dd moles = {P.min_moles_hi, P.min_moles_lo};
  double vol = P.min_volume;
  for (int q = P.pm_ptr[pm], k = 0; q < P.pm_ptr[pm + 1]; ++q, ++k) {
    if (!P.pm_counted[q]) continue;
    double x = s_state[st0 + k];
    mass = dd_add(mass, dd_from(x));
    moles = dd_add(moles, dd_div_d(dd_from(x), P.pm_mw[q]));
    vol += x / P.pm_dens[q];
  }
  *mass_out = mass;
  *moles_out = moles;
  *vol_out = vol;
}

__host__ __device__ __forceinline__ void mp_phase_update(const DevProgram& P, int i, const double* s_state,
                                                double* a) {
  dd mass, moles;
  double vol;
  mp_phase_sums(P, i, s_state, &mass, &moles, &vol);
  double mass_d = dd_round(mass);
  a[0] = mass_d;
  a[1] = dd_round(dd_div(dd_from(mass_d), moles));
  double r = P.pi_dp[i] / 2.0;
  a[2] = vol * 3.0 / (4.0 * kPi) / cr_powi(r, 3);
}

// sums of the state after the sub models ran (the getters' partials see it)
__host__ __device__ __forceinline__ void mp_phase_partial_sums(const DevProgram& P, int i,
                                                      const double* s_state, double* a) {
  dd mass, moles;
  double vol;
  mp_phase_sums(P, i, s_state, &mass, &moles, &vol);
  a[3] = mass.hi;
  a[4] = mass.lo;
  a[5] = moles.hi;
  a[6] = moles.lo;
}

// --- ZSR water of one (sub model, phase) item (:216-301) -------------------
__host__ __device__ __forceinline__ void mp_zsr_water(const DevProgram& P, int z, const double* s_sub_env,
                                             double* s_state) {
  // a_w = PPM_TO_RH * state: a double product (both operands double)
  dd aw = dd_from(s_sub_env[P.zi_sub[z]] * s_state[P.zi_gw[z]]);
  double water = 1.0e-30;
  for (int x = P.zi_pair_ptr[z]; x < P.zi_pair_ptr[z + 1]; ++x) {
    double low = P.zp_lowrh[x];
    dd jaw = dd_gt(aw, low) ? aw : dd_from(low);
    double jaw_d = dd_round(jaw);
    dd mol = dd_from(0.0);
    for (int k = 0; k < P.zp_ny[x]; ++k) mol = dd_add(mol, dd_from(P.zp_y[8 * x + k] * cr_powi(jaw_d, k)));
    mol = dd_mul(mol, mol);
    // ion concentrations are double expressions (state / int / MW / 1000)
    dd cat = dd_from(s_state[P.zp_cat[x]] / (double)P.zp_ncat[x] / P.zp_catmw[x] / 1000.0);
    dd an = dd_from(s_state[P.zp_an[x]] / (double)P.zp_nan[x] / P.zp_anmw[x] / 1000.0);
    dd e_ac = dd_from(cr_exp(-100.0 * cat.hi));
    dd e_aa = dd_from(cr_exp(-100.0 * an.hi));
    dd conc = dd_div(dd_add(dd_mul(cat, e_ac), dd_mul(an, e_aa)), dd_add(e_ac, e_aa));
    water = dd_round(dd_add(dd_from(water), dd_mul(dd_div(conc, mol), dd_from(1000.0))));
  }
  s_state[P.zi_water[z]] = water;
}

// --- ZSR parameter Jacobian of one item (:314-~440), added into s_jp -------
__host__ __device__ __forceinline__ void mp_zsr_jac(const DevProgram& P, int z, const double* s_sub_env,
                                           const double* s_state, double* s_jp) {
  double ppm = s_sub_env[P.zi_sub[z]];
  dd aw = dd_from(ppm * s_state[P.zi_gw[z]]);
  for (int x = P.zi_pair_ptr[z]; x < P.zi_pair_ptr[z + 1]; ++x) {
    double low = P.zp_lowrh[x];
    bool above = dd_gt(aw, low);
    dd jaw = above ? aw : dd_from(low);
    double jaw_d = dd_round(jaw);
    dd djaw = above ? dd_from(ppm) : dd_from(0.0);
    const double* Y = P.zp_y + 8 * x;
    dd mol = dd_from(Y[0]), dmol = dd_from(0.0);
    for (int k = 1; k < P.zp_ny[x]; ++k) {
      mol = dd_add(mol, dd_from(Y[k] * cr_powi(jaw_d, k)));
      dmol = dd_add(dmol, dd_from(Y[k] * k * cr_powi(jaw_d, k - 1)));
    }
    dmol = dd_mul(dmol, djaw);
    dd cat = dd_from(s_state[P.zp_cat[x]] / (double)P.zp_ncat[x] / P.zp_catmw[x] / 1000.0);
    double dcat = 1.0 / (double)P.zp_ncat[x] / P.zp_catmw[x] / 1000.0;
    dd an = dd_from(s_state[P.zp_an[x]] / (double)P.zp_nan[x] / P.zp_anmw[x] / 1000.0);
    double dan = 1.0 / (double)P.zp_nan[x] / P.zp_anmw[x] / 1000.0;
    dd e_ac = dd_from(cr_exp(-100.0 * cat.hi));
    dd e_aa = dd_from(cr_exp(-100.0 * an.hi));
    dd conc = dd_div(dd_add(dd_mul(cat, e_ac), dd_mul(an, e_aa)), dd_add(e_ac, e_aa));
    dd den = dd_mul(dd_add(e_ac, e_aa), dd_add(e_ac, e_aa));
    // (1 - ALPHA*anion + ALPHA*cation), ALPHA = -100
    dd kc = dd_add(dd_add(dd_from(1.0), dd_mul(dd_from(100.0), an)), dd_mul(dd_from(-100.0), cat));
    dd ka = dd_add(dd_add(dd_from(1.0), dd_mul(dd_from(100.0), cat)), dd_mul(dd_from(-100.0), an));
    dd dcdc = dd_div(dd_add(dd_mul(e_ac, e_ac), dd_mul(dd_mul(e_ac, e_aa), kc)), den);
    dd dcda = dd_div(dd_add(dd_mul(e_aa, e_aa), dd_mul(dd_mul(e_ac, e_aa), ka)), den);
    double mol_d = dd_round(mol);
    double p3 = cr_powi(mol_d, 3), p2 = cr_powi(mol_d, 2);
    // J[...] += value: each a double += (long double) rounded to double
    // -2.0 * conc / pow(m, 3) * 1000.0 * dm and 1.0 / pow(m, 2) * 1000.0 * dc * dion,
    // evaluated left to right as the reference does
    dd v_gw = dd_mul(dd_mul(dd_div_d(dd_mul(dd_from(-2.0), conc), p3), dd_from(1000.0)), dmol);
    double t2 = 1.0 / p2 * 1000.0;
    dd v_an = dd_mul(dd_mul(dd_from(t2), dcda), dd_from(dan));
    dd v_cat = dd_mul(dd_mul(dd_from(t2), dcdc), dd_from(dcat));
    s_jp[P.zp_jgw[x]] = dd_round(dd_add(dd_from(s_jp[P.zp_jgw[x]]), v_gw));
    s_jp[P.zp_jan[x]] = dd_round(dd_add(dd_from(s_jp[P.zp_jan[x]]), v_an));
    s_jp[P.zp_jcat[x]] = dd_round(dd_add(dd_from(s_jp[P.zp_jcat[x]]), v_cat));
  }
}

// --- per-instance quantities shared by f and J ------------------------------
struct SpInst {
  double nconc, kp, kc, ev0;  // number conc., ppm per kg/m3, k_c(r), k_c*Keq*MW/mass
  dd cond, evap;              // k_c*gas and k_c*Keq*MW/mass*act*aero (long double in CAMP)
};

__host__ __device__ __forceinline__ SpInst mp_simpol_inst(const DevProgram& P, int i, const double* s_state,
                                                 const double* s_k, const double* s_aero) {
  const double* e = s_k + P.sp_env[i];
  const double* f = P.rx_float + P.sp_flt[i];
  const double* a = s_aero + kAeroSlots * P.sp_pi[i];
  SpInst s;
  s.nconc = a[2];
  s.kp = e[3];
  s.kc = mp_cond_rate(f[2], e[0], P.sp_radius[i], e[1]);
  // evap = cond_rate * (EQUIL_CONST * avg_MW / mass): the factor in double
  s.ev0 = e[2] * a[1] / a[0];
  dd evap = dd_mul(dd_from(s.kc), dd_from(s.ev0));
  s.cond = dd_mul(dd_from(s.kc), dd_from(s_state[P.sp_gas[i]]));
  s.evap = dd_mul(evap, dd_from(s_state[P.sp_aero[i]]));
  return s;
}

// f slots of one SIMPOL instance (rxn_SIMPOL_phase_transfer.c:355-399)
__host__ __device__ __forceinline__ void mp_simpol_f(const DevProgram& P, int i, const double* s_state,
                                            const double* s_k, const double* s_aero,
                                            double* s_slot) {
  SpInst s = mp_simpol_inst(P, i, s_state, s_k, s_aero);
  double* o = s_slot + 2 * P.sp_slot[i];
  dd ne = dd_mul(dd_from(s.nconc), s.evap), nc = dd_mul(dd_from(s.nconc), s.cond);
  dd v[4] = {ne, dd_neg(nc), dd_neg(dd_div_d(ne, s.kp)), dd_div_d(nc, s.kp)};
  for (int q = 0; q < 4; ++q) {
    o[2 * q] = v[q].hi;
    o[2 * q + 1] = v[q].lo;
  }
}

// aqueous equilibrium: forward and reverse rates of one instance, or
// skipped (water below the threshold) -- calc_standard_rate (:278-310)
struct AqRates {
  bool skip;
  dd rf, rr;
};
__host__ __device__ __forceinline__ AqRates mp_aqeq_rates(const DevProgram& P, int i,
                                                 const double* s_state, const double* s_k) {
  AqRates r;
  const double* f = P.rx_float + P.aq_flt[i];
  int nr = P.aq_nr[i], np = P.aq_np[i];
  const int* sp = P.aq_sp + 6 * i;
  double w = s_state[P.aq_water[i]];
  r.skip = w < 1.0e-4 * P.aq_small_water[i];
  dd water = dd_from(w);
  dd react = dd_mul(dd_from(s_state[sp[0]]), dd_from(f[5]));
  for (int q = 1; q < nr; ++q) react = dd_mul(react, dd_div(dd_from(s_state[sp[q]] * f[5 + q]), water));
  dd prod = dd_mul(dd_from(s_state[sp[3]]), dd_from(f[5 + nr]));
  for (int q = 1; q < np; ++q)
    prod = dd_mul(prod, dd_div(dd_from(s_state[sp[3 + q]] * f[5 + nr + q]), water));
  r.rf = dd_mul(dd_from(s_k[P.aq_env[i]]), react);
  r.rr = dd_mul(dd_from(f[2]), prod);
  return r;
}

__host__ __device__ __forceinline__ void mp_aqeq_f(const DevProgram& P, int i, const double* s_state,
                                          const double* s_k, double* s_slot) {
  int nr = P.aq_nr[i], np = P.aq_np[i];
  double* o = s_slot + 2 * P.aq_slot[i];
  AqRates r = mp_aqeq_rates(P, i, s_state, s_k);
  dd rate = dd_sub(r.rf, r.rr);
  bool zero = r.skip || (rate.hi == 0.0 && rate.lo == 0.0);
  const double* f = P.rx_float + P.aq_flt[i];
  for (int q = 0; q < nr + np; ++q) {
    double m = f[5 + q];
    dd a = dd_div_d(r.rf, m), b = dd_div_d(r.rr, m);
    if (q < nr) a = dd_neg(a);
    else b = dd_neg(b);
    o[4 * q] = zero ? 0.0 : a.hi;
    o[4 * q + 1] = zero ? 0.0 : a.lo;
    o[4 * q + 2] = zero ? 0.0 : b.hi;
    o[4 * q + 3] = zero ? 0.0 : b.lo;
  }
}


// --- Jacobian scalars of one SIMPOL instance (rxn_SIMPOL_phase_transfer.c:
// 415-~590), kSpScalars doubles: N, kp, kc, evap0 (dd), cond (dd), evap
// (dd), d_cond_d_radius, d_evap_d_radius, d_evap_d_mass, d_evap_d_MW.  evap0
// is k_c*Keq*MW/mass before the activity coefficient and the aerosol factor.
__host__ __device__ __forceinline__ void mp_simpol_jscalars(const DevProgram& P, int i,
                                                   const double* s_state, const double* s_k,
                                                   const double* s_aero, double* o) {
  const double* e = s_k + P.sp_env[i];
  const double* f = P.rx_float + P.sp_flt[i];
  const double* a = s_aero + kAeroSlots * P.sp_pi[i];
  double gas = s_state[P.sp_gas[i]], aero = s_state[P.sp_aero[i]];
  double kc = mp_cond_rate(f[2], e[0], P.sp_radius[i], e[1]);
  dd evap0 = dd_mul(dd_from(kc), dd_from(e[2] * a[1] / a[0]));
  dd cond = dd_mul(dd_from(kc), dd_from(gas));
  dd evap = dd_mul(evap0, dd_from(aero));
  double dcdr = mp_dcond_rate_dr(f[2], e[0], P.sp_radius[i], e[1]) * gas;
  double dedr = dcdr / gas * e[2] * a[1] / a[0] * aero;
  double dedm = dd_round(dd_div_d(dd_neg(evap), a[0]));
  double dedw = dd_round(dd_div_d(evap, a[1]));
  o[0] = a[2];
  o[1] = e[3];
  o[2] = kc;
  o[3] = evap0.hi;
  o[4] = evap0.lo;
  o[5] = cond.hi;
  o[6] = cond.lo;
  o[7] = evap.hi;
  o[8] = evap.lo;
  o[9] = dcdr;
  o[10] = dedr;
  o[11] = dedm;
  o[12] = dedw;
}

// --- Jacobian scalars of one aqueous-equilibrium instance (:403-~430):
// forward (dd), reverse (dd), water, skip flag
__host__ __device__ __forceinline__ void mp_aqeq_jscalars(const DevProgram& P, int i,
                                                 const double* s_state, const double* s_k,
                                                 double* o) {
  const double* f = P.rx_float + P.aq_flt[i];
  int nr = P.aq_nr[i], np = P.aq_np[i];
  const int* sp = P.aq_sp + 6 * i;
  double w = s_state[P.aq_water[i]];
  dd water = dd_from(w);
  dd fwd = dd_from(s_k[P.aq_env[i]]);
  for (int q = 0; q < nr; ++q) fwd = dd_mul(fwd, dd_div(dd_from(s_state[sp[q]] * f[5 + q]), water));
  dd rev = dd_from(f[2]);
  for (int q = 0; q < np; ++q)
    rev = dd_mul(rev, dd_div(dd_from(s_state[sp[3 + q]] * f[5 + nr + q]), water));
  o[0] = fwd.hi;
  o[1] = fwd.lo;
  o[2] = rev.hi;
  o[3] = rev.lo;
  o[4] = w;
  o[5] = w < 1.0e-4 * P.aq_small_water[i] ? 1.0 : 0.0;
}

// --- one special Jacobian record {kind, instance, p1, p2}, c, added to D ----
// PRODUCTION contributions add, LOSS contributions subtract (Jacobian.c:
// 223-241 keeps them in separate long double sums and outputs prod - loss).
__host__ __device__ __forceinline__ dd mp_jac_record(const DevProgram& P, int4 a, double c,
                                            const double* s_state, const double* s_aero,
                                            const double* s_jsc, dd D) {
  const double* aq0 = s_jsc + kSpScalars * P.n_sp;
  if (a.x <= kJsSimpolAP) {
    const double* s = s_jsc + kSpScalars * a.y;
    double N = s[0], kp = s[1], kc = s[2];
    dd evap0 = {s[3], s[4]}, cond = {s[5], s[6]}, evap = {s[7], s[8]};
    double dcdr = s[9], dedr = s[10], dedm = s[11], dedw = s[12];
    switch (a.x) {
      case kJsSimpolGG:  // LOSS N * cond_rate
        D = dd_sub(D, dd_mul(dd_from(N), dd_from(kc)));
        break;
      case kJsSimpolAG:  // PRODUCTION N * cond_rate / kp
        D = dd_add(D, dd_div_d(dd_mul(dd_from(N), dd_from(kc)), kp));
        break;
      case kJsSimpolGA:  // PRODUCTION N * evap_rate * act
        D = dd_add(D, dd_mul(dd_from(N), evap0));
        break;
      case kJsSimpolAA:  // LOSS N * evap_rate * act / kp
        D = dd_sub(D, dd_div_d(dd_mul(dd_from(N), evap0), kp));
        break;
      default: {  // phase element: species a.z of the instance's phase
        int pi = P.sp_pi[a.y];
        const double* ag = s_aero + kAeroSlots * pi;
        double dens = P.pm_dens[a.z], mw = P.pm_mw[a.z];
        double factor = 3.0 / (4.0 * kPi) / cr_powi(P.pi_dp[pi] / 2.0, 3);
        double numj = 1.0 / dens * factor;
        double massj = 1.0;
        dd lm = {ag[3], ag[4]}, mo = {ag[5], ag[6]};
        double mwj = dd_round(dd_div(dd_sub(mo, dd_mul(dd_from(1.0 / mw), lm)), dd_mul(mo, mo)));
        double radj = 0.0;
        if (a.x == kJsSimpolGP) {
          D = dd_add(D, dd_from(N * dedr * radj));
          D = dd_sub(D, dd_from(N * dcdr * radj));
          D = dd_add(D, dd_mul(evap, dd_from(numj)));
          D = dd_sub(D, dd_mul(cond, dd_from(numj)));
          D = dd_add(D, dd_from(N * dedm * massj));
          D = dd_add(D, dd_from(N * dedw * mwj));
        } else {
          D = dd_sub(D, dd_from(N * dedr / kp * radj));
          D = dd_add(D, dd_from(N * dcdr / kp * radj));
          D = dd_sub(D, dd_mul(dd_div_d(evap, kp), dd_from(numj)));
          D = dd_add(D, dd_mul(dd_div_d(cond, kp), dd_from(numj)));
          D = dd_sub(D, dd_from(N * dedm / kp * massj));
          D = dd_sub(D, dd_from(N * dedw / kp * mwj));
        }
      }
    }
    return D;
  }
  const double* s = aq0 + kAqScalars * a.y;
  if (s[5] != 0.0) return D;  // water below threshold: no contribution
  dd fwd = {s[0], s[1]}, rev = {s[2], s[3]};
  dd water = dd_from(s[4]);
  const int* sp = P.aq_sp + 6 * a.y;
  int nr = P.aq_nr[a.y], np = P.aq_np[a.y];
  switch (a.x) {
    case kJsAqRR:
      if (fwd.hi != 0.0) D = dd_sub(D, dd_mul(dd_div_d(dd_div_d(fwd, s_state[sp[a.z]]), c), water));
      break;
    case kJsAqRP:
      if (fwd.hi != 0.0) D = dd_add(D, dd_mul(dd_div_d(dd_div_d(fwd, s_state[sp[a.z]]), c), water));
      break;
    case kJsAqPR:
      if (rev.hi != 0.0) D = dd_add(D, dd_mul(dd_div_d(dd_div_d(rev, s_state[sp[a.z]]), c), water));
      break;
    case kJsAqPP:
      if (rev.hi != 0.0) D = dd_sub(D, dd_mul(dd_div_d(dd_div_d(rev, s_state[sp[a.z]]), c), water));
      break;
    case kJsAqRW:  // LOSS -fwd/m*(nr-1), PRODUCTION -rev/m*(np-1)
      D = dd_sub(D, dd_mul(dd_div_d(dd_neg(fwd), c), dd_from((double)(nr - 1))));
      D = dd_add(D, dd_mul(dd_div_d(dd_neg(rev), c), dd_from((double)(np - 1))));
      break;
    case kJsAqPW:  // PRODUCTION -fwd/m*(nr-1), LOSS -rev/m*(np-1)
      D = dd_add(D, dd_mul(dd_div_d(dd_neg(fwd), c), dd_from((double)(nr - 1))));
      D = dd_sub(D, dd_mul(dd_div_d(dd_neg(rev), c), dd_from((double)(np - 1))));
      break;
  }
  return D;
}

// All special Jacobian records of one reaction element, added to D.
__host__ __device__ __forceinline__ dd mp_jac_records(const DevProgram& P, int e, const double* s_state,
                                             const double* s_aero, const double* s_jsc, dd D) {
  for (int q = P.js_ptr[e]; q < P.js_ptr[e + 1]; ++q)
    D = mp_jac_record(P, P.js_a[q], P.js_c[q], s_state, s_aero, s_jsc, D);
  return D;
}

// --- per-cell LDS of the multi-phase extras ----------------------------------
struct MpLds {
  double *sub_env, *aero, *slot, *jsc, *jp;
};

// Per-cell extras kept in LDS (environment, aerosol state, f slots, the
// parameter Jacobian); the per-instance Jacobian scalars (used only while J
// is assembled) get their own mp_jsc_doubles(P) region, in LDS or in
// per-wavefront global scratch as the caller places it.
__host__ __device__ inline size_t mp_lds_doubles(const DevProgram& P) {
  if (!P.mp) return 0;
  return (size_t)P.n_sub_env + kAeroSlots * (size_t)P.n_pi + 2 * (size_t)P.n_slots + P.pnnz;
}
__host__ __device__ inline size_t mp_jsc_doubles(const DevProgram& P) {
  if (!P.mp) return 0;
  return kSpScalars * (size_t)P.n_sp + kAqScalars * (size_t)P.n_aq;
}

__device__ __forceinline__ MpLds mp_lds(const DevProgram& P, double* base, double* jsc) {
  MpLds L;
  L.sub_env = base;
  L.aero = L.sub_env + P.n_sub_env;
  L.slot = L.aero + kAeroSlots * P.n_pi;
  L.jp = L.slot + 2 * P.n_slots;
  L.jsc = jsc;
  return L;
}

// sub_model_update_env_state for one cell: PPM_TO_RH of every ZSR model
__device__ __forceinline__ void mp_cell_env(const DevProgram& P, int lane, double T, double Pa,
                                            MpLds L) {
  for (int k = lane; k < P.n_sub_env; k += kWave) L.sub_env[k] = mp_ppm_to_rh(T, Pa);
  wsync();
}

// aero_rep_update_state then sub_model_calculate for one cell
// (camp_solver.c:913-917 in f(), 1065-1072 in Jac()): the representation
// caches mass, MW and number concentration from the water currently in the
// state, then ZSR writes the new water.  Jac() runs f() first, so its own
// update sees the water that f() left.
__device__ __forceinline__ void mp_refresh(const DevProgram& P, int lane, double* s_state, MpLds L) {
  lane = phase_lane(lane);
  for (int i = lane; i < P.n_pi; i += kWave) mp_phase_update(P, i, s_state, L.aero + kAeroSlots * i);
  wsync();
  for (int z = lane; z < P.n_zi; z += kWave) mp_zsr_water(P, z, L.sub_env, s_state);
  wsync();
}

// The per-cell start of f(): refresh, then the contributions of the
// multi-phase reactions into their f slots.
__device__ __forceinline__ void mp_f_prologue(const DevProgram& P, int lane, double* s_state,
                                              const double* s_k, MpLds L) {
  lane = phase_lane(lane);
  mp_refresh(P, lane, s_state, L);
  for (int i = lane; i < P.n_sp; i += kWave) mp_simpol_f(P, i, s_state, s_k, L.aero, L.slot);
  for (int i = lane; i < P.n_aq; i += kWave) mp_aqeq_f(P, i, s_state, s_k, L.slot);
  wsync();
}

// After mp_f_prologue and the refresh of Jac(): the sums the partial derivatives use, the
// per-instance Jacobian scalars and the ZSR parameter Jacobian, whose
// element 0 is then set to 1 (the map's dummy, camp_solver.c:1099).
__device__ __forceinline__ void mp_j_prologue(const DevProgram& P, int lane, const double* s_state,
                                              const double* s_k, MpLds L) {
  lane = phase_lane(lane);
  for (int i = lane; i < P.n_pi; i += kWave)
    mp_phase_partial_sums(P, i, s_state, L.aero + kAeroSlots * i);
  for (int q = lane; q < P.pnnz; q += kWave) L.jp[q] = 0.0;
  wsync();
  for (int z = lane; z < P.n_zi; z += kWave) mp_zsr_jac(P, z, L.sub_env, s_state, L.jp);
  for (int i = lane; i < P.n_sp; i += kWave)
    mp_simpol_jscalars(P, i, s_state, s_k, L.aero, L.jsc + kSpScalars * i);
  for (int i = lane; i < P.n_aq; i += kWave)
    mp_aqeq_jscalars(P, i, s_state, s_k, L.jsc + kSpScalars * P.n_sp + kAqScalars * i);
  wsync();
  if (lane == 0) L.jp[0] = 1.0;
  wsync();
}

// One record of a J_rxn row stream (Program::mj_rec) added to D: a
// mass-action record as jac_elem_mass_action adds it, a special record as
// mp_jac_record does, padding not at all.
__device__ __forceinline__ dd mp_jac_stream_record(const DevProgram& P, int4 a, int4 b,
                                                   const double* s_state, const double* s_k,
                                                   double h, const double* s_aero,
                                                   const double* s_jsc, dd D) {
  const double c = hilo_to_double(b.w, b.z);
  if (a.x == -1) {
    double v = s_k[a.y];
    if (a.z >= 0) v *= s_state[a.z];
    if (a.w >= 0) v *= s_state[a.w];
    if (c == -1.0) return dd_add(D, dd{-v, 0.0});
    const int g = b.y;
    if (g >= 0 && !(-v * s_state[b.x] * c * h <= s_state[g])) return D;
    return dd_add(D, dd{c * v, 0.0});
  }
  if (a.x == 0) return D;
  return mp_jac_record(P, a, c, s_state, s_aero, s_jsc, D);
}

// J of a multi-phase cell: the reaction Jacobian into Jr (global scratch,
// rnnz values), then J = J_rxn * J_params over the map (camp_solver.c:
// 1100-1106), one solver element per lane.
__device__ __forceinline__ void cell_jac_mp(const DevProgram& P, int lane, const double* s_state,
                                            const double* s_k, double h, MpLds L, double* Jr,
                                            double* J) {
  lane = phase_lane(lane);
  vm_drain();
  // elements in the program's order, whose rows share their record kinds;
  // each lane reads its element's records as a stream, the next record in
  // flight while the current one is added
  for (int r = 0; r < P.mj_rows; ++r) {
    const int e = P.mj_elist[r * kWave + lane];
    const int2 rh = ld_uniform(P.mj_rhdr, r);
    const int4* rec = P.mj_rec + 2 * ((size_t)rh.x * kWave + lane);
    dd D = {0.0, 0.0};
    int4 a = rec[0], b = rec[1];
    for (int k = 0; k < rh.y; ++k) {
      const int4 ca = a, cb = b;
      const int4* nx = rec + 2 * (size_t)(k + 1) * kWave;
      a = nx[0];
      b = nx[1];
      D = mp_jac_stream_record(P, ca, cb, s_state, s_k, h, L.aero, L.jsc, D);
    }
    if (e >= 0) Jr[e] = D.hi;
  }
  __threadfence_block();
  wsync();
  for (int e = lane; e < P.nnz; e += kWave) {
    double v = 0.0;
    for (int q = P.jm_ptr[e]; q < P.jm_ptr[e + 1]; ++q) v += Jr[P.jm_rxn[q]] * L.jp[P.jm_param[q]];
    J[e] = v;
  }
}
}  // namespace camp
