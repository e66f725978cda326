// Packed layouts of the multi-phase components, as views over the int and
// float parameter arrays CAMP's Fortran layer packs:
//   * aerosol phase            aero_phase_solver.c:26-32
//   * modal/binned-mass rep    aero_reps/aero_rep_modal_binned_mass.c:27-81
//   * SIMPOL phase transfer    rxns/rxn_SIMPOL_phase_transfer.c:31-74
//   * aqueous equilibrium      rxns/rxn_aqueous_equilibrium.c:27-68
//   * ZSR aerosol water        sub_models/sub_model_ZSR_aerosol_water.c:30-68
// Views index the data after the type slot.  Species ids come 1-based from
// the packers and are returned 0-based.  Every view is a plain accessor set,
// usable on the host and (for the fixed-size parts) on the device.
#pragma once

#include <hip/hip_runtime.h>

namespace camp {

constexpr int kRxnAqueousEquilibrium = 7;   // rxn_solver.c:26
constexpr int kRxnSimpolPhaseTransfer = 10; // rxn_solver.c:28
constexpr int kAeroRepModalBinnedMass = 2;  // aero_rep_solver.c:20
constexpr int kSubModelZsrAerosolWater = 2; // sub_model_solver.c:17
constexpr int kConstant = 2, kPssa = 3;     // camp_common.h:35-37
constexpr int kBinned = 1, kModal = 2;      // aero_rep_modal_binned_mass.c:18-19
constexpr int kZsrJacobson = 1;             // sub_model_ZSR_aerosol_water.c:20

// kinds of the special Jacobian records of the multi-phase kernel program:
// SIMPOL (gas/gas, aerosol/gas, gas/aerosol, aerosol/aerosol, gas or
// aerosol row against a phase species) and aqueous equilibrium (reactant or
// product row against a reactant, product or water column)
enum SpecialJac {
  kJsSimpolGG = 1, kJsSimpolAG, kJsSimpolGA, kJsSimpolAA, kJsSimpolGP, kJsSimpolAP,
  kJsAqRR, kJsAqRP, kJsAqPR, kJsAqPP, kJsAqRW, kJsAqPW
};

struct PhaseView {  // int: [n, type...]; float: [MW..., density...]
  const int* d;
  const double* f;
  __host__ __device__ int n() const { return d[0]; }
  __host__ __device__ int type(int i) const { return d[1 + i]; }
  __host__ __device__ bool counted(int i) const {
    int t = type(i);
    return t == 1 || t == kConstant || t == kPssa;
  }
  __host__ __device__ double mw(int i) const { return f[i]; }
  __host__ __device__ double density(int i) const { return f[n() + i]; }
};

struct BinnedRepView {
  int* d;
  double* f;
  int nsec() const { return d[0]; }
  int iloc(int s) const { return d[4 + s] - 1; }
  int floc(int s) const { return d[4 + nsec() + s] - 1; }
  int type(int s) const { return d[iloc(s)]; }
  int nbin(int s) const { return d[iloc(s) + 1]; }
  int nph(int s) const { return d[iloc(s) + 2]; }
  int phase_state(int s, int p, int b) const { return d[iloc(s) + 3 + b * nph(s) + p] - 1; }
  int phase_model(int s, int p, int b) const {
    return d[iloc(s) + 3 + nbin(s) * nph(s) + b * nph(s) + p] - 1;
  }
  int& phase_njac(int s, int p, int b) const {
    return d[iloc(s) + 3 + 2 * nbin(s) * nph(s) + b * nph(s) + p];
  }
  double bin_dp(int s, int b) const { return f[floc(s) + 3 * b]; }
  double radius(int s, int b) const { return f[floc(s) + 3 * b + 2]; }
};

struct SimpolView {
  int* d;
  double* f;
  int n() const { return d[0]; }
  int gas() const { return d[1] - 1; }
  int aero(int i) const { return d[2 + i] - 1; }
  int act(int i) const { return d[2 + n() + i] - 1; }
  int phase(int i) const { return d[2 + 2 * n() + i] - 1; }
  int rep(int i) const { return d[2 + 3 * n() + i] - 1; }
  int& deriv_id(int x) const { return d[2 + 4 * n() + x]; }
  int& gas_act_jac(int i) const { return d[3 + 5 * n() + i]; }
  int& aero_act_jac(int i) const { return d[3 + 6 * n() + i]; }
  int& jac_id(int x) const { return d[3 + 7 * n() + x]; }
  int iloc(int i) const { return d[4 + 10 * n() + i] - 1; }
  int floc(int i) const { return d[4 + 11 * n() + i] - 1; }
  int njac(int i) const { return d[iloc(i)]; }
  int& phase_jac(int i, int s, int k) const { return d[iloc(i) + 1 + s * njac(i) + k]; }
  // float: DELTA_H, DELTA_S, DIFF_COEFF, PRE_C_AVG, B1..B4, CONV, MW
};

struct AqEqView {
  int* d;
  double* f;
  int nr() const { return d[0]; }
  int np() const { return d[1]; }
  int nph() const { return d[2]; }
  int react(int x) const { return d[3 + x] - 1; }
  int prod(int x) const { return d[3 + nr() * nph() + x] - 1; }
  int water(int p) const { return d[3 + (nr() + np()) * nph() + p] - 1; }
  int act(int p) const { return d[3 + (nr() + np() + 1) * nph() + p] - 1; }
  int& deriv_id(int x) const { return d[3 + (nr() + np() + 2) * nph() + x]; }
  int& jac_id(int x) const { return d[3 + (2 * (nr() + np()) + 2) * nph() + x]; }
  double mftm(int x) const { return f[5 + x]; }
  double& small_water(int p) const { return f[5 + 2 * nr() + 2 * np() + p]; }
  double& small_conc(int p) const { return f[5 + 2 * nr() + 2 * np() + nph() + p]; }
  // float: A, C, k_reverse, (scratch) water, (scratch) activity coeff
};

struct ZsrView {
  int* d;
  double* f;
  int nph() const { return d[0]; }
  int gas_water() const { return d[1] - 1; }
  int npair() const { return d[2]; }
  int phase(int p) const { return d[5 + p] - 1; }  // state id of the phase's water
  int piloc(int x) const { return d[5 + nph() + x] - 1; }
  int pfloc(int x) const { return d[5 + nph() + npair() + x] - 1; }
  int type(int x) const { return d[piloc(x)]; }
  int n_cation(int x) const { return d[piloc(x) + 1]; }
  int n_anion(int x) const { return d[piloc(x) + 2]; }
  int cation(int x) const { return d[piloc(x) + 3]; }  // offset from the water id
  int anion(int x) const { return d[piloc(x) + 4]; }
  int ny(int x) const { return d[piloc(x) + 5]; }
  int& gw_jac(int p, int x) const { return d[piloc(x) + 6 + p]; }
  int& cation_jac(int p, int x) const { return d[piloc(x) + 6 + nph() + p]; }
  int& anion_jac(int p, int x) const { return d[piloc(x) + 6 + 2 * nph() + p]; }
  double low_rh(int x) const { return f[pfloc(x)]; }
  double cation_mw(int x) const { return f[pfloc(x) + 1]; }
  double anion_mw(int x) const { return f[pfloc(x) + 2]; }
  double y(int x, int k) const { return f[pfloc(x) + 3 + k]; }
};

}  // namespace camp
