// Persistent per-cell BDF kernel.
//
// Each wavefront pops cells from an atomic queue and integrates each cell
// from t0 to t1 on its own: variable-order (1-5) variable-step BDF in
// Nordsieck form with modified Newton and WRMS error control -- the CVODE
// algorithm that CAMP's solver_run drives (camp_solver.c:544-693; CVODE:
// Hindmarsh et al., ACM TOMS 31(3), 2005) -- but with order, step size,
// Newton iteration, Jacobian reuse and error test decided per cell instead
// of once for the whole block of cells.  Everything a step touches lives
// in LDS (Nordsieck history, work vectors, rate constants, LU factors);
// only the cell's saved Jacobian history (J_solver, J_state, J_deriv,
// camp_solver.c:1111-1114) is in HBM, because CAMP keeps it across calls.
#include <cstdio>
#include <mutex>

#include "camp_device.hpp"
#include "camp_kernels.hpp"
#include "camp_solver_impl.hpp"

namespace camp {

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

enum { CV_SUCCESS = 0, CV_TOO_MUCH_WORK = -1, CV_TOO_MUCH_ACC = -2, CV_ERR_FAILURE = -3,
       CV_CONV_FAILURE = -4, CV_LSETUP_FAIL = -6, CV_LSOLVE_FAIL = -7, CV_RHSFUNC_FAIL = -8,
       CV_FIRST_RHSFUNC_ERR = -9, CV_REPTD_RHSFUNC_ERR = -10, CV_UNREC_RHSFUNC_ERR = -11,
       CV_ILL_INPUT = -22 };
enum { FIRST_CALL = 1, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN, CONV_FAIL, TRY_AGAIN, RHSFUNC_RECVR };
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

__device__ __forceinline__ double rpowerr(double b, double e) { return b <= 0.0 ? 0.0 : pow(b, e); }
__device__ __forceinline__ double rpoweri(double b, int e) {
  double p = 1.0;
  for (int i = 1; i <= (e < 0 ? -e : e); ++i) p *= b;
  return e < 0 ? 1.0 / p : p;
}

// A wave-uniform double kept in scalar registers: every assignment reads the
// first active lane's value (all lanes hold the same one), so the value
// lives in an SGPR pair between uses instead of occupying two VGPRs for the
// whole solve (the integrator's step-control scalars).
struct UDouble {
  double v;
  UDouble() = default;
  __device__ UDouble(double x) : v(uni(x)) {}
  __device__ UDouble& operator=(double x) {
    v = uni(x);
    return *this;
  }
  __device__ UDouble& operator+=(double x) { return *this = v + x; }
  __device__ UDouble& operator-=(double x) { return *this = v - x; }
  __device__ UDouble& operator*=(double x) { return *this = v * x; }
  __device__ UDouble& operator/=(double x) { return *this = v / x; }
  __device__ operator double() const { return v; }
  __device__ static double uni(double x) {
    int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
  }
};

// MP selects the multi-phase instantiation: gas-phase programs run a kernel
// without the aerosol code (smaller code and register footprint).
template <bool MP>
struct CellCv {
  // the kernel argument's program (kernarg memory: uniform, read-only, so
  // its fields are reloaded with scalar loads rather than kept in registers)
  const DevProgram& P;
  __device__ explicit CellCv(const DevProgram& prog) : P(prog) {}
  int lane, nv;
  double *s_state, *s_k, *s_rate, *znb, *ewt, *acor, *y, *tempv, *ftemp, *last, *acinit, *gh1, *gh2,
      *w, *lu, *s_fb;
  double *Js, *Jy, *Jf;
  MpLds mpl;   // multi-phase extras (LDS); unused for gas-phase programs
  double* Jr;  // reaction Jacobian scratch (global) of multi-phase programs
  UDouble tn, h, hscale, hprime, next_h, eta, etamax, etaq, etaqm1, etaqp1, hu, init_h, rtol;
  double *l, *tq, *tau;  // small coefficient arrays, in LDS (uniform values)
  UDouble rl1, gamma, gammap, gamrat, crate, acnrm, saved_tq5;
  int q, qprime, L, qwait, jcur, use_acinit, guess_on, maxcf, max_steps, projected;
  int nst, nfe, nsetups, nni, ncfn, netf, nje, nstlp, nstlj;
  // optional shader-clock counters (cycles in f, Jac, LU factor, LU solve)
  bool prof;
  long long c_f, c_j, c_lu, c_sol, c_fr, c_fg;
  __device__ __forceinline__ long long tick() const { return prof ? clock64() : 0; }

  __device__ __forceinline__ double* zn(int j) const { return znb + (size_t)j * nv; }

  __device__ __forceinline__ double wrms(const double* x) {
    double s = 0.0;
    for (int i = lane; i < nv; i += kWave) {
      double p = x[i] * ewt[i];
      s += p * p;
    }
    s = wave_sum(s);
    return sqrt(s / nv);
  }

  __device__ __forceinline__ void ewt_set(const double* ycur) {
    for (int i = lane; i < nv; i += kWave)
      ewt[i] = 1.0 / (rtol * fabs(ycur[i]) + P.abs_tol_dep[i]);
    wsync();
  }

  // CAMP f(): camp_solver.c:843-958 for one cell (use_deriv_est = 1)
  __device__ __forceinline__ int f(const double* yv, double* out) {
    nfe++;
    long long t0 = tick();
    double ts = next_h > 0.0 ? next_h : init_h;
    int bad = cell_set_state(P, lane, yv, s_state);
    if (!bad) {
      if (MP) mp_f_prologue(P, lane, s_state, s_k, mpl);
      cell_rates(P, lane, s_state, s_k, s_rate);
      wsync();
      long long t1 = tick();
      if (prof) c_fr += t1 - t0;
      cell_deriv(P, lane, s_state, s_rate, yv, Js, Jy, Jf, true, ts, out, s_fb, mpl.slot);
      wsync();
      if (prof) c_fg += clock64() - t1;
    }
    wsync();
    if (prof) c_f += clock64() - t0;
    return bad;
  }

  // CAMP Jac(): pure f into fy, J into the cell's J_solver, save history
  __device__ __forceinline__ int jac(const double* yv, double* fy) {
    long long t0 = tick();
    double ts = next_h > 0.0 ? next_h : init_h;
    if (cell_set_state(P, lane, yv, s_state)) return 1;
    if (MP) {
      mp_f_prologue(P, lane, s_state, s_k, mpl);
      cell_rates(P, lane, s_state, s_k, s_rate);
      wsync();
      cell_deriv(P, lane, s_state, s_rate, yv, Js, Jy, Jf, false, ts, fy, s_fb, mpl.slot);
      wsync();
      mp_refresh(P, lane, s_state, mpl);
      mp_j_prologue(P, lane, s_state, s_k, mpl);
      cell_jac_mp(P, lane, s_state, s_k, ts, mpl, Jr, Js);
    } else {
      cell_rates(P, lane, s_state, s_k, s_rate);
      wsync();
      cell_deriv(P, lane, s_state, s_rate, yv, Js, Jy, Jf, false, ts, fy, s_fb);
      cell_jac(P, lane, s_state, s_k, ts, Js);
    }
    for (int i = lane; i < nv; i += kWave) {
      Jy[i] = yv[i];
      Jf[i] = fy[i];
    }
    __threadfence_block();
    wsync();
    if (prof) c_j += clock64() - t0;
    return 0;
  }

  // The history updates below work on one row of the Nordsieck array at a
  // time, staged in registers: all of its loads are issued together and each
  // value is stored once, instead of a chain of dependent load-store steps
  // through the array (which may live in global memory).  Every element
  // sees the same operations in the same order as the column-wise loops of
  // CVODE (cvRescale, cvPredict, cvRestore, cvCompleteStep).
  __device__ __forceinline__ void hist_load(int i, double* z) const {
#pragma unroll
    for (int j = 0; j <= QMAX; ++j)
      if (j <= q) z[j] = zn(j)[i];
  }

  __device__ __forceinline__ void rescale() {
    double fac[QMAX + 1];
    double factor = eta;
#pragma unroll
    for (int j = 1; j <= QMAX; ++j) {
      fac[j] = factor;
      factor *= eta;
    }
    for (int i = lane; i < nv; i += kWave) {
      double z[QMAX + 1];
      hist_load(i, z);
#pragma unroll
      for (int j = 1; j <= QMAX; ++j)
        if (j <= q) zn(j)[i] = z[j] * fac[j];
    }
    h = hscale * eta;
    next_h = h;
    hscale = h;
    wsync();
  }

  // zn[j-1] += zn[j] (SIGN 1) or -= (SIGN -1), k = 1..q, j = q..k
  template <int SIGN>
  __device__ __forceinline__ void pascal() {
    for (int i = lane; i < nv; i += kWave) {
      double z[QMAX + 1];
      hist_load(i, z);
#pragma unroll
      for (int k = 1; k <= QMAX; ++k)
#pragma unroll
        for (int j = QMAX; j >= k; --j)
          if (j <= q) z[j - 1] = SIGN > 0 ? z[j - 1] + z[j] : z[j - 1] - z[j];
#pragma unroll
      for (int j = 0; j < QMAX; ++j)
        if (j < q) zn(j)[i] = z[j];
    }
    wsync();
  }

  __device__ __forceinline__ void predict() {
    tn += h;
    pascal<1>();
  }

  __device__ __forceinline__ void restore(double saved_t) {
    tn = saved_t;
    pascal<-1>();
  }

  __device__ __forceinline__ void set_tq(double hsum, double alpha0, double alpha0_hat, double xi_inv,
                         double xistar_inv) {
    double A1 = 1.0 - alpha0_hat + alpha0;
    double A2 = 1.0 + q * A1;
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        double C = xistar_inv / l[q];
        double A3 = alpha0 + 1.0 / q;
        double A4 = alpha0_hat + xi_inv;
        double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      double A5 = alpha0 - (1.0 / (q + 1));
      double A6 = alpha0_hat - xi_inv;
      double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
  }

  __device__ __forceinline__ void set() {
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

  __device__ __forceinline__ void increase_bdf() {
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
    for (int i = lane; i < nv; i += kWave) {
      double v = A1 * zn(QMAX)[i];
      zn(L)[i] = v;
      for (int j = 2; j <= q; ++j) zn(j)[i] += l[j] * v;
    }
    wsync();
  }

  __device__ __forceinline__ void decrease_bdf() {
    double hsum, xi;
    for (int j = 0; j <= QMAX; ++j) l[j] = 0.0;
    l[2] = 1.0;
    hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int i = lane; i < nv; i += kWave)
      for (int j = 2; j < q; ++j) zn(j)[i] += (-l[j]) * zn(q)[i];
    wsync();
  }

  __device__ __forceinline__ void adjust_params() {
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

  // guess_helper, camp_solver.c:1321-1408 (called with h_n = h > 0)
  // y_n_min: the minimum of y_n, which the caller has computed
  __device__ __forceinline__ int guess_helper(double t_n, double h_n, const double* y_n, const double* y_n1,
                              const double* hf, double* tmp1, double* corr, double y_n_min) {
    if (y_n_min > -1.0e-30) return 0;
    double c = 1.0 / h_n;
    for (int i = lane; i < nv; i += kWave) {
      tmp1[i] = y_n1[i];
      corr[i] = c * hf[i];
    }
    wsync();
    double t_0 = t_n - h_n;
    double t_j = 0.0;
    for (int iter = 0; iter < GUESS_MAX_ITER && t_0 + t_j < t_n; iter++) {
      double h_j = t_n - (t_0 + t_j);
      double best = h_j;
      int found = 0;
      for (int i = lane; i < nv; i += kWave) {
        double t_star = -tmp1[i] / corr[i];
        if ((t_star > 0.0 || (t_star == 0.0 && corr[i] < 0.0)) && t_star < best) {
          best = t_star;
          found = 1;
        }
      }
      best = wave_min(best);
      found = wave_or(found);
      h_j = best;
      if (found) h_j *= 0.95 + 0.1 * iter / (double)GUESS_MAX_ITER;
      h_j = t_n < t_0 + t_j + h_j ? t_n - (t_0 + t_j) : h_j;
      for (int i = lane; i < nv; i += kWave) tmp1[i] += h_j * corr[i];
      wsync();
      t_j += h_j;
      if (f(tmp1, corr)) return -1;
    }
    for (int i = lane; i < nv; i += kWave) corr[i] = (tmp1[i] - y_n[i]) * 0.999;
    wsync();
    return 1;
  }

  __device__ __forceinline__ int lsetup(int convfail, const double* ypred, double* fpred) {
    double dgamma = fabs(gamma / gammap - 1.0);
    int jbad = (nst == 0) || (nst > nstlj + MSBJ) ||
               ((convfail == CV_FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) ||
               (convfail == CV_FAIL_OTHER);
    if (!jbad) {
      jcur = 0;
    } else {
      nje++;
      nstlj = nst;
      jcur = 1;
      if (jac(ypred, fpred)) return 1;
    }
    long long t0 = tick();
    int rc = cell_lu_factor(P, lane, Js, gamma, lu);
    if (prof) c_lu += clock64() - t0;
    return rc;
  }

  __device__ __forceinline__ void lsolve(double* b) {
    long long t0 = tick();
    cell_lu_solve(P, lane, lu, b, w);
    if (prof) c_sol += clock64() - t0;
    if (gamrat != 1.0) {
      double s = 2.0 / (1.0 + gamrat);
      for (int i = lane; i < nv; i += kWave) b[i] *= s;
      wsync();
    }
  }

  // Non-negativity safeguard (not in CVODE): CAMP's f() rejects y < -SMALL
  // (camp_solver.c:816-822), and cutting the step is CVODE's only answer,
  // so a species driven to ~0 by a fast loss stalls the integration.
  // Wherever f() would see the iterate y = zn0 + corr, components below
  // -SMALL by at most their tolerance (-y_i * ewt_i <= 1) are projected to
  // TINY (solver_run's floor, camp_solver.c:554-567) through corr.
  // Returns -1 (nothing below -SMALL), 0 (projected) or 1 (a larger
  // excursion: the recoverable f() failure it implies); wave-uniform.
  __device__ __forceinline__ int project(double* corr) {
    double worst = -1.0;
    for (int i = lane; i < nv; i += kWave)
      if (y[i] < -1.0e-30) worst = fmax(worst, -y[i] * ewt[i]);
    worst = -wave_min(-worst);
    if (worst > 1.0) return 1;
    if (worst < 0.0) return -1;
    for (int i = lane; i < nv; i += kWave)
      if (y[i] < -1.0e-30) {
        corr[i] = 1.0e-60 - zn(0)[i];
        y[i] = 1.0e-60;
      }
    wsync();
    return 0;
  }

  __device__ __forceinline__ int newton_iteration() {
    int m = 0;
    double del = 0.0, delp = 0.0;
    for (;;) {
      for (int i = lane; i < nv; i += kWave) {
        double t = rl1 * zn(1)[i] + acor[i];
        tempv[i] = gamma * ftemp[i] - t;
      }
      wsync();
      lsolve(tempv);
      nni++;
      // the correction's WRMS norm, the corrected state and the per-lane sum
      // of the new correction's norm in one pass over the rows (one round
      // trip for the error weights and zn[0] instead of two)
      double s_del = 0.0, s_acor = 0.0;
      for (int i = lane; i < nv; i += kWave) {
        double t = tempv[i], wt = ewt[i];
        double p = t * wt;
        s_del += p * p;
        double a = acor[i] + t;
        acor[i] = a;
        y[i] = zn(0)[i] + a;
        double pa = a * wt;
        s_acor += pa * pa;
      }
      del = sqrt(wave_sum(s_del) / nv);
      wsync();
      if (m > 0) crate = fmax(CRDOWN * crate, del / delp);
      double dcon = del * fmin(1.0, crate) / tq[4];
      if (dcon <= 1.0) {
        int pr = project(acor);
        if (pr > 0) return jcur ? RHSFUNC_RECVR : TRY_AGAIN;
        projected = pr == 0;
        // (without a projection acor is the one summed above)
        acnrm = pr < 0 ? (m == 0 ? del : sqrt(wave_sum(s_acor) / nv)) : wrms(acor);
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

  __device__ __forceinline__ int nls(int nflag) {
    int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
    int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) ||
                    (nst >= nstlp + MSBP) || (fabs(gamrat - 1.0) > DGMAX);
    use_acinit = 0;
    if (guess_on) {
      double mn = 1.0e308;
      for (int i = lane; i < nv; i += kWave) {
        double z = zn(0)[i];
        acinit[i] = z - last[i];
        mn = fmin(mn, z);
      }
      wsync();
      int gh = guess_helper(tn, h, zn(0), last, acinit, gh1, gh2, wave_min(mn));
      if (gh == 1) {
        for (int i = lane; i < nv; i += kWave) acinit[i] = gh2[i];
        wsync();
        use_acinit = 1;
      }
    }
    for (;;) {
      for (int i = lane; i < nv; i += kWave) y[i] = use_acinit ? zn(0)[i] + acinit[i] : zn(0)[i];
      wsync();
      double mn = 1.0e308;
      for (int i = lane; i < nv; i += kWave) mn = fmin(mn, y[i]);
      if (wave_min(mn) < -1.0e-30) {
        if (!use_acinit) {
          for (int i = lane; i < nv; i += kWave) acinit[i] = 0.0;
          wsync();
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
      for (int i = lane; i < nv; i += kWave) acor[i] = use_acinit ? acinit[i] : 0.0;
      wsync();
      int ier = newton_iteration();
      if (ier != TRY_AGAIN) return ier;
      callSetup = 1;
      convfail = CV_FAIL_BAD_J;
    }
  }

  __device__ __forceinline__ int handle_nflag(int* nflagPtr, double saved_t, int* ncfPtr) {
    int nflag = *nflagPtr;
    if (nflag == CV_SUCCESS) return DO_ERROR_TEST;
    ncfn++;
    restore(saved_t);
    if (nflag == CV_LSETUP_FAIL || nflag == CV_LSOLVE_FAIL || nflag == CV_RHSFUNC_FAIL) return nflag;
    (*ncfPtr)++;
    etamax = 1.0;
    if (*ncfPtr == maxcf) {
      if (nflag == CONV_FAIL) return CV_CONV_FAILURE;
      if (nflag == RHSFUNC_RECVR) return CV_REPTD_RHSFUNC_ERR;
    }
    eta = ETACF;
    *nflagPtr = PREV_CONV_FAIL;
    rescale();
    return PREDICT_AGAIN;
  }

  __device__ __forceinline__ int do_error_test(int* nflagPtr, double saved_t, int* nefPtr, double* dsmPtr) {
    double dsm = acnrm * tq[2];
    *dsmPtr = dsm;
    if (dsm <= 1.0) return CV_SUCCESS;
    (*nefPtr)++;
    netf++;
    *nflagPtr = PREV_ERR_FAIL;
    restore(saved_t);
    if (*nefPtr == maxcf) return CV_ERR_FAILURE;
    etamax = 1.0;
    if (*nefPtr <= MXNEF1) {
      eta = 1.0 / (rpowerr(BIAS2 * dsm, 1.0 / L) + ADDON);
      eta = fmax(ETAMIN, fmax(eta, 0.0));
      if (*nefPtr >= SMALL_NEF) eta = fmin(eta, ETAMXF);
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
    if (f(zn(0), tempv)) return CV_UNREC_RHSFUNC_ERR;
    for (int i = lane; i < nv; i += kWave) zn(1)[i] = h * tempv[i];
    wsync();
    return TRY_AGAIN;
  }

  __device__ __forceinline__ void complete_step() {
    nst++;
    hu = h;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if ((q == 1) && (nst > 1)) tau[2] = tau[1];
    tau[1] = h;
    qwait--;
    bool save = (qwait == 1) && (q != QMAX);
    double lj[QMAX + 1];
#pragma unroll
    for (int j = 0; j <= QMAX; ++j) lj[j] = j <= q ? l[j] : 0.0;
    for (int i = lane; i < nv; i += kWave) {
      double a = acor[i];
      double z[QMAX + 1];
      hist_load(i, z);
#pragma unroll
      for (int j = 0; j <= QMAX; ++j)
        if (j <= q) z[j] += lj[j] * a;
      // after a projection, components at TINY restart flat
      if (projected && y[i] == 1.0e-60) {
        z[0] = 1.0e-60;
#pragma unroll
        for (int j = 1; j <= QMAX; ++j) z[j] = 0.0;
      }
#pragma unroll
      for (int j = 0; j <= QMAX; ++j)
        if (j <= q) zn(j)[i] = z[j];
      if (save) zn(QMAX)[i] = a;
      last[i] = z[0];
    }
    if (save) saved_tq5 = tq[5];
    wsync();
  }

  __device__ __forceinline__ void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = fmin(eta, etamax);
      hprime = h * eta;
    }
  }

  __device__ __forceinline__ void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = qwait > 2 ? qwait : 2;
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
      double ddn = wrms(zn(q)) * tq[1];
      etaqm1 = 1.0 / (rpowerr(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    etaqp1 = 0.0;
    if (q != QMAX && saved_tq5 != 0.0) {
      double cquot = (tq[5] / saved_tq5) * rpoweri(h / tau[2], L);
      for (int i = lane; i < nv; i += kWave) tempv[i] = -cquot * zn(QMAX)[i] + acor[i];
      wsync();
      double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (rpowerr(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    double etam = fmax(etaqm1, fmax(etaq, etaqp1));
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
      for (int i = lane; i < nv; i += kWave) zn(QMAX)[i] = acor[i];
      wsync();
    }
    set_eta();
  }

  __device__ __forceinline__ int step() {
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
    for (int i = lane; i < nv; i += kWave) acor[i] *= tq[2];
    wsync();
    return CV_SUCCESS;
  }

  // Measurement: the cycles of f(y), J(y), the LU factorisation (gamma =
  // 10) and one triangular solve at the cell's staged state, averaged over
  // reps repetitions, into pr[0..3].  Integrates nothing.
  __device__ int bench_phases(int reps, double* pr) {
    for (int i = lane; i < nv; i += kWave) y[i] = zn(0)[i];
    wsync();
    next_h = init_h = 1.0;
    if (jac(y, ftemp)) return -101;
    long long cf = 0, cj = 0, clu = 0, cs = 0;
    for (int r = 0; r < reps; ++r) {
      long long t0 = clock64();
      f(y, ftemp);
      long long t1 = clock64();
      jac(y, ftemp);
      long long t2 = clock64();
      if (cell_lu_factor(P, lane, Js, 10.0, lu)) return -102;
      long long t3 = clock64();
      for (int i = lane; i < nv; i += kWave) tempv[i] = ftemp[i];
      wsync();
      long long t4 = clock64();
      lsolve(tempv);
      long long t5 = clock64();
      cf += t1 - t0;
      cj += t2 - t1;
      clu += t3 - t2;
      cs += t5 - t4;
    }
    if (lane == 0) {
      pr[0] = (double)cf / reps;
      pr[1] = (double)cj / reps;
      pr[2] = (double)clu / reps;
      pr[3] = (double)cs / reps;
    }
    return -100;
  }

  // CVodeReInit + CVodeSetInitStep + CVode(tout, CV_NORMAL); zn(0) holds y0
  // on entry, y holds y(tout) on successful exit.
  __device__ __forceinline__ int solve(double t0, double tout) {
    for (int i = lane; i < nv; i += kWave) last[i] = zn(0)[i];
    tn = t0;
    q = 1;
    L = 2;
    qwait = L;
    qprime = 1;
    etamax = ETAMX1;
    hu = 0.0;
    nst = nfe = ncfn = netf = nni = nsetups = nstlp = nje = nstlj = 0;
    saved_tq5 = 0.0;
    crate = 1.0;
    gamrat = 1.0;
    gammap = 1.0;
    jcur = 0;
    projected = 0;
    eta = 1.0;
    for (int i = 0; i < 8; ++i) tau[i] = 0.0;
    for (int i = 0; i < 6; ++i) tq[i] = 0.0;
    init_h = tout - t0;
    next_h = 0.0;
    wsync();
    ewt_set(zn(0));
    if (f(zn(0), zn(1))) return CV_FIRST_RHSFUNC_ERR;
    h = init_h;
    for (int i = lane; i < nv; i += kWave) zn(1)[i] *= h;
    wsync();
    hscale = h;
    hprime = h;
    int nstloc = 0;
    for (;;) {
      next_h = h;
      if (nst > 0) ewt_set(zn(0));
      if (nstloc >= max_steps) return CV_TOO_MUCH_WORK;
      double nrm = wrms(zn(0));
      if (2.220446049250313e-16 * nrm > 1.0) return CV_TOO_MUCH_ACC;
      int kflag = step();
      if (kflag != CV_SUCCESS) return kflag;
      nstloc++;
      if ((tn - tout) * h >= 0.0) {
        double s = (tout - tn) / h;
        for (int i = lane; i < nv; i += kWave) {
          double v = zn(q)[i];
          for (int j = q - 1; j >= 0; --j) v = zn(j)[i] + s * v;
          y[i] = v;
        }
        wsync();
        next_h = hprime;
        return CV_SUCCESS;
      }
    }
  }
};

}  // namespace

// MP: multi-phase program; PL: the placement (kPlaceLds, kPlaceGlobal:
// the Nordsieck history, 7 of the 10 work vectors and the J history in
// global memory, kPlaceGlobalLu: the LU factors as well; multi-phase
// programs run kPlaceGlobal with global LU factors).  Each variant is
// compiled for the wavefronts per SIMD its LDS footprint lets onto a CU
// (1, 2 and 3 for the CB05-scale mechanism), which bounds its registers.
// BENCH: the measurement variant (bench_phases instead of the solve).
template <bool MP, int PL, bool BENCH = false>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(MP || PL == kPlaceLds ? 1 : 2)))
    k_cell_bdf(CellBdfArgs a) {
  extern __shared__ double smem[];
  const DevProgram& P = a.P;
  constexpr bool GH = MP || PL != kPlaceLds;
  constexpr bool GL = MP || PL == kPlaceGlobalLu;
  constexpr int kPlace = MP ? kPlaceGlobal : PL;
  // the wavefront index is uniform: a scalar, so the cell's LDS and scratch
  // pointers derived from it live in scalar registers
  int lane = threadIdx.x % kWave, wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int nv = P.nv;
  size_t per_cell = cell_bdf_lds_doubles(P, kPlace);
  double* base = smem + (size_t)wv * per_cell;
  double* scratch = GH ? a.wave_scratch + ((size_t)blockIdx.x * (blockDim.x / kWave) + wv) *
                                              cell_bdf_wave_scratch(P, kPlace)
                       : nullptr;
  CellCv<MP> cv(P);
  cv.lane = lane;
  cv.nv = nv;
  cv.s_state = base;
  cv.s_k = cv.s_state + P.ns;
  cv.mpl = mp_lds(P, cv.s_k + P.n_env,
                  MP ? scratch + P.lu_nnz + kWave + 1 + P.rnnz : nullptr);
  cv.s_rate = cv.s_k + P.n_env + mp_lds_doubles(P);
  double* v = cv.s_rate + 2 * (size_t)P.n_terms;
  // With GH the Nordsieck history and the work vectors other than y, acor
  // and w live in the wavefront's global scratch (after the LU where it is
  // global, and for multi-phase programs the J_rxn and Jacobian scalars);
  // else all in LDS.
  double* g = MP   ? scratch + P.lu_nnz + kWave + 1 + P.rnnz + mp_jsc_doubles(P)
              : GL ? scratch + P.lu_nnz + kWave + 1
                   : scratch;
  double*& vh = GH ? g : v;
  constexpr bool kNewtonLds = !MP && PL == kPlaceGlobal;
  cv.znb = vh;
  vh += 6 * (size_t)nv;
  cv.last = vh; vh += nv;
  cv.acinit = vh; vh += nv;
  cv.ewt = vh; vh += nv;
  // gas-phase cells with the history in global memory keep the Newton
  // iteration's two work vectors in LDS as well (six cells still fit a CU)
  if (kNewtonLds) {
    cv.tempv = v; v += nv;
    cv.ftemp = v; v += nv;
  } else {
    cv.tempv = vh; vh += nv;
    cv.ftemp = vh; vh += nv;
  }
  cv.gh1 = vh; vh += nv;
  cv.gh2 = vh; vh += nv;
  cv.acor = v; v += nv;
  cv.y = v; v += nv;
  // the column solve's scratch w (nv > 128 only: smaller programs solve in
  // registers) stays in LDS except for gas-phase cells with the history in
  // global memory, whose LDS it would push past six cells per CU
  if (kNewtonLds) {
    cv.w = vh; vh += nv;
  } else {
    cv.w = v; v += nv;
  }
  // Without GH the J history (J_solver, J_state, J_deriv) is staged in LDS
  // for the whole solve; with GH it is used in place in HBM.
  double *h_js = nullptr, *h_jy = nullptr, *h_jf = nullptr;
  // the f gather's scratch follows the LU values (gas: also the LU's sinks)
  if (MP) {
    cv.lu = scratch;
    cv.Jr = scratch + P.lu_nnz + kWave + 1;
    cv.s_fb = v; v += fb_sink_doubles(P);
  } else {
    cv.Jr = nullptr;
    if (GL) {
      cv.lu = scratch;
    } else {
      cv.lu = v;
      v += P.lu_nnz;
    }
    cv.s_fb = v; v += fb_sink_doubles(P);
    if (!GH) {
      h_js = v; v += P.nnz;
      h_jy = v; v += nv;
      h_jf = v; v += nv;
    }
  }
  cv.l = v;
  cv.tq = v + 6;
  cv.tau = v + 12;
  cv.rtol = a.rtol;
  cv.guess_on = a.guess_on;
  cv.maxcf = a.max_conv_fails;
  cv.max_steps = a.max_steps;
  cv.prof = !BENCH && a.prof != nullptr;

  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(a.counter, 1);
    q = __builtin_amdgcn_readfirstlane(q);  // lane 0's ticket, as a scalar
    if (q >= a.nc) break;
    int cell = a.cell_list ? a.cell_list[q] : q;
    double t1 = a.cell_t1 ? a.cell_t1[cell] : a.t1;
    int nsteps = a.cell_nsteps ? a.cell_nsteps[cell] : 1;
    for (int step = 0; step < nsteps; ++step) {
      if (a.poison) {
        const double nan = __builtin_nan("");
        for (size_t i = lane; i < per_cell; i += kWave) base[i] = nan;
        if (GH)
          for (size_t i = lane; i < cell_bdf_wave_scratch(P, kPlace); i += kWave) scratch[i] = nan;
        __threadfence_block();
        wsync();
      }
      double* state_c = a.state + (size_t)cell * P.ns;
      double* renv_c = a.rxn_env + (size_t)cell * P.n_env;
      double* g_js = a.Js + (size_t)cell * P.nnz;
      double* g_jy = a.Jy + (size_t)cell * nv;
      double* g_jf = a.Jf + (size_t)cell * nv;
      if (GH) {
        cv.Js = g_js;
        cv.Jy = g_jy;
        cv.Jf = g_jf;
      } else {
        for (int e = lane; e < P.nnz; e += kWave) h_js[e] = g_js[e];
        for (int i = lane; i < nv; i += kWave) {
          h_jy[i] = g_jy[i];
          h_jf[i] = g_jf[i];
        }
        cv.Js = h_js;
        cv.Jy = h_jy;
        cv.Jf = h_jf;
      }
      // stage state (constants clamped in place) and rate data
      for (int i = lane; i < P.ns; i += kWave) {
        double s = state_c[i];
        cv.s_state[i] = s;
        if (a.var_type[i] != 1) {
          s = s > 1.0e-60 ? s : 1.0e-60;
          cv.s_state[i] = s;
          state_c[i] = s;
        }
      }
      for (int i = lane; i < P.n_env; i += kWave) cv.s_k[i] = renv_c[i];
      wsync();
      // rate constants for the cell's environment (rxn_update_env_state)
      double T = a.env[2 * cell], Pa = a.env[2 * cell + 1];
      cell_rate_constants(P, lane, T, Pa, cv.s_k);
      wsync();
      if (MP) mp_cell_env(P, lane, T, Pa, cv.mpl);
      for (int i = lane; i < P.n_env; i += kWave) renv_c[i] = cv.s_k[i];
      for (int i = lane; i < nv; i += kWave) {
        double s = cv.s_state[P.dep_state[i]];
        cv.zn(0)[i] = s > 1.0e-60 ? s : 1.0e-60;
      }
      wsync();
      cv.c_f = cv.c_j = cv.c_lu = cv.c_sol = cv.c_fr = cv.c_fg = 0;
      long long c_start = cv.tick();
      int flag;
      if constexpr (BENCH)
        flag = cv.bench_phases(a.bench_reps, a.prof + (size_t)cell * kCellProf);
      else
        flag = cv.solve(a.t0, t1);
      if (flag >= 0) {
        for (int i = lane; i < nv; i += kWave) {
          double yv = cv.y[i];
          state_c[P.dep_state[i]] = yv > 0.0 ? yv : 0.0;
          if (MP) cv.s_state[P.dep_state[i]] = yv > 0.0 ? yv : 0.0;
        }
        if (MP) {
          // the post-solve sub_model_calculate (camp_solver.c:685-687) on the
          // final state; the sub models' outputs go back to the state array
          wsync();
          for (int z = lane; z < P.n_zi; z += kWave) {
            mp_zsr_water(P, z, cv.mpl.sub_env, cv.s_state);
            state_c[P.zi_water[z]] = cv.s_state[P.zi_water[z]];
          }
        }
      }
      if (!GH) {
        wsync();
        for (int e = lane; e < P.nnz; e += kWave) g_js[e] = h_js[e];
        for (int i = lane; i < nv; i += kWave) {
          g_jy[i] = h_jy[i];
          g_jf[i] = h_jf[i];
        }
      }
      if (lane == 0) {
        double* st = a.stats + (size_t)cell * kCellStats;
        st[0] = cv.nst; st[1] = cv.nfe; st[2] = cv.nsetups; st[3] = cv.netf;
        st[4] = cv.nni; st[5] = cv.ncfn; st[6] = cv.nje; st[7] = flag;
        st[8] = cv.hu; st[9] = cv.next_h;
        if (a.fail_count && flag < 0) a.fail_count[cell] += 1;
        if (cv.prof) {
          double* pr = a.prof + (size_t)cell * kCellProf;
          pr[0] = (double)cv.c_f; pr[1] = (double)cv.c_j; pr[2] = (double)cv.c_lu;
          pr[3] = (double)cv.c_sol; pr[4] = (double)(clock64() - c_start);
          pr[5] = (double)cv.c_fr; pr[6] = (double)cv.c_fg;
        }
      }
      wsync();
    }
  }
}

// Per-cell LDS beyond 64 KiB per block (multi-phase programs, or several
// cells per workgroup) is opted into per kernel; gfx950 has 160 KiB per CU.
static const void* cell_bdf_fn(const DevProgram& P, int place) {
  if (P.mp) return (const void*)k_cell_bdf<true, kPlaceGlobal>;
  switch (cell_placement(P, place)) {
    case kPlaceGlobal: return (const void*)k_cell_bdf<false, kPlaceGlobal>;
    case kPlaceGlobalLu: return (const void*)k_cell_bdf<false, kPlaceGlobalLu>;
    default: return (const void*)k_cell_bdf<false, kPlaceLds>;
  }
}
static void allow_lds(const void* fn, size_t lds) {
  if (lds > 65536)
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

void launch_cell_bdf(const CellBdfArgs& a, int grid, int block, size_t lds, hipStream_t s) {
  allow_lds(cell_bdf_fn(a.P, a.place), lds);
  if (a.P.mp)
    hipLaunchKernelGGL((k_cell_bdf<true, kPlaceGlobal>), dim3(grid), dim3(block), lds, s, a);
  else if (cell_placement(a.P, a.place) == kPlaceGlobal)
    hipLaunchKernelGGL((k_cell_bdf<false, kPlaceGlobal>), dim3(grid), dim3(block), lds, s, a);
  else if (cell_placement(a.P, a.place) == kPlaceGlobalLu)
    hipLaunchKernelGGL((k_cell_bdf<false, kPlaceGlobalLu>), dim3(grid), dim3(block), lds, s, a);
  else
    hipLaunchKernelGGL((k_cell_bdf<false, kPlaceLds>), dim3(grid), dim3(block), lds, s, a);
}

void launch_cell_bdf_bench(const CellBdfArgs& a, int grid, int block, size_t lds,
                           hipStream_t s) {
  allow_lds(cell_bdf_fn(a.P, a.place), lds);
  if (a.P.mp) {
    allow_lds((const void*)k_cell_bdf<true, kPlaceGlobal, true>, lds);
    hipLaunchKernelGGL((k_cell_bdf<true, kPlaceGlobal, true>), dim3(grid), dim3(block), lds, s,
                       a);
  } else if (cell_placement(a.P, a.place) == kPlaceGlobal) {
    allow_lds((const void*)k_cell_bdf<false, kPlaceGlobal, true>, lds);
    hipLaunchKernelGGL((k_cell_bdf<false, kPlaceGlobal, true>), dim3(grid), dim3(block), lds, s,
                       a);
  } else if (cell_placement(a.P, a.place) == kPlaceLds) {
    allow_lds((const void*)k_cell_bdf<false, kPlaceLds, true>, lds);
    hipLaunchKernelGGL((k_cell_bdf<false, kPlaceLds, true>), dim3(grid), dim3(block), lds, s, a);
  } else {
    die("phase measurement: placements 0 and 1 only");
  }
}

int cell_bdf_blocks_per_cu(const DevProgram& P, int place, int block, size_t lds) {
  if (block > 256 || lds > 160 * 1024) return 0;
  const void* fn = cell_bdf_fn(P, place);
  allow_lds(fn, lds);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, block, lds) != hipSuccess) return 0;
  return n;
}

// ---------------------------------------------------------------------
// The no-allocation entry point's scratch: the queue counter and the
// wavefronts' global scratch of camp_solver_run_device_noalloc.  Host code
// never uses the array's name as an address: each device's copy is looked
// up once, with that device current, with hipGetSymbolAddress.
// ---------------------------------------------------------------------
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

// This device's copy of the scratch (nullptr, with a message, on failure).
double* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    std::fprintf(stderr, "camp_solver_run_device_noalloc: no current HIP device\n");
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      std::fprintf(stderr, "camp_solver_run_device_noalloc: hipGetSymbolAddress failed: %s\n",
                   hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace camp

// camp_solver_run_device without any allocation (see include/camp_solver.h).
extern "C" int camp_solver_run_device_noalloc(void* solver_data, double* d_state, double* d_env,
                                              double t_initial, double t_final) {
  using namespace camp;
  Solver* s = static_cast<Solver*>(solver_data);
  if (!s || !s->m.initialized) return CAMP_NOALLOC_NOT_READY;
  // the solver's own device buffers are made by its first device call
  if (!s->stream) {
    std::fprintf(stderr, "camp_solver_run_device_noalloc: the solver's device buffers do not "
                         "exist yet (any earlier call that uses the device makes them)\n");
    return CAMP_NOALLOC_NOT_READY;
  }
  if (s->mode != CAMP_MODE_PER_CELL || s->profiling || s->eval_jac) {
    std::fprintf(stderr, "camp_solver_run_device_noalloc: per-cell mode only, without phase "
                         "timing or the Jacobian check\n");
    return CAMP_NOALLOC_UNSUPPORTED;
  }
  const size_t per_wave = cell_bdf_wave_scratch(s->P, s->place);
  if (per_wave > kNoAllocScratchDoubles - kNoAllocHeaderDoubles) {
    std::fprintf(stderr, "camp_solver_run_device_noalloc: one wavefront needs %zu doubles of "
                         "scratch, more than the %zu available\n",
                 per_wave, kNoAllocScratchDoubles - kNoAllocHeaderDoubles);
    return CAMP_NOALLOC_TOO_LARGE;
  }
  double* sc = noalloc_scratch();
  if (!sc) return CAMP_NOALLOC_NOT_READY;
  s->upload_renv();
  if (s->m.n_rxn == 0) return CAMP_SOLVER_SUCCESS;
  int rc = s->run_cells(d_state, d_env, t_initial, t_final, sc, kNoAllocScratchDoubles);
  HIP_OK(hipStreamSynchronize(s->stream));
  return rc;
}
