// HIP kernels for the CAMP chemistry hot path on gfx950 (MI355X).
//
//  * plain path (one launch per RHS / Jacobian / LU / solve, the reference's
//    shared-control block semantics): k_rates, k_deriv, k_jac,
//    k_lu_factor, k_lu_solve and the BDF vector kernels (k_vec_*);
//  * per-cell path: k_cell_bdf, a persistent kernel in which each
//    wavefront takes cells from an atomic queue and integrates each with
//    its own BDF order, step size and Newton control (CVODE algorithm).
//
// Data layout in HBM is cell-major (the reference's own layout): a cell's
// state, rate data, Jacobian and LU values are contiguous, so a wavefront
// moves a whole cell with coalesced 512-byte transactions and stages it in
// LDS.
#include "camp_device.hpp"
#include "camp_kernels.hpp"

namespace camp {

// ---------------------------------------------------------------------
// Rate constants for all cells (thread per cell x reaction).
// ---------------------------------------------------------------------
__global__ void k_rates(DevProgram P, int nc, const double* __restrict__ env,
                        double* __restrict__ rxn_env) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nc * P.n_rxn) return;
  int c = (int)(g / P.n_rxn), r = (int)(g % P.n_rxn);
  rate_constant(P, r, env[2 * c], env[2 * c + 1], rxn_env + (size_t)c * P.n_env);
}

// Stage one cell's state and rate data in LDS (wavefront-cooperative).
__device__ __forceinline__ void stage_cell(const DevProgram& P, int lane, const double* state_c,
                                           const double* renv_c, double* s_state, double* s_k) {
  for (int i = lane; i < P.ns; i += kWave) s_state[i] = state_c[i];
  for (int i = lane; i < P.n_env; i += kWave) s_k[i] = renv_c[i];
  wsync();
}

// ---------------------------------------------------------------------
// Plain path: f(y) for all cells.  One wavefront per cell, grid-stride.
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_deriv(DevProgram P, int nc, const double* __restrict__ y,
                                               const double* __restrict__ state,
                                               const double* __restrict__ env,
                                               const double* __restrict__ rxn_env,
                                               const double* __restrict__ Js,
                                               const double* __restrict__ Jy,
                                               const double* __restrict__ Jf,
                                               double* __restrict__ fout, double h, int use_est,
                                               int* __restrict__ fail) {
  extern __shared__ double smem[];
  int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  int wpb = blockDim.x / kWave;
  double* s_state = smem + (size_t)wv * fj_lds_doubles(P);
  double* s_k = s_state + P.ns;
  double* s_rate = s_k + P.n_env + mp_lds_doubles(P);
  MpLds L = mp_lds(P, s_k + P.n_env, s_rate + 2 * (size_t)P.n_terms);
  double* s_fb = s_rate + 2 * (size_t)P.n_terms + mp_jsc_doubles(P);
  for (int c = blockIdx.x * wpb + wv; c < nc; c += gridDim.x * wpb) {
    stage_cell(P, lane, state + (size_t)c * P.ns, rxn_env + (size_t)c * P.n_env, s_state, s_k);
    const double* yc = y + (size_t)c * P.nv;
    if (cell_set_state(P, lane, yc, s_state)) {
      if (lane == 0) atomicOr(fail, 1);
      continue;
    }
    if (P.mp) {
      mp_cell_env(P, lane, env[2 * c], env[2 * c + 1], L);
      mp_f_prologue(P, lane, s_state, s_k, L);
    }
    cell_rates(P, lane, s_state, s_k, s_rate);
    wsync();
    cell_deriv(P, lane, s_state, s_rate, yc, Js + (size_t)c * P.nnz, Jy + (size_t)c * P.nv,
               Jf + (size_t)c * P.nv, use_est != 0, h, fout + (size_t)c * P.nv, s_fb, L.slot);
    wsync();
  }
}

// Plain path: J(y) for all cells into J (nnz per cell, CSC element order).
__global__ void __launch_bounds__(256) k_jac(DevProgram P, int nc, const double* __restrict__ y,
                                             const double* __restrict__ state,
                                             const double* __restrict__ env,
                                             const double* __restrict__ rxn_env,
                                             double* __restrict__ J, double h,
                                             double* __restrict__ Jr) {
  extern __shared__ double smem[];
  int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  int wpb = blockDim.x / kWave;
  double* s_state = smem + (size_t)wv * fj_lds_doubles(P);
  double* s_k = s_state + P.ns;
  double* s_rate = s_k + P.n_env + mp_lds_doubles(P);
  MpLds L = mp_lds(P, s_k + P.n_env, s_rate + 2 * (size_t)P.n_terms);
  for (int c = blockIdx.x * wpb + wv; c < nc; c += gridDim.x * wpb) {
    stage_cell(P, lane, state + (size_t)c * P.ns, rxn_env + (size_t)c * P.n_env, s_state, s_k);
    if (cell_set_state(P, lane, y + (size_t)c * P.nv, s_state)) continue;
    if (P.mp) {
      mp_cell_env(P, lane, env[2 * c], env[2 * c + 1], L);
      mp_refresh(P, lane, s_state, L);  // the f() that Jac() runs first
      mp_refresh(P, lane, s_state, L);  // Jac()'s own update
      mp_j_prologue(P, lane, s_state, s_k, L);
      cell_jac_mp(P, lane, s_state, s_k, h, L, Jr + (size_t)c * P.rnnz, J + (size_t)c * P.nnz);
    } else {
      cell_jac(P, lane, s_state, s_k, h, J + (size_t)c * P.nnz);
    }
    wsync();
  }
}

// Plain path: LU of I - gamma*J for all cells (LU staged in LDS).
__global__ void __launch_bounds__(64) k_lu_factor(DevProgram P, int nc, const double* __restrict__ J,
                                                  double gamma, double* __restrict__ LU,
                                                  int* __restrict__ fail) {
  extern __shared__ double smem[];
  int lane = threadIdx.x;
  double* s_lu = smem;
  for (int c = blockIdx.x; c < nc; c += gridDim.x) {
    int bad = cell_lu_factor(P, lane, J + (size_t)c * P.nnz, gamma, s_lu);
    if (bad) {
      if (lane == 0) atomicOr(fail, 1);
      continue;
    }
    double* out = LU + (size_t)c * P.lu_nnz;
    for (int p = lane; p < P.lu_nnz; p += kWave) out[p] = s_lu[p];
    wsync();
  }
}

// Plain path: solve in place, then scale by s (CVDls gamma-ratio scaling).
__global__ void __launch_bounds__(64) k_lu_solve(DevProgram P, int nc, const double* __restrict__ LU,
                                                 double* __restrict__ b, double s) {
  extern __shared__ double smem[];
  int lane = threadIdx.x;
  double* s_lu = smem;  // LU values, then the factorisation's sinks and the solve's zero slot
  double* s_b = s_lu + P.lu_nnz + kWave + 1;
  double* s_w = s_b + P.nv;
  for (int c = blockIdx.x; c < nc; c += gridDim.x) {
    const double* in = LU + (size_t)c * P.lu_nnz;
    for (int p = lane; p < P.lu_nnz; p += kWave) s_lu[p] = in[p];
    double* bc = b + (size_t)c * P.nv;
    for (int i = lane; i < P.nv; i += kWave) s_b[i] = bc[i];
    wsync();
    cell_lu_solve(P, lane, s_lu, s_b, s_w);
    for (int i = lane; i < P.nv; i += kWave) bc[i] = s != 1.0 ? s_b[i] * s : s_b[i];
    wsync();
  }
}

// ---------------------------------------------------------------------
// Vector kernels of the host-driven (block) BDF.  n = nv * nc entries,
// cell-contiguous like CAMP's N_Vector.
// ---------------------------------------------------------------------
__global__ void k_vec_lin(long n, double a, const double* __restrict__ x, double b,
                          const double* __restrict__ y, double* __restrict__ z) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = a * x[i] + b * y[i];
}

__global__ void k_vec_scale(long n, double a, double* __restrict__ x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = a * x[i];
}

// Nordsieck prediction (sign=+1) or restore (sign=-1): for k=1..q, for
// j=q..k: zn[j-1] += sign * zn[j].  zn is (6, n) contiguous.
__global__ void k_vec_pascal(long n, int q, double* __restrict__ zn, int sign) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double z[6];
  for (int j = 0; j <= q; ++j) z[j] = zn[j * n + i];
  for (int k = 1; k <= q; ++k)
    for (int j = q; j >= k; --j) z[j - 1] = sign > 0 ? z[j - 1] + z[j] : z[j - 1] - z[j];
  for (int j = 0; j < q; ++j) zn[j * n + i] = z[j];
}

// zn[j] *= eta^j for j=1..q (factor accumulated as in cvRescale)
__global__ void k_vec_rescale(long n, int q, double eta, double* __restrict__ zn) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double factor = eta;
  for (int j = 1; j <= q; ++j) {
    zn[j * n + i] *= factor;
    factor *= eta;
  }
}

// zn[j] += l[j] * acor for j = 0..q; after a projection, components
// projected to TINY restart flat there (zn[1..q] cleared)
__global__ void k_vec_complete(long n, int q, Coef6 l6, const double* __restrict__ acor,
                               double* __restrict__ zn, int save, double* __restrict__ last,
                               int projected, const double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = acor[i];
  for (int j = 0; j <= q; ++j) zn[j * n + i] += l6.v[j] * a;
  if (save) zn[5 * n + i] = a;
  if (projected && y[i] == 1.0e-60) {
    zn[i] = 1.0e-60;
    for (int j = 1; j <= q; ++j) zn[j * n + i] = 0.0;
  }
  last[i] = zn[i];
}

// Newton residual: tempv = gamma*ftemp - (rl1*zn1 + acor)
__global__ void k_vec_resid(long n, double rl1, const double* __restrict__ zn1,
                            const double* __restrict__ acor, double gamma,
                            const double* __restrict__ ftemp, double* __restrict__ tempv) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t = rl1 * zn1[i] + acor[i];
  tempv[i] = gamma * ftemp[i] - t;
}

// acor += b; y = zn0 + acor
__global__ void k_vec_newton_update(long n, const double* __restrict__ b, double* __restrict__ acor,
                                    const double* __restrict__ zn0, double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = acor[i] + b[i];
  acor[i] = a;
  y[i] = zn0[i] + a;
}

// ewt = 1 / (rtol*|y| + atol[dep])
__global__ void k_vec_ewt(long n, int nv, double rtol, const double* __restrict__ atol,
                          const double* __restrict__ y, double* __restrict__ ewt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ewt[i] = 1.0 / (rtol * fabs(y[i]) + atol[i % nv]);
}

// y = sum_j s^j zn[j] (Horner, CVodeGetDky with k = 0)
__global__ void k_vec_interp(long n, int q, double s, const double* __restrict__ zn,
                             double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = zn[q * n + i];
  for (int j = q - 1; j >= 0; --j) v = zn[j * n + i] + s * v;
  y[i] = v;
}

// Partial sums of (x*w)^2 per block; out[blockIdx.x]
__global__ void __launch_bounds__(256) k_vec_wsq(long n, const double* __restrict__ x,
                                                 const double* __restrict__ w,
                                                 double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double p = x[i] * w[i];
    s += p * p;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Partial minima per block
__global__ void __launch_bounds__(256) k_vec_min(long n, const double* __restrict__ x,
                                                 double* __restrict__ out) {
  __shared__ double red[4];
  double s = 1.0e308;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    s = fmin(s, x[i]);
  for (int off = kWave / 2; off > 0; off >>= 1) s = fmin(s, __shfl_xor(s, off, kWave));
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}

// Per-block max over i with x_i < -SMALL of -x_i * w_i (-1 when none)
__global__ void __launch_bounds__(256) k_vec_negmax(long n, const double* __restrict__ x,
                                                    const double* __restrict__ w,
                                                    double* __restrict__ out) {
  __shared__ double red[4];
  double s = -1.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (x[i] < -1.0e-30) s = fmax(s, -x[i] * w[i]);
  for (int off = kWave / 2; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, kWave));
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// y_i < -SMALL -> y_i = TINY through the correction (acor_i = TINY - zn0_i)
__global__ void k_vec_project(long n, const double* __restrict__ zn0, double* __restrict__ acor,
                              double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && y[i] < -1.0e-30) {
    acor[i] = 1.0e-60 - zn0[i];
    y[i] = 1.0e-60;
  }
}

// dst[r][j] = src[r * stride][j] for r < n rows of len doubles
__global__ void k_gather_rows(long n, int len, long stride, const double* __restrict__ src,
                              double* __restrict__ dst) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * len) return;
  long r = i / len, j = i - r * len;
  dst[i] = src[r * stride * len + j];
}

// Guess-helper jump search: per block, min over i of the candidate t_star
// (-tmp1/corr) that limits the jump; out[2*b] = min t_star, out[2*b+1] = its
// index (as double), considering only t_star < h_j.
__global__ void __launch_bounds__(256) k_vec_tstar(long n, const double* __restrict__ tmp1,
                                                   const double* __restrict__ corr, double h_j,
                                                   double* __restrict__ out) {
  __shared__ double rv[4], ri[4];
  double best = h_j, bi = -1.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double t_star = -tmp1[i] / corr[i];
    if ((t_star > 0.0 || (t_star == 0.0 && corr[i] < 0.0)) && t_star < best) {
      best = t_star;
      bi = (double)i;
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    double ob = __shfl_xor(best, off, kWave), oi = __shfl_xor(bi, off, kWave);
    if (ob < best || (ob == best && oi >= 0.0 && (bi < 0.0 || oi < bi))) {
      best = ob;
      bi = oi;
    }
  }
  if (threadIdx.x % kWave == 0) {
    rv[threadIdx.x / kWave] = best;
    ri[threadIdx.x / kWave] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = rv[0], i0 = ri[0];
    for (int k = 1; k < 4; ++k)
      if (rv[k] < b || (rv[k] == b && ri[k] >= 0.0 && (i0 < 0.0 || ri[k] < i0))) {
        b = rv[k];
        i0 = ri[k];
      }
    out[2 * blockIdx.x] = b;
    out[2 * blockIdx.x + 1] = i0;
  }
}

// Between the AoS state (ns per cell) and the solver vector (nv per cell):
// constants clamped to TINY in place; gather y = max(state, TINY); scatter
// state = max(y, 0) (camp_solver.c:554-567, 672-683).
__global__ void k_clamp_constants(DevProgram P, int nc, const int* __restrict__ var_type,
                                  double* __restrict__ state) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nc * P.ns) return;
  int i = (int)(g % P.ns);
  double v = state[g];
  if (var_type[i] != 1) state[g] = v > 1.0e-60 ? v : 1.0e-60;
}
__global__ void k_state_gather(DevProgram P, int nc, const double* __restrict__ state,
                               double* __restrict__ y) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nc * P.nv) return;
  int c = (int)(g / P.nv), d = (int)(g % P.nv);
  double v = state[(size_t)c * P.ns + P.dep_state[d]];
  y[g] = v > 1.0e-60 ? v : 1.0e-60;
}
__global__ void k_state_scatter(DevProgram P, int nc, const double* __restrict__ y,
                                double* __restrict__ state) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nc * P.nv) return;
  int c = (int)(g / P.nv), d = (int)(g % P.nv);
  double v = y[g];
  state[(size_t)c * P.ns + P.dep_state[d]] = v > 0.0 ? v : 0.0;
}

__global__ void k_scatter_rates(int n_cells, int n_upd, const int* upd_off, const double* rates,
                                int n_env, double* renv) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n_cells * n_upd) return;
  long c = i / n_upd;
  int u = (int)(i - c * n_upd);
  renv[c * n_env + upd_off[u]] = rates[i];
}

__global__ void k_fill(long n, double v, double* x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

}  // namespace camp
