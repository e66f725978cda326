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

// ---------------------------------------------------------------------
// Lane-group f(y) and J(y) (gas-phase programs): G lanes per cell (G = 8,
// 16 or 32), 64 / G cells per wavefront, so that a wavefront instruction
// serves 64 / G cells and a cell's LDS footprint stays small (its state,
// its rate constants and, for the blended f, per-species running sums).
// Every lane walks a fixed number of records (Program::gf_rec / gj_rec),
// the same for all cells, so a wavefront's record loads touch G addresses;
// a record's rate constant and concentrations come from LDS.  The
// arithmetic of a record is that of the per-wavefront kernels (cell_rates +
// f_record, cell_jac), so the results agree with them up to the grouping of
// split species.
// ---------------------------------------------------------------------
// kval[c][i] = rxn_env[c][k_slot[i]]: the rate constants the terms use,
// compact (n_k per cell), refreshed whenever the rate constants change.
__global__ void k_kval(DevProgram P, int nc, const double* __restrict__ rxn_env,
                       double* __restrict__ kval) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nc * P.n_k) return;
  long c = g / P.n_k;
  int i = (int)(g - c * P.n_k);
  kval[g] = rxn_env[c * P.n_env + P.k_slot[i]];
}

template <int G>
__host__ __device__ constexpr int g_index() {
  return G == 8 ? 0 : G == 16 ? 1 : 2;
}

// Stage the wavefront's cells in LDS at stride group_cell_stride(P, fsum):
// the state variables that are not solver variables from the state, the
// solver variables from y, and (after the f gather's running sums when
// fsum) the compact rate constants.  Loads are coalesced (consecutive cells
// are contiguous).  Returns non-zero in every lane if some cell has
// y < -SMALL (camp_solver.c:816-822).
template <int G, bool FSUM>
__device__ __forceinline__ int g_stage(const DevProgram& P, int nc, int base, int lane,
                                       const double* state, const double* y,
                                       const double* kval, double* s) {
  constexpr int C = kWave / G;
  const int SC = group_cell_stride(P, FSUM);
  const int ncell = nc - base < C ? nc - base : C;
  const int ko = P.ns + (FSUM ? 2 * P.nv + 2 : 0);
  for (int i = lane; i < ncell * P.n_cs; i += kWave) {
    int cs = i / P.n_cs, id = P.cs_ids[i - cs * P.n_cs];
    s[cs * SC + id] = state[(size_t)(base + cs) * P.ns + id];
  }
  const double* kb = kval + (size_t)base * P.n_k;
  for (int i = lane; i < ncell * P.n_k; i += kWave) {
    int cs = i / P.n_k;
    s[cs * SC + ko + (i - cs * P.n_k)] = kb[i];
  }
  const double* yb = y + (size_t)base * P.nv;
  int bad = 0;
  for (int i = lane; i < ncell * P.nv; i += kWave) {
    int cs = i / P.nv, d = i - cs * P.nv;
    double v = yb[i];
    bad |= v < -1.0e-30;
    s[cs * SC + P.dep_state[d]] = v > -1.0e-30 ? v : 1.0e-60;
  }
  wsync();
  return wave_or(bad);
}

// Operands of one f record from the cell's LDS block.
struct GfOps {
  int tgt, last, g, fl, h0, h1, h2;
  double coef, k, y0, y1, y2, sg;
};
__device__ __forceinline__ GfOps gf_ops(int4 x, int4 w, const double* sy, const double* sk) {
  GfOps o;
  o.tgt = x.x & 0x3FFFFFFF;
  o.last = (x.x >> 30) & 1;
  o.g = x.y;
  o.coef = hilo_to_double(x.w, x.z);
  o.fl = (unsigned)w.x >> 28;
  o.h0 = w.y >= 0;
  o.h1 = w.z >= 0;
  o.h2 = w.w >= 0;
  o.k = sk[w.x & 0x0FFFFFFF];
  o.y0 = sy[o.h0 ? w.y : 0];
  o.y1 = sy[o.h1 ? w.z : 0];
  o.y2 = sy[o.h2 ? w.w : 0];
  o.sg = sy[o.g >= 0 ? o.g : 0];
  return o;
}
// The term's rate k * prod(reactants) (cell_rates), then f_record for a
// whole species: guard, error-free product and sum.  After the species'
// last record its sum Dh + Dl (and magnitude sum S) goes to the species:
// with the blend (EST) to the cell's running-sum block in LDS for the
// final pass, else straight to f (the unblended f is Dh + Dl); the sums
// then restart.
template <bool EST>
__device__ __forceinline__ void gf_acc(const GfOps& o, double h, int nv, double* fb, double* fc,
                                       bool valid, double& Dh, double& Dl, double& S) {
  dd rate = {o.k, 0.0};
  if (o.fl & 2) {
    rate = {o.k * o.y0, 0.0};
  } else {
    rate = o.h0 ? dd_mul_d(rate, o.y0) : rate;
    rate = o.h1 ? dd_mul_d(rate, o.y1) : rate;
    rate = o.h2 ? dd_mul_d(rate, o.y2) : rate;
  }
  bool keep = o.g < 0 || -rate.hi * o.coef * h <= o.sg;
  double cf = keep ? o.coef : 0.0;
  double p = rate.hi * cf;
  double e = fma(rate.hi, cf, -p);
  e = fma(rate.lo, cf, e);
  double sm = Dh + p;
  double bb = sm - Dh;
  double err = (Dh - (sm - bb)) + (p - bb);
  Dh = sm;
  Dl = Dl + (err + e);
  if (EST) S += fabs(p);
  if (o.last) {
    if (EST) {
      fb[o.tgt] = Dh + Dl;
      fb[nv + 1 + o.tgt] = S;
    } else if (valid && o.tgt < nv) {
      fc[o.tgt] = Dh + Dl;
    }
    Dh = 0.0;
    Dl = 0.0;
    S = 0.0;
  }
}

template <int G, bool EST>
__global__ void __launch_bounds__(64) k_deriv_g(DevProgram P, int nc, const double* __restrict__ y,
                                                const double* __restrict__ state,
                                                const double* __restrict__ kval,
                                                const double* __restrict__ Js,
                                                const double* __restrict__ Jy,
                                                const double* __restrict__ Jf,
                                                double* __restrict__ fout, double h,
                                                const double* __restrict__ hp,
                                                int* __restrict__ fail) {
  extern __shared__ double smem[];
  constexpr int C = kWave / G;
  if (hp) h = *hp;  // time step from device memory (graph replays)
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave, wpb = blockDim.x / kWave;
  const int cs = lane / G, j = lane % G;
  const int SC = group_cell_stride(P, EST), nv = P.nv;
  double* s = smem + (size_t)wv * C * SC;
  double* sy = s + cs * SC;                         // this lane's cell: state
  double* fb = sy + P.ns;                           // EST: value [nv + 1], magnitudes [nv + 1]
  const double* sk = sy + P.ns + (EST ? 2 * nv + 2 : 0);  // compact rate constants
  const int4* rec = P.gf_rec[g_index<G>()] + 2 * j;
  const int rows = P.g_f_rows[g_index<G>()];
  constexpr int RS = 2 * G;                         // int4 per row
  for (int base = (blockIdx.x * wpb + wv) * C; base < nc; base += gridDim.x * wpb * C) {
    if (g_stage<G, EST>(P, nc, base, lane, state, y, kval, s) && lane == 0) atomicOr(fail, 1);
    const int c = base + cs;
    const bool valid = c < nc;
    double* fc = fout + (size_t)(valid ? c : base) * nv;
    if (!EST && valid)
      for (int q = j; q < P.n_fzero; q += G) fc[P.f_zero[q]] = 0.0;
    double Dh = 0.0, Dl = 0.0, S = 0.0;
    // four rows per iteration, the next four rows' records in flight (the
    // program has four rows of padding past its end)
    int4 a0 = rec[0], b0 = rec[1], a1 = rec[RS], b1 = rec[RS + 1];
    int4 a2 = rec[2 * RS], b2 = rec[2 * RS + 1], a3 = rec[3 * RS], b3 = rec[3 * RS + 1];
    for (int r = 0; r < rows; r += 4) {
      int4 x0 = a0, w0 = b0, x1 = a1, w1 = b1, x2 = a2, w2 = b2, x3 = a3, w3 = b3;
      const int4* nx = rec + (size_t)(r + 4) * RS;
      a0 = nx[0]; b0 = nx[1]; a1 = nx[RS]; b1 = nx[RS + 1];
      a2 = nx[2 * RS]; b2 = nx[2 * RS + 1]; a3 = nx[3 * RS]; b3 = nx[3 * RS + 1];
      GfOps o0 = gf_ops(x0, w0, sy, sk), o1 = gf_ops(x1, w1, sy, sk),
            o2 = gf_ops(x2, w2, sy, sk), o3 = gf_ops(x3, w3, sy, sk);
      gf_acc<EST>(o0, h, nv, fb, fc, valid, Dh, Dl, S);
      gf_acc<EST>(o1, h, nv, fb, fc, valid, Dh, Dl, S);
      gf_acc<EST>(o2, h, nv, fb, fc, valid, Dh, Dl, S);
      gf_acc<EST>(o3, h, nv, fb, fc, valid, Dh, Dl, S);
    }
    if (EST) {
      wsync();
      // per species: the Jacobian-based estimate and the precision-loss
      // blend (time_derivative.c:58-64), lanes of the cell over its species
      if (valid) {
        for (int i = j; i < nv; i += G) {
          double v = 0.0;
          if (P.f_count[i] > 0) {
            const double Dhi = fb[i], Sm = fb[nv + 1 + i];
            if (Sm != 0.0) {
              const double* js = Js + (size_t)c * P.nnz;
              const double* jy = Jy + (size_t)c * nv;
              double acc = 0.0;
              for (int q = P.r_ptr[i]; q < P.r_ptr[i + 1]; ++q) {
                int col = P.r_col[q];
                acc += js[P.r_elem[q]] * (sy[P.dep_state[col]] - jy[col]);
              }
              double est = Jf[(size_t)c * nv + i] + acc;
              double inv = 1.0 / Sm;
              double sf = inv / (inv + 1.0e-14 / fabs(Dhi));
              v = sf * Dhi + (1.0 - sf) * est;
            }
          }
          fc[i] = v;
        }
      }
    }
    wsync();
  }
}

template <int G>
__global__ void __launch_bounds__(64) k_jac_g(DevProgram P, int nc, const double* __restrict__ y,
                                              const double* __restrict__ state,
                                              const double* __restrict__ kval,
                                              double* __restrict__ J, double h) {
  extern __shared__ double smem[];
  constexpr int C = kWave / G;
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave, wpb = blockDim.x / kWave;
  const int cs = lane / G, j = lane % G;
  const int SC = group_cell_stride(P, false);
  const int nnz = P.nnz;
  double* s = smem + (size_t)wv * C * (SC + nnz);
  double* sy = s + cs * SC;
  const double* sk = sy + P.ns;
  // the wavefront's cells' Jacobians, in LDS until they are written out
  double* sJ = s + (size_t)C * SC;
  double* Jc = sJ + (size_t)cs * nnz;
  const int4* rec = P.gj_rec[g_index<G>()] + 2 * j;
  const int rows = P.g_j_rows[g_index<G>()];
  constexpr int RS = 2 * G;
  for (int base = (blockIdx.x * wpb + wv) * C; base < nc; base += gridDim.x * wpb * C) {
    g_stage<G, false>(P, nc, base, lane, state, y, kval, s);
    dd D = {0.0, 0.0};
    int4 a0 = rec[0], b0 = rec[1], a1 = rec[RS], b1 = rec[RS + 1];
    int4 a2 = rec[2 * RS], b2 = rec[2 * RS + 1], a3 = rec[3 * RS], b3 = rec[3 * RS + 1];
    for (int r = 0; r < rows; r += 4) {
      int4 x0 = a0, w0 = b0, x1 = a1, w1 = b1, x2 = a2, w2 = b2, x3 = a3, w3 = b3;
      const int4* nx = rec + (size_t)(r + 4) * RS;
      a0 = nx[0]; b0 = nx[1]; a1 = nx[RS]; b1 = nx[RS + 1];
      a2 = nx[2 * RS]; b2 = nx[2 * RS + 1]; a3 = nx[3 * RS]; b3 = nx[3 * RS + 1];
      GjOps o0 = gj_ops(x0, w0, sy, sk), o1 = gj_ops(x1, w1, sy, sk),
            o2 = gj_ops(x2, w2, sy, sk), o3 = gj_ops(x3, w3, sy, sk);
      gj_acc(o0, h, nnz, Jc, D);
      gj_acc(o1, h, nnz, Jc, D);
      gj_acc(o2, h, nnz, Jc, D);
      gj_acc(o3, h, nnz, Jc, D);
    }
    wsync();
    // the cells base .. base+C-1 are consecutive in J: one contiguous,
    // coalesced write of their Jacobians (every element has a record)
    const int n_cells = nc - base < C ? nc - base : C;
    const size_t n = (size_t)n_cells * nnz;
    double* out = J + (size_t)base * nnz;
    for (size_t i = lane; i < n; i += kWave) out[i] = sJ[i];
    wsync();
  }
}

// Launchers: f and J of all cells with G lanes per cell, one wavefront per
// workgroup (a cell's LDS needs let many such workgroups onto a CU).
template <int G>
static void launch_group_fj(bool jac, const DevProgram& P, int nc, const double* y,
                            const double* state, const double* kval, const double* Js,
                            const double* Jy, const double* Jf, double* out, double h,
                            const double* hp, int use_est, int* fail, hipStream_t st) {
  constexpr int C = kWave / G;
  int blocks = (nc + C - 1) / C;
  blocks = blocks < 1 ? 1 : blocks > (1 << 22) ? (1 << 22) : blocks;
  if (jac) {
    const size_t lds = group_jac_lds_bytes(P, G);
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)k_jac_g<G>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_jac_g<G>), dim3(blocks), dim3(kWave), lds, st, P, nc, y, state, kval,
                       out, h);
  } else if (use_est) {
    hipLaunchKernelGGL((k_deriv_g<G, true>), dim3(blocks), dim3(kWave),
                       group_lds_bytes(P, G, true), st, P, nc, y, state, kval, Js, Jy, Jf, out,
                       h, hp, fail);
  } else {
    hipLaunchKernelGGL((k_deriv_g<G, false>), dim3(blocks), dim3(kWave),
                       group_lds_bytes(P, G, false), st, P, nc, y, state, kval, Js, Jy, Jf, out,
                       h, hp, fail);
  }
}

void launch_group_deriv(int lanes, const DevProgram& P, int nc, const double* y,
                        const double* state, const double* kval, const double* Js,
                        const double* Jy, const double* Jf, double* fout, double h,
                        const double* hp, int use_est, int* fail, hipStream_t st) {
  if (lanes == 8)
    launch_group_fj<8>(false, P, nc, y, state, kval, Js, Jy, Jf, fout, h, hp, use_est, fail, st);
  else if (lanes == 16)
    launch_group_fj<16>(false, P, nc, y, state, kval, Js, Jy, Jf, fout, h, hp, use_est, fail, st);
  else
    launch_group_fj<32>(false, P, nc, y, state, kval, Js, Jy, Jf, fout, h, hp, use_est, fail, st);
}

void launch_group_jac(int lanes, const DevProgram& P, int nc, const double* y,
                      const double* state, const double* kval, double* J, double h,
                      hipStream_t st) {
  if (lanes == 8)
    launch_group_fj<8>(true, P, nc, y, state, kval, nullptr, nullptr, nullptr, J, h, nullptr, 0,
                       nullptr, st);
  else if (lanes == 16)
    launch_group_fj<16>(true, P, nc, y, state, kval, nullptr, nullptr, nullptr, J, h, nullptr, 0,
                        nullptr, st);
  else
    launch_group_fj<32>(true, P, nc, y, state, kval, nullptr, nullptr, nullptr, J, h, nullptr, 0,
                        nullptr, st);
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
                                                 double* __restrict__ b, double s,
                                                 const double* __restrict__ sp) {
  extern __shared__ double smem[];
  int lane = threadIdx.x;
  if (sp) s = *sp;  // scale from device memory (graph replays)
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
                            const double* __restrict__ ftemp, double* __restrict__ tempv,
                            const double* __restrict__ ctl) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (ctl) {  // rl1 and gamma from device memory (graph replays)
    rl1 = ctl[0];
    gamma = ctl[1];
  }
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

// ---------------------------------------------------------------------
// Finite-difference check of the Jacobian, the role of check_Jac
// (camp_solver.c:1146-1257) with its pass rule on the best difference
// quotient: for each column j with y_j >= SMALL (smaller ones are skipped,
// as the reference does) and each element (i, j), the difference quotients
// d of f_i at the steps h = y_j * s, s = 1e-12, 1e-11, ..., 1
// (JAC_CHECK_ADV_MIN..MAX) are Richardson extrapolations 2 D(h/2) - D(h)
// of forward differences D(h) = (f(y + h e_j) - f(y)) / h, each with an
// error estimate err = |D(h/2) - D(h)| (truncation) + 64 eps (4 |f(h/2)| +
// 3 |f(0)| + |f(h)|) / h (rounding); the quotient with the smallest err is
// kept, and the element passes when |J - d| < max(1.2 err, 1e-9) or
// |J - d| / |d| < 1e-4 (JAC_CHECK_GSL_ABS_TOL, JAC_CHECK_GSL_REL_TOL).
// (The reference passes an element at the first step that meets the rule;
// at the smallest steps the rounding error admits almost any value, so the
// best quotient is used instead.)  One wavefront per cell; cell_fails[c] =
// failing elements (-1: f() failed at y).  LDS per wavefront:
// fj_lds_doubles(P) + 4 nv + 2 nnz.
// ---------------------------------------------------------------------
__device__ __forceinline__ int fd_f(const DevProgram& P, int lane, const double* yp, double* s_state,
                                   double* s_k, MpLds& L, double* s_rate, double* s_fb,
                                   double h_ts, double* out) {
  if (cell_set_state(P, lane, yp, s_state)) return 1;
  if (P.mp) mp_f_prologue(P, lane, s_state, s_k, L);
  cell_rates(P, lane, s_state, s_k, s_rate);
  wsync();
  cell_deriv(P, lane, s_state, s_rate, yp, nullptr, nullptr, nullptr, false, h_ts, out, s_fb,
             L.slot);
  wsync();
  return 0;
}

__global__ void __launch_bounds__(64) k_check_jac(DevProgram P, int nc, const double* __restrict__ y,
                                                  const double* __restrict__ state,
                                                  const double* __restrict__ env,
                                                  const double* __restrict__ rxn_env,
                                                  const double* __restrict__ J, double h_ts,
                                                  int* __restrict__ cell_fails) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x;
  double* s_state = smem;
  double* s_k = s_state + P.ns;
  double* s_rate = s_k + P.n_env + mp_lds_doubles(P);
  MpLds L = mp_lds(P, s_k + P.n_env, s_rate + 2 * (size_t)P.n_terms);
  double* s_fb = s_rate + 2 * (size_t)P.n_terms + mp_jsc_doubles(P);
  double* f0 = smem + fj_lds_doubles(P);
  double* f1 = f0 + P.nv;
  double* f2 = f1 + P.nv;
  double* yp = f2 + P.nv;
  double* best_d = yp + P.nv;       // per element: the quotient with the
  double* best_e = best_d + P.nnz;  // smallest error estimate, and that error
  for (int c = blockIdx.x; c < nc; c += gridDim.x) {
    stage_cell(P, lane, state + (size_t)c * P.ns, rxn_env + (size_t)c * P.n_env, s_state, s_k);
    if (P.mp) mp_cell_env(P, lane, env[2 * c], env[2 * c + 1], L);
    const double* yc = y + (size_t)c * P.nv;
    const double* Jc = J + (size_t)c * P.nnz;
    for (int i = lane; i < P.nv; i += kWave) yp[i] = yc[i];
    for (int e = lane; e < P.nnz; e += kWave) {
      best_d[e] = 0.0;
      best_e[e] = __builtin_huge_val();
    }
    wsync();
    if (fd_f(P, lane, yp, s_state, s_k, L, s_rate, s_fb, h_ts, f0)) {
      if (lane == 0) cell_fails[c] = -1;
      continue;
    }
    int nfail = 0;
    for (int j = 0; j < P.nv; ++j) {
      const double x = yc[j];
      if (x < 1.0e-30) continue;
      const int e0 = P.colptr[j], e1 = P.colptr[j + 1];
      double sc = 1.0e-12;
      bool bad = false;
      for (int k = 0; k <= 12 && !bad; ++k, sc *= 10.0) {
        const double h = x * sc;
        if (lane == 0) yp[j] = x + h;
        wsync();
        bad = fd_f(P, lane, yp, s_state, s_k, L, s_rate, s_fb, h_ts, f1);
        if (lane == 0) yp[j] = x + 0.5 * h;
        wsync();
        bad |= fd_f(P, lane, yp, s_state, s_k, L, s_rate, s_fb, h_ts, f2);
        if (lane == 0) yp[j] = x;
        wsync();
        if (bad) break;
        for (int e = e0 + lane; e < e1; e += kWave) {
          const int i = P.rowval[e];
          const double d1 = (f1[i] - f0[i]) / h;
          const double d2 = (f2[i] - f0[i]) / (0.5 * h);
          const double err = fabs(d2 - d1) +
                             64.0 * 2.220446049250313e-16 *
                                 (4.0 * fabs(f2[i]) + 3.0 * fabs(f0[i]) + fabs(f1[i])) / h;
          if (err < best_e[e]) {
            best_e[e] = err;
            best_d[e] = 2.0 * d2 - d1;
          }
        }
        wsync();
      }
      int nf = 0;
      for (int e = e0 + lane; e < e1; e += kWave) {
        const double d = best_d[e];
        const double tol = fmax(1.2 * best_e[e], 1.0e-9);
        const double rel = d != 0.0 ? fabs((Jc[e] - d) / d) : 1.0;
        nf += !(fabs(Jc[e] - d) < tol || rel < 1.0e-4);
      }
      nfail += (int)wave_sum((double)nf);
    }
    if (lane == 0) cell_fails[c] = nfail;
    wsync();
  }
}

__global__ void k_fill_int(long n, int v, int* x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_fill(long n, double v, double* x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

}  // namespace camp
