// Launch interface between the host solver (camp_solver.cpp) and the HIP
// kernels (camp_kernels.hip, camp_cell_bdf.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "camp_device.hpp"
#include "camp_model.hpp"
#include "camp_mp_device.hpp"

namespace camp {

constexpr int kCellStats = 10;  // nst nfe nsetups netf nni ncfn nje flag hu next_h
// shader cycles per cell: f, Jac, LU factor, LU solve, whole solve, and the
// two parts of f (state update + rates, species gather + estimate blend)
constexpr int kCellProf = 7;

// Where a cell's working set lives in the persistent kernel.
//  kPlaceLds     everything in LDS: state, rate data, per-term rates,
//                Nordsieck history (6 vectors), 10 work vectors, the f
//                gather's scratch, the LU factors, the J history and 24
//                slots for the BDF coefficient arrays l, tq and tau (gas
//                phase only: 4 cells per CU for the CB05-scale mechanism);
//  kPlaceGlobal  the Nordsieck history and most work vectors in
//                per-wavefront global scratch (gas phase: all but y, acor,
//                tempv and ftemp; multi-phase: all but y, acor and w) and
//                the J history used in place in HBM, so that more cells fit
//                a CU's LDS (6 for the CB05-scale mechanism; compiled for
//                two wavefronts per SIMD).  Multi-phase programs always run
//                so, with their LU factors, reaction Jacobian and
//                per-instance Jacobian scalars in the scratch as well.
//  kPlaceGlobalLu  the Nordsieck history, 7 of the 10 work vectors and the
//                LU factors (with the factorisation's sink slots and the
//                solve's zero slot) in the scratch (gas phase: 12 cells per
//                CU for the CB05-scale mechanism, compiled for two
//                wavefronts per SIMD).
constexpr int kPlaceLds = 0, kPlaceGlobal = 1, kPlaceGlobalLu = 2, kNumPlacements = 3;
__host__ __device__ inline int cell_placement(const DevProgram& P, int place) {
  return P.mp ? kPlaceGlobal : place;
}
// LU factors in the wavefront's global scratch
__host__ __device__ inline bool cell_lu_global(const DevProgram& P, int place) {
  return P.mp || place == kPlaceGlobalLu;
}
// work vectors of a cell in LDS: all 16 with the history in LDS; for
// gas-phase cells with the history in global memory y, acor and the Newton
// iteration's tempv and ftemp; otherwise y, acor and the solve's w
__host__ __device__ inline int cell_lds_vectors(const DevProgram& P, int place) {
  const int pl = cell_placement(P, place);
  return pl == kPlaceLds ? 16 : (!P.mp && pl == kPlaceGlobal) ? 4 : 3;
}
__host__ __device__ inline size_t cell_bdf_lds_doubles(const DevProgram& P, int place) {
  const bool g = cell_placement(P, place) != kPlaceLds;
  return (size_t)P.ns + P.n_env + mp_lds_doubles(P) + 2 * (size_t)P.n_terms +
         (size_t)cell_lds_vectors(P, place) * P.nv + fb_sink_doubles(P) +
         (cell_lu_global(P, place) ? 0 : (size_t)P.lu_nnz) +
         (g ? 0 : (size_t)P.nnz + 2 * (size_t)P.nv) + 24;
}
// Global scratch doubles per wavefront: the LU factors where they are global
// (with the factorisation's 64 sink slots and the solve's zero slot), for
// multi-phase programs then the reaction Jacobian and the per-instance
// Jacobian scalars, then (kPlaceGlobal, kPlaceGlobalLu) the Nordsieck
// history and the work vectors not in LDS (13 vectors; 12 for gas-phase
// kPlaceGlobal).
__host__ __device__ inline size_t cell_bdf_wave_scratch(const DevProgram& P, int place) {
  if (P.mp) return (size_t)P.lu_nnz + kWave + 1 + P.rnnz + mp_jsc_doubles(P) + 13 * (size_t)P.nv;
  if (place == kPlaceGlobalLu) return (size_t)P.lu_nnz + kWave + 1 + 13 * (size_t)P.nv;
  return place == kPlaceGlobal ? 12 * (size_t)P.nv : 0;
}

// The no-allocation entry point's scratch (a file-scope device array of
// kNoAllocScratchDoubles doubles, 64 MiB): the first kNoAllocHeaderDoubles
// (256 bytes) hold the queue counter, the rest the wavefronts' scratch.
constexpr size_t kNoAllocScratchDoubles = (size_t)8 << 20;
constexpr size_t kNoAllocHeaderDoubles = 32;

struct CellBdfArgs {
  DevProgram P;
  int nc;
  double* state;        // [nc][ns] in/out
  const double* env;    // [nc][2]
  double* rxn_env;      // [nc][n_env] in/out (rate constants written back)
  double *Js, *Jy, *Jf; // per-cell J history
  const int* var_type;  // [ns]
  double rtol, t0, t1;
  int max_steps, max_conv_fails, guess_on;
  int* counter;         // work queue head (zeroed before launch)
  double* stats;        // [nc][kCellStats]
  // Optional indirection used by the request server: queue position q
  // solves cell cell_list[q] (nullptr: cell q) over [t0, cell_t1[cell]]
  // (nullptr: [t0, t1]).  nc is then the length of cell_list.
  const int* cell_list = nullptr;
  const double* cell_t1 = nullptr;
  double* prof = nullptr;  // [cells][kCellProf] when phase timing is on
  int* fail_count = nullptr;  // per cell: incremented on each failed solve
  double* wave_scratch = nullptr;  // [grid waves][cell_bdf_wave_scratch] (kPlaceGlobal)
  // Optional: cell solves cell_nsteps[cell] consecutive timesteps over
  // [t0, t1] in this launch, each exactly as a separate launch would
  // (state written back and restaged between them); nullptr: one.
  const int* cell_nsteps = nullptr;
  int place = kPlaceLds;  // cell_placement(P, place) selects the kernel variant
  // Check of per-cell independence: when non-zero, every wavefront fills its
  // whole working memory (the cell's LDS region and, for multi-phase
  // programs, its global scratch) with quiet NaNs before it stages each
  // cell, so that any value a cell reads before writing shows in its result.
  int poison = 0;
  int bench_reps = 0;  // launch_cell_bdf_bench: repetitions per phase
};

void launch_cell_bdf(const CellBdfArgs& a, int grid, int block, size_t lds, hipStream_t s);
// the measurement variant: per cell, the cycles of f, J, LU and one solve
// into prof[cell][0..3] (gas-phase programs, placements 0 and 1)
void launch_cell_bdf_bench(const CellBdfArgs& a, int grid, int block, size_t lds, hipStream_t s);
// workgroups per CU of the variant (program kind and placement) at this
// block size and LDS, 0 if it does not fit
int cell_bdf_blocks_per_cu(const DevProgram& P, int place, int block, size_t lds);

__global__ void k_rates(DevProgram P, int nc, const double* env, double* rxn_env);
__global__ void k_deriv(DevProgram P, int nc, const double* y, const double* state,
                        const double* env, const double* rxn_env, const double* Js,
                        const double* Jy, const double* Jf, double* fout, double h, int use_est,
                        int* fail);
// Jr: reaction-Jacobian scratch [nc][rnnz] of multi-phase programs (else unused)
__global__ void k_jac(DevProgram P, int nc, const double* y, const double* state,
                      const double* env, const double* rxn_env, double* J, double h, double* Jr);
// LDS doubles per wavefront of k_deriv and k_jac, and wavefronts per block
// (one for multi-phase programs, whose per-cell extras are large)
__host__ __device__ inline size_t fj_lds_doubles(const DevProgram& P) {
  return (size_t)P.ns + P.n_env + mp_lds_doubles(P) + 2 * (size_t)P.n_terms + mp_jsc_doubles(P) +
         fb_lds_doubles(P);
}
inline int fj_waves_per_block(const DevProgram& P) { return P.mp ? 1 : 4; }

// Lane-group f(y) and J(y) of gas-phase programs (lanes per cell = 8, 16
// or 32; camp_kernels.hip) and the compact rate-constant array they read.
__global__ void k_kval(DevProgram P, int nc, const double* rxn_env, double* kval);
// hp (may be null): the time step read from device memory instead of h
void launch_group_deriv(int lanes, const DevProgram& P, int nc, const double* y,
                        const double* state, const double* kval, const double* Js,
                        const double* Jy, const double* Jf, double* fout, double h,
                        const double* hp, int use_est, int* fail, hipStream_t st);
void launch_group_jac(int lanes, const DevProgram& P, int nc, const double* y,
                      const double* state, const double* kval, double* J, double h,
                      hipStream_t st);
// Per-cell LDS stride of the lane-group kernels: the state, (the blended f)
// the gather's running sums 2 nv + 2, the compact rate constants; odd, so
// that the cells of a wavefront start in different LDS banks.
__host__ __device__ inline int group_cell_stride(const DevProgram& P, bool fsum) {
  return (P.ns + (fsum ? 2 * P.nv + 2 : 0) + P.n_k) | 1;
}
inline size_t group_lds_bytes(const DevProgram& P, int lanes, bool fsum) {
  return (size_t)(kWave / lanes) * group_cell_stride(P, fsum) * sizeof(double);
}
// The J kernel also stages its cells' Jacobians (nnz each) and writes them
// out contiguously.
inline size_t group_jac_lds_bytes(const DevProgram& P, int lanes) {
  return group_lds_bytes(P, lanes, false) + (size_t)(kWave / lanes) * P.nnz * sizeof(double);
}
__global__ void k_lu_factor(DevProgram P, int nc, const double* J, double gamma, double* LU,
                            int* fail);
// sp (may be null): the scale read from device memory instead of s
__global__ void k_lu_solve(DevProgram P, int nc, const double* LU, double* b, double s,
                           const double* sp);
__global__ void k_vec_lin(long n, double a, const double* x, double b, const double* y,
                          double* z);
__global__ void k_vec_scale(long n, double a, double* x);
__global__ void k_vec_pascal(long n, int q, double* zn, int sign);
__global__ void k_vec_rescale(long n, int q, double eta, double* zn);
struct Coef6 {
  double v[6];
};
__global__ void k_vec_complete(long n, int q, Coef6 l6, const double* acor, double* zn, int save,
                               double* last, int projected, const double* y);
// ctl (may be null): rl1 = ctl[0], gamma = ctl[1] from device memory
__global__ void k_vec_resid(long n, double rl1, const double* zn1, const double* acor,
                            double gamma, const double* ftemp, double* tempv, const double* ctl);
__global__ void k_vec_newton_update(long n, const double* b, double* acor, const double* zn0,
                                    double* y);
__global__ void k_vec_ewt(long n, int nv, double rtol, const double* atol, const double* y,
                          double* ewt);
__global__ void k_vec_interp(long n, int q, double s, const double* zn, double* y);
__global__ void k_vec_wsq(long n, const double* x, const double* w, double* out);
__global__ void k_vec_min(long n, const double* x, double* out);
__global__ void k_gather_rows(long n, int len, long stride, const double* src, double* dst);
__global__ void k_vec_negmax(long n, const double* x, const double* w, double* out);
__global__ void k_vec_project(long n, const double* zn0, double* acor, double* y);
__global__ void k_vec_tstar(long n, const double* tmp1, const double* corr, double h_j,
                            double* out);
__global__ void k_fill(long n, double v, double* x);
__global__ void k_fill_int(long n, int v, int* x);
__global__ void k_check_jac(DevProgram P, int nc, const double* y, const double* state,
                            const double* env, const double* rxn_env, const double* J, double h_ts,
                            int* cell_fails);
inline size_t check_jac_lds_doubles(const DevProgram& P) {
  return fj_lds_doubles(P) + 4 * (size_t)P.nv + 2 * (size_t)P.nnz;
}
// renv[c][upd_off[u]] = rates[c][u] for c < n_cells, u < n_upd
__global__ void k_scatter_rates(int n_cells, int n_upd, const int* upd_off, const double* rates,
                                int n_env, double* renv);
__global__ void k_clamp_constants(DevProgram P, int nc, const int* var_type, double* state);
__global__ void k_state_gather(DevProgram P, int nc, const double* state, double* y);
__global__ void k_state_scatter(DevProgram P, int nc, const double* y, double* state);

}  // namespace camp
