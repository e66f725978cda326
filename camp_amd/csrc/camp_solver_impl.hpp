// Host-side solver object shared by the C ABI (camp_solver.cpp) and the
// request server (camp_server.cpp): the Model, its kernel program uploaded to
// HBM, per-cell persistent buffers and the launch of the per-cell kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/camp_solver.h"
#include "camp_kernels.hpp"
#include "camp_model.hpp"

namespace camp {

[[noreturn]] inline void die(const char* what) {
  std::fprintf(stderr, "\n\nERROR %s\n\n", what);
  std::fflush(stderr);
  std::exit(EXIT_FAILURE);
}

#define HIP_OK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      std::exit(EXIT_FAILURE);                                                            \
    }                                                                                     \
  } while (0)

template <class T>
inline T* dev_upload(const std::vector<T>& v, std::vector<void*>& owned) {
  T* p = nullptr;
  size_t n = std::max<size_t>(v.size(), 1);
  HIP_OK(hipMalloc(&p, n * sizeof(T)));
  if (!v.empty()) HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  owned.push_back(p);
  return p;
}

template <class T>
inline T* dev_alloc(size_t n, std::vector<void*>& owned) {
  T* p = nullptr;
  HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  HIP_OK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  owned.push_back(p);
  return p;
}

struct BlockStats {
  long nst = 0, nfe = 0, nsetups = 0, netf = 0, nni = 0, ncfn = 0, nje = 0;
  int flag = 0;
  double hu = 0.0, next_h = 0.0;
};

struct Solver {
  Model m;
  int mode = CAMP_MODE_PER_CELL;
  int cells_per_block = 1;
  // per-cell kernel variant (cell_placement): history in global memory, the
  // fastest for the CB05-scale mechanism (six cells per CU instead of four)
  int place = kPlaceGlobal;
  int guess_on = 1;
  bool renv_dirty = true;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_kernel_ms = 0.0;
  std::vector<void*> owned;  // device allocations
  DevProgram P{};
  // per-cell persistent buffers
  double *d_state = nullptr, *d_env = nullptr, *d_renv = nullptr;
  double *d_Js = nullptr, *d_Jy = nullptr, *d_Jf = nullptr, *d_stats = nullptr;
  int *d_counter = nullptr, *d_fail = nullptr, *d_var_type = nullptr;
  // Queue order of the per-cell kernel (a permutation of the cells; empty:
  // index order): the caller's (explicit_order), else longest first by the
  // previous call's step counts.  Results do not depend on it; it only
  // decides which cells the launch starts first.
  std::vector<int> h_order;
  int* d_order = nullptr;
  bool explicit_order = false;
  bool order_valid = false;  // d_order holds an order for the next launch
  // per-cell statistics copied back after each per-cell launch (sized at
  // set-up, so that a launch allocates no host memory either)
  std::vector<double> h_stats;
  int poison = 0;  // CellBdfArgs::poison for per-cell runs
  // solver_set_eval_jac: finite-difference check of every evaluated Jacobian
  bool eval_jac = false;
  long jac_eval_fails = 0;
  int* d_jac_fails = nullptr;
  // camp_eval_deriv / camp_eval_jac with the lane-group kernels (the plain
  // path's) instead of the per-wavefront ones (the per-cell kernel's code)
  bool eval_group = false;
  // lanes per cell of the block path's f and J kernels (8, 16 or 32)
  int f_lanes = 16, j_lanes = 16;
  // block path: Newton iterations replayed as HIP graphs (BlockCv)
  bool use_graphs = true;
  hipGraphExec_t graph_a = nullptr, graph_b = nullptr;
  const double* graph_b_state = nullptr;
  int graph_b_lanes = 0;
  double *h_ctl = nullptr, *h_red_pin = nullptr, *d_ctl = nullptr;  // pinned / device
  int* h_fail_pin = nullptr;
  // block-path buffers
  double *d_zn = nullptr, *d_ewt = nullptr, *d_acor = nullptr, *d_y = nullptr,
         *d_tempv = nullptr, *d_ftemp = nullptr, *d_last = nullptr, *d_acinit = nullptr,
         *d_gh1 = nullptr, *d_gh2 = nullptr, *d_LU = nullptr, *d_red = nullptr,
         *d_jtmp = nullptr;
  // compact rate constants of the lane-group f/J kernels (gas-phase
  // programs): [nc][n_k], refreshed with the rate constants
  double* d_kval = nullptr;
  double* kval() {
    if (!d_kval) d_kval = dev_alloc<double>((size_t)m.nc * std::max(P.n_k, 1), owned);
    return d_kval;
  }
  // f(y) and J(y) of all cells on the solver's stream (lane-group kernels for
  // gas-phase programs, one wavefront per cell for multi-phase programs);
  // refresh_kval after the rate constants change
  void refresh_kval();
  void launch_deriv(const double* y, const double* st, const double* env, double* out, double h,
                    int use_est, int* fail);
  void launch_jac(const double* y, const double* st, const double* env, double* J, double h,
                  double* jr);
  std::vector<double> h_red;
  BlockStats bstats;
  bool stats_from_cells = false;
  // per-cell phase timing (per-cell mode): shader cycles, and the summed
  // seconds reported as RHS_time__s / Jac_time__s
  bool profiling = false;
  double* d_prof = nullptr;
  // per-wavefront global scratch of the kPlaceGlobal variants (grow-only)
  double* d_wave_scratch = nullptr;
  size_t wave_scratch_doubles = 0;
  double* wave_scratch(size_t waves, int at_place) {
    size_t n = waves * cell_bdf_wave_scratch(P, at_place);
    if (n == 0) return nullptr;
    if (n > wave_scratch_doubles) {
      if (d_wave_scratch) HIP_OK(hipFree(d_wave_scratch));
      HIP_OK(hipMalloc(&d_wave_scratch, n * sizeof(double)));
      wave_scratch_doubles = n;
    }
    return d_wave_scratch;
  }
  double rhs_time_s = 0.0, jac_time_s = 0.0;

  ~Solver() {
    if (graph_a) (void)hipGraphExecDestroy(graph_a);
    if (graph_b) (void)hipGraphExecDestroy(graph_b);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_red_pin) (void)hipHostFree(h_red_pin);
    if (h_fail_pin) (void)hipHostFree(h_fail_pin);
    for (void* p : owned) (void)hipFree(p);
    if (d_wave_scratch) (void)hipFree(d_wave_scratch);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void setup_device();
  // Device buffers are created on first use, so that solver_initialize (and
  // the pattern getters) work on a host without a GPU.
  void ensure_device() {
    if (!stream) setup_device();
  }
  void upload_renv();
  // Per-cell kernel arguments with the solver's program and controls filled
  // in (buffers left to the caller), and the launch shape for nq queued cells.
  CellBdfArgs cell_args() const;
  // (at_place < 0: the solver's own placement)
  void cell_launch_shape(long nq, int* grid, int* block, size_t* lds, int at_place = -1) const;
  int run(double* d_state_in, double* d_env_in, double t0, double t1);
  // noalloc: the no-allocation entry point's scratch (device pointer and
  // capacity in doubles): queue counter and wavefront scratch come from it,
  // the wavefront count is capped to fit it, and nothing is allocated
  int run_cells(double* st, double* env, double t0, double t1, double* noalloc = nullptr,
                size_t noalloc_doubles = 0);
  int run_block(double* st, double* env, double t0, double t1);
};


}  // namespace camp
