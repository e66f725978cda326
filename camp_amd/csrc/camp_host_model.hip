// Host-model interface: the gather / solve / scatter that the reference's
// MONARCH interface wraps around camp_core%solve
// (test/monarch/camp_monarch_interface.F90:343-509), on the GPU.
//
// The mapped tracers' blocks of the host model's single-precision tracer
// array (each tracer is one contiguous ni*nj*nk block in Fortran order),
// the temperature, pressure, air density and water-vapour blocks are copied
// to HBM as they are; one kernel gathers every cell of the requested range
// into the solver's cell-major state and env, the per-cell solver runs, and
// one kernel scatters the mapped species back into the tracer blocks, which
// are copied back whole (cells outside the range are unchanged on the
// device).  Per call the host link moves 4 bytes per mapped tracer and cell
// each way plus the four meteorological fields.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/camp_host_model.h"
#include "../../include/camp_solver.h"
#include "camp_solver_impl.hpp"

namespace camp {
namespace {

struct HostModel {
  Solver* S = nullptr;
  int ni = 0, nj = 0, nk = 0, n_tracers = 0, n_water = 0, n_map = 0, water = 0;  // water 0-based
  long n3 = 0;
  std::vector<int> tracer, camp;  // 0-based
  int* d_camp = nullptr;
  float *d_conc = nullptr, *d_T = nullptr, *d_P = nullptr, *d_rho = nullptr, *d_wv = nullptr;
  double *d_state = nullptr, *d_env = nullptr;  // [n_cells][ns], [n_cells][2]
  std::vector<void*> owned;
  int ns() const { return S->m.prog.ns; }
  long cells() const { return S->m.nc; }
};

// The cells of one call: W->E i0 .. i0+nI-1, S->N j0 .. j0+nJ-1 and the
// host model's levels k0 .. k0+nK-1 of its k loop (0-based; k_flip is
// nk-1-k).  Cell z of the range is (k-k0)*nI*nJ + (j-j0)*nI + (i-i0).
struct Range {
  int i0, nI, j0, nJ, k0, nK;
  __host__ __device__ long cells() const { return (long)nI * nJ * nK; }
};

__device__ __forceinline__ void hm_cell(int ni, int nj, int nk, Range r, long z, long* idx3,
                                        long* idxw) {
  const long plane = (long)r.nI * r.nJ;
  const int k = r.k0 + (int)(z / plane);
  const long rem = z % plane;
  const int j = r.j0 + (int)(rem / r.nI);
  const int i = r.i0 + (int)(rem % r.nI);
  const int kf = nk - 1 - k;  // k_flip
  *idx3 = i + (long)ni * (j + (long)nj * kf);  // (i, j, k_flip) of ni x nj x nk
  *idxw = i + (long)ni * (k + (long)nk * j);   // (i, k, j) of ni x nk x nj
}

// Cell z of the range into state row z and env row z (clear: the whole row
// first, the reference's single-cell branch).
__global__ void k_hm_gather(int ni, int nj, int nk, Range r, long n3, int n_map, const int* camp,
                            const float* conc, const float* T, const float* P, const float* rho,
                            const float* wv, int water, int ns, int clear, double* state,
                            double* env) {
  const long n = r.cells();
  for (long z = blockIdx.x * (long)blockDim.x + threadIdx.x; z < n;
       z += (long)gridDim.x * blockDim.x) {
    long c3, cw;
    hm_cell(ni, nj, nk, r, z, &c3, &cw);
    double* s = state + z * ns;
    if (clear)
      for (int v = 0; v < ns; ++v) s[v] = 0.0;
    // state(map_camp_id(:)) = 0; state(map_camp_id(:)) += conc(...): the
    // right-hand side is evaluated before the assignment, so each mapped
    // species gets its tracer's value (the last one where a species is
    // mapped twice)
    for (int m = 0; m < n_map; ++m) s[camp[m]] = (double)conc[(long)m * n3 + c3];
    // water_conc * air_density is a single-precision product, then scaled
    // in double precision
    const float wr = __fmul_rn(wv[c3], rho[cw]);
    s[water] = (double)wr * 1.0e9;
    env[2 * z] = (double)T[c3];
    env[2 * z + 1] = (double)P[cw];
  }
}

__global__ void k_hm_scatter(int ni, int nj, int nk, Range r, long n3, int n_map, const int* camp,
                             int ns, const double* state, float* conc) {
  const long n = r.cells();
  for (long z = blockIdx.x * (long)blockDim.x + threadIdx.x; z < n;
       z += (long)gridDim.x * blockDim.x) {
    long c3, cw;
    hm_cell(ni, nj, nk, r, z, &c3, &cw);
    const double* s = state + z * ns;
    for (int m = 0; m < n_map; ++m) conc[(long)m * n3 + c3] = (float)s[camp[m]];
  }
}

[[noreturn]] void hm_die(const char* msg, long a, long b) {
  std::fprintf(stderr, "camp host-model interface: %s (%ld, %ld)\n", msg, a, b);
  std::exit(EXIT_FAILURE);
}

HostModel* HM(void* p) { return static_cast<HostModel*>(p); }

Range range_of(const HostModel& h, int i_start, int i_end, int j_start, int j_end) {
  if (i_start < 1 || i_end > h.ni || i_start > i_end)
    hm_die("W->E range outside the grid", i_start, i_end);
  if (j_start < 1 || j_end > h.nj || j_start > j_end)
    hm_die("S->N range outside the grid", j_start, j_end);
  return Range{i_start - 1, i_end - i_start + 1, j_start - 1, j_end - j_start + 1, 0, h.nk};
}

// Copy the host model's fields for one call to HBM.
void upload(HostModel& h, const float* temperature, const float* conc, const float* water_conc,
            int water_vapor_index, const float* air_density, const float* pressure) {
  if (water_vapor_index < 1 || water_vapor_index > h.n_water)
    hm_die("water vapour index outside the water array", water_vapor_index, h.n_water);
  hipStream_t st = h.S->stream;
  const size_t b = (size_t)h.n3 * sizeof(float);
  for (int m = 0; m < h.n_map; ++m)
    HIP_OK(hipMemcpyAsync(h.d_conc + (size_t)m * h.n3, conc + (size_t)h.tracer[m] * h.n3, b,
                          hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(h.d_T, temperature, b, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(h.d_wv, water_conc + (size_t)(water_vapor_index - 1) * h.n3, b,
                        hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(h.d_rho, air_density, b, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(h.d_P, pressure, b, hipMemcpyHostToDevice, st));
}

void gather(HostModel& h, Range r, int clear, double* state, double* env) {
  const long n = r.cells();
  const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_hm_gather, dim3(blocks), dim3(256), 0, h.S->stream, h.ni, h.nj, h.nk, r,
                     h.n3, h.n_map, h.d_camp, h.d_conc, h.d_T, h.d_P, h.d_rho, h.d_wv, h.water,
                     h.ns(), clear, state, env);
  HIP_OK(hipGetLastError());
}

void scatter(HostModel& h, Range r, const double* state) {
  const long n = r.cells();
  const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_hm_scatter, dim3(blocks), dim3(256), 0, h.S->stream, h.ni, h.nj, h.nk, r,
                     h.n3, h.n_map, h.d_camp, h.ns(), state, h.d_conc);
  HIP_OK(hipGetLastError());
}

}  // namespace
}  // namespace camp

using namespace camp;

extern "C" void* camp_hostmodel_new(void* solver_data, int ni, int nj, int nk, int n_tracers,
                                    int n_water, int n_map, const int* map_tracer_id,
                                    const int* map_camp_id, int water_camp_id) {
  Solver* S = static_cast<Solver*>(solver_data);
  if (!S || !S->m.initialized) hm_die("the solver is not initialised", 0, 0);
  if (ni < 1 || nj < 1 || nk < 1 || n_tracers < 1 || n_water < 1 || n_map < 0)
    hm_die("grid or array sizes must be positive", ni, nj);
  HostModel* h = new HostModel();
  h->S = S;
  h->ni = ni;
  h->nj = nj;
  h->nk = nk;
  h->n_tracers = n_tracers;
  h->n_water = n_water;
  h->n_map = n_map;
  h->n3 = (long)ni * nj * nk;
  const int ns = S->m.prog.ns;
  for (int m = 0; m < n_map; ++m) {
    if (map_tracer_id[m] < 1 || map_tracer_id[m] > n_tracers)
      hm_die("tracer id outside the tracer array", map_tracer_id[m], n_tracers);
    if (map_camp_id[m] < 1 || map_camp_id[m] > ns)
      hm_die("CAMP species id outside the state", map_camp_id[m], ns);
    h->tracer.push_back(map_tracer_id[m] - 1);
    h->camp.push_back(map_camp_id[m] - 1);
  }
  if (water_camp_id < 1 || water_camp_id > ns)
    hm_die("gas-phase water id outside the state", water_camp_id, ns);
  h->water = water_camp_id - 1;
  S->ensure_device();
  h->d_camp = dev_upload(h->camp, h->owned);
  h->d_conc = dev_alloc<float>((size_t)std::max(n_map, 1) * h->n3, h->owned);
  h->d_T = dev_alloc<float>((size_t)h->n3, h->owned);
  h->d_P = dev_alloc<float>((size_t)h->n3, h->owned);
  h->d_rho = dev_alloc<float>((size_t)h->n3, h->owned);
  h->d_wv = dev_alloc<float>((size_t)h->n3, h->owned);
  h->d_state = dev_alloc<double>((size_t)h->cells() * ns, h->owned);  // zeroed
  h->d_env = dev_alloc<double>((size_t)h->cells() * 2, h->owned);
  return h;
}

extern "C" int camp_hostmodel_integrate(void* hm, double time_step, int i_start, int i_end,
                                        int j_start, int j_end, const float* temperature,
                                        float* conc, const float* water_conc,
                                        int water_vapor_index, const float* air_density,
                                        const float* pressure) {
  HostModel& h = *HM(hm);
  Range r = range_of(h, i_start, i_end, j_start, j_end);
  const long n = r.cells();
  upload(h, temperature, conc, water_conc, water_vapor_index, air_density, pressure);
  int rc = CAMP_SOLVER_SUCCESS;
  if (h.cells() == 1 && n > 1) {
    // the reference's single-cell branch: the cells one at a time in its
    // loop order (i, then j, then k innermost), the state cleared before
    // each gather; a failed cell makes the call fail (the reference asserts)
    for (int i = r.i0; i < r.i0 + r.nI; ++i)
      for (int j = r.j0; j < r.j0 + r.nJ; ++j)
        for (int k = 0; k < h.nk; ++k) {
          Range one{i, 1, j, 1, k, 1};
          gather(h, one, 1, h.d_state, h.d_env);
          if (camp_solver_run_device(h.S, h.d_state, h.d_env, 0.0, time_step) != 0)
            rc = CAMP_SOLVER_FAIL;
          scatter(h, one, h.d_state);
        }
  } else {
    if (n != h.cells()) hm_die("grid cell number mismatch: got, expected", n, h.cells());
    gather(h, r, 0, h.d_state, h.d_env);
    rc = camp_solver_run_device(h.S, h.d_state, h.d_env, 0.0, time_step);
    scatter(h, r, h.d_state);
  }
  hipStream_t st = h.S->stream;
  const size_t b = (size_t)h.n3 * sizeof(float);
  for (int m = 0; m < h.n_map; ++m)
    HIP_OK(hipMemcpyAsync(conc + (size_t)h.tracer[m] * h.n3, h.d_conc + (size_t)m * h.n3, b,
                          hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  return rc;
}

extern "C" void camp_hostmodel_set_state(void* hm, const double* state) {
  HostModel& h = *HM(hm);
  HIP_OK(hipMemcpy(h.d_state, state, (size_t)h.cells() * h.ns() * sizeof(double),
                   hipMemcpyHostToDevice));
}

extern "C" void camp_hostmodel_gather(void* hm, int i_start, int i_end, int j_start, int j_end,
                                      const float* temperature, const float* conc,
                                      const float* water_conc, int water_vapor_index,
                                      const float* air_density, const float* pressure,
                                      double* state_out, double* env_out) {
  HostModel& h = *HM(hm);
  Range r = range_of(h, i_start, i_end, j_start, j_end);
  const long n = r.cells();
  const bool single = h.cells() == 1 && n > 1;
  if (!single && n != h.cells()) hm_die("grid cell number mismatch: got, expected", n, h.cells());
  upload(h, temperature, conc, water_conc, water_vapor_index, air_density, pressure);
  std::vector<void*> tmp;
  double* st = dev_alloc<double>((size_t)n * h.ns(), tmp);
  double* env = dev_alloc<double>((size_t)n * 2, tmp);
  if (!single)
    HIP_OK(hipMemcpyAsync(st, h.d_state, (size_t)n * h.ns() * sizeof(double),
                          hipMemcpyDeviceToDevice, h.S->stream));
  gather(h, r, single ? 1 : 0, st, env);
  HIP_OK(hipMemcpyAsync(state_out, st, (size_t)n * h.ns() * sizeof(double),
                        hipMemcpyDeviceToHost, h.S->stream));
  HIP_OK(hipMemcpyAsync(env_out, env, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost,
                        h.S->stream));
  HIP_OK(hipStreamSynchronize(h.S->stream));
  for (void* p : tmp) HIP_OK(hipFree(p));
}

extern "C" void camp_hostmodel_free(void* hm) {
  HostModel* h = HM(hm);
  if (!h) return;
  HIP_OK(hipStreamSynchronize(h->S->stream));
  for (void* p : h->owned) (void)hipFree(p);
  delete h;
}
