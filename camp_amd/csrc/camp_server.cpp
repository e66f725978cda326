// Request server: continuous batching of CAMP chemistry solves on one GPU.
//
// Per mechanism an HBM arena holds cell slots (state, env, rate data,
// Jacobian history, statistics, end time, step count, failure counter).  A
// request is admitted when a contiguous range of slots is free.  Admitted
// requests form one first-come queue of cells: whenever a batch stream is
// free, the scheduler takes the next cells of the oldest request's
// mechanism -- from that request and the ones after it, a request being
// split across launches when it is larger than what is left -- up to a
// chunk of twice the cells the GPU holds at once, and launches the
// persistent per-cell kernel over them
// (each cell runs all of its request's timesteps).  One mechanism never
// holds every stream while another has cells waiting.  Launches allocate
// nothing: cell lists and multi-phase scratch are per stream and sized when
// the mechanism is registered.  A request's inputs are copied in on the stream of
// its first chunk (later chunks on other streams wait for that copy's
// event).  When the last chunk of a request completes its results are
// copied out on that stream and the request completes with the copy's
// event.  Chunks keep every launch short, so the streams' launches overlap
// and a request never waits for an unrelated large one to finish.  One
// thread makes every HIP call.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "../../include/camp_server.h"
#include "camp_solver_impl.hpp"

namespace camp {
namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { kQueued, kReady, kRunning, kDone };

struct Request {
  long id = 0;
  int mech = 0, n = 0, n_steps = 0;
  double dt = 0.0;
  const double *in_state = nullptr, *in_env = nullptr, *in_rates = nullptr;
  double *out_state = nullptr, *out_stats = nullptr;
  Clock::time_point t_submit, t_done;
  long slot0 = -1;
  long launched = 0, finished = 0;  // cells handed to launches / completed
  hipEvent_t up_ev = nullptr, out_ev = nullptr;
  int state = kQueued;
  bool failed = false;
};

// A run of consecutive cells of one request in a launch.
struct Piece {
  Request* r;
  long first, count;
};

struct Arena {
  Solver* S = nullptr;
  long cap = 0;
  int ns = 0, n_env = 0, nnz = 0, nv = 0, n_upd = 0;
  double *state = nullptr, *env = nullptr, *renv = nullptr, *Js = nullptr, *Jy = nullptr,
         *Jf = nullptr, *stats = nullptr, *t1 = nullptr, *rates = nullptr;
  int *fail = nullptr, *upd_off = nullptr, *nsteps = nullptr;
  int* h_fail = nullptr;  // pinned [cap]
  std::vector<void*> owned;
  std::map<long, long> free_ranges;  // start -> length
  long resident = 0;
  // per stream: the launch's cell list (device + pinned host, chunk
  // entries) and, for multi-phase programs, the wavefronts' global scratch
  // for the largest launch -- allocated when the mechanism is registered,
  // so that launching allocates and frees nothing
  std::vector<int*> d_list, h_list;
  std::vector<double*> scratch;
  long chunk = 0;
  int place = 0;  // the solver's cell placement when registered
  // replay pool (pinned host memory)
  long pool_n = 0;
  double *pool_state = nullptr, *pool_env = nullptr, *pool_rates = nullptr, *pool_out = nullptr;

  long alloc(long n) {
    for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
      if (it->second < n) continue;
      long s = it->first, len = it->second;
      free_ranges.erase(it);
      if (len > n) free_ranges[s + n] = len - n;
      resident += n;
      return s;
    }
    return -1;
  }
  void release(long s, long n) {
    resident -= n;
    auto next = free_ranges.lower_bound(s);
    if (next != free_ranges.end() && next->first == s + n) {
      n += next->second;
      free_ranges.erase(next);
    }
    if (!free_ranges.empty()) {
      auto prev = free_ranges.lower_bound(s);
      if (prev != free_ranges.begin()) {
        --prev;
        if (prev->first + prev->second == s) {
          prev->second += n;
          return;
        }
      }
    }
    free_ranges[s] = n;
  }
};

struct Lane {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  int* d_counter = nullptr;
  bool busy = false;
  int mech = -1;  // mechanism of the launch in flight
  long cells = 0;
  std::vector<Piece> batch;
};

// Mechanisms a server can hold: the arena table never reallocates, so the
// scheduler thread reads it without the lock while a caller registers
// another mechanism.
constexpr int kMaxMechs = 16;

struct Server {
  int dev = 0;
  long max_batch = 0;
  std::unique_ptr<Arena> mechs_[kMaxMechs];
  int n_mechs = 0;  // guarded by mu
  Arena* mech(int m) const { return mechs_[m].get(); }
  std::vector<Lane> lanes;
  std::thread thr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  long next_id = 0;
  std::deque<Request*> inbox;                        // guarded by mu
  std::unordered_map<long, std::unique_ptr<Request>> reqs;  // guarded by mu
  // scheduler-thread state: admitted requests with cells not yet launched
  // (arrival order), and requests whose results are being copied out
  std::deque<Request*> pending, ready;
  std::vector<Request*> copying;
  // counters (guarded by mu)
  double n_done = 0, n_failed = 0, cell_steps = 0, batches = 0, batch_cells = 0, peak = 0;

  void loop();
  bool poll_lanes();
  bool poll_copies();
  bool admit();
  bool launch();
  void upload(Arena& A, Request* r, hipStream_t st);
};

void Server::upload(Arena& A, Request* r, hipStream_t st) {
  long s = r->slot0, n = r->n;
  size_t d = sizeof(double);
  HIP_OK(hipMemcpyAsync(A.state + s * A.ns, r->in_state, n * A.ns * d, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(A.env + s * 2, r->in_env, n * 2 * d, hipMemcpyHostToDevice, st));
  if (A.n_upd > 0) {
    HIP_OK(hipMemcpyAsync(A.rates + s * A.n_upd, r->in_rates, n * A.n_upd * d,
                          hipMemcpyHostToDevice, st));
    long m = n * A.n_upd;
    hipLaunchKernelGGL(k_scatter_rates, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                       (int)n, A.n_upd, A.upd_off, A.rates + s * A.n_upd, A.n_env,
                       A.renv + s * A.n_env);
  }
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, r->dt,
                     A.t1 + s);
  HIP_OK(hipMemsetD32Async((hipDeviceptr_t)(A.nsteps + s), r->n_steps, n, st));
  // a fresh request starts with an empty Jacobian history, as a new solver
  HIP_OK(hipMemsetAsync(A.Js + s * A.nnz, 0, n * A.nnz * d, st));
  HIP_OK(hipMemsetAsync(A.Jy + s * A.nv, 0, n * A.nv * d, st));
  HIP_OK(hipMemsetAsync(A.Jf + s * A.nv, 0, n * A.nv * d, st));
  HIP_OK(hipMemsetAsync(A.fail + s, 0, n * sizeof(int), st));
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventCreateWithFlags(&r->up_ev, hipEventDisableTiming));
  HIP_OK(hipEventRecord(r->up_ev, st));
}

bool Server::admit() {
  bool any = false;
  // strict arrival order per mechanism: a request that does not fit yet
  // holds back later requests of its mechanism (no starvation of large ones)
  std::vector<char> blocked(kMaxMechs, 0);
  for (auto it = pending.begin(); it != pending.end();) {
    Request* r = *it;
    Arena& A = *mech(r->mech);
    if (blocked[r->mech]) {
      ++it;
      continue;
    }
    long s = A.alloc(r->n);
    if (s < 0) {
      blocked[r->mech] = 1;
      ++it;
      continue;
    }
    r->slot0 = s;
    r->state = kReady;
    ready.push_back(r);
    it = pending.erase(it);
    any = true;
    std::lock_guard<std::mutex> g(mu);
    peak = std::max(peak, (double)A.resident);
  }
  return any;
}

bool Server::launch() {
  bool any = false;
  for (size_t li = 0; li < lanes.size(); ++li) {
    Lane& L = lanes[li];
    if (L.busy || ready.empty()) continue;
    // The oldest ready request's mechanism goes next -- unless every other
    // stream is busy with that mechanism and another mechanism has ready
    // cells: then the oldest request of another mechanism goes, so that a
    // long run of one mechanism's launches never holds every stream.
    int m = ready.front()->mech;
    int same = 0, busy = 0;
    for (const Lane& o : lanes)
      if (o.busy) {
        ++busy;
        same += o.mech == m;
      }
    if (lanes.size() > 1 && same == busy && busy == (int)lanes.size() - 1) {
      for (Request* r : ready)
        if (r->mech != m) {
          m = r->mech;
          break;
        }
    }
    Arena& A = *mech(m);
    // A chunk: up to twice the cells the GPU holds at once, so that the
    // launch fills the GPU and its tail overlaps the next launch on another
    // stream.  (Capping chunks in cell-timesteps instead, so that multi-step
    // requests make launches no longer than others, was measured: launches
    // of multi-step requests then held fewer cells than the GPU holds,
    // capacity fell by a third and the latency spread grew.)
    long cap = std::min(max_batch, A.chunk);
    L.batch.clear();
    long total = 0, work = 0;
    for (auto it = ready.begin(); it != ready.end() && total < cap;) {
      Request* r = *it;
      if (r->mech != m) {
        ++it;
        continue;
      }
      long take = std::min(r->n - r->launched, cap - total);
      L.batch.push_back({r, r->launched, take});
      r->launched += take;
      total += take;
      work += take * r->n_steps;
      if (r->launched == r->n) it = ready.erase(it);
      else ++it;
    }
    int* h_list = A.h_list[li];
    int* d_list = A.d_list[li];
    long q = 0;
    for (const Piece& pc : L.batch) {
      Request* r = pc.r;
      for (long c = 0; c < pc.count; ++c) h_list[q++] = (int)(r->slot0 + pc.first + c);
      if (!r->up_ev) upload(A, r, L.st);
      else HIP_OK(hipStreamWaitEvent(L.st, r->up_ev, 0));
      r->state = kRunning;
    }
    HIP_OK(hipMemcpyAsync(d_list, h_list, total * sizeof(int), hipMemcpyHostToDevice, L.st));
    HIP_OK(hipMemsetAsync(L.d_counter, 0, sizeof(int), L.st));
    CellBdfArgs a = A.S->cell_args();
    a.nc = (int)total;
    a.state = A.state;
    a.env = A.env;
    a.rxn_env = A.renv;
    a.Js = A.Js;
    a.Jy = A.Jy;
    a.Jf = A.Jf;
    a.t0 = 0.0;
    a.t1 = 0.0;
    a.counter = L.d_counter;
    a.stats = A.stats;
    a.cell_list = d_list;
    a.cell_t1 = A.t1;
    a.cell_nsteps = A.nsteps;
    a.fail_count = A.fail;
    a.wave_scratch = A.scratch[li];
    a.place = A.place;
    int grid = 0, block = 0;
    size_t lds = 0;
    A.S->cell_launch_shape(total, &grid, &block, &lds, A.place);
    launch_cell_bdf(a, grid, block, lds, L.st);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(L.ev, L.st));
    L.busy = true;
    L.mech = m;
    L.cells = work;
    any = true;
    std::lock_guard<std::mutex> g(mu);
    batches += 1;
    batch_cells += (double)total;
  }
  return any;
}

// A launch has completed: a request whose every cell is now done gets its
// results copied out on the same stream, completing with that copy's event.
bool Server::poll_lanes() {
  bool any = false;
  for (Lane& L : lanes) {
    if (!L.busy) continue;
    hipError_t e = hipEventQuery(L.ev);
    if (e == hipErrorNotReady) continue;
    HIP_OK(e);
    L.busy = false;
    any = true;
    for (const Piece& pc : L.batch) {
      Request* r = pc.r;
      r->finished += pc.count;
      if (r->finished < r->n) continue;
      Arena& A = *mech(r->mech);
      size_t d = sizeof(double);
      HIP_OK(hipMemcpyAsync(r->out_state, A.state + r->slot0 * A.ns, r->n * A.ns * d,
                            hipMemcpyDeviceToHost, L.st));
      if (r->out_stats) {
        HIP_OK(hipMemcpyAsync(r->out_stats, A.stats + r->slot0 * kCellStats,
                              r->n * kCellStats * d, hipMemcpyDeviceToHost, L.st));
      }
      HIP_OK(hipMemcpyAsync(A.h_fail + r->slot0, A.fail + r->slot0, r->n * sizeof(int),
                            hipMemcpyDeviceToHost, L.st));
      HIP_OK(hipEventCreateWithFlags(&r->out_ev, hipEventDisableTiming));
      HIP_OK(hipEventRecord(r->out_ev, L.st));
      copying.push_back(r);
    }
    std::lock_guard<std::mutex> g(mu);
    cell_steps += (double)L.cells;
    L.batch.clear();
  }
  return any;
}

bool Server::poll_copies() {
  bool any = false;
  for (size_t i = 0; i < copying.size();) {
    Request* r = copying[i];
    hipError_t e = hipEventQuery(r->out_ev);
    if (e == hipErrorNotReady) {
      ++i;
      continue;
    }
    HIP_OK(e);
    auto now = Clock::now();
    (void)hipEventDestroy(r->out_ev);
    (void)hipEventDestroy(r->up_ev);
    r->out_ev = r->up_ev = nullptr;
    copying[i] = copying.back();
    copying.pop_back();
    any = true;
    std::lock_guard<std::mutex> g(mu);
    Arena& A = *mech(r->mech);
    for (long c = 0; c < r->n; ++c) r->failed |= A.h_fail[r->slot0 + c] != 0;
    A.release(r->slot0, r->n);
    r->t_done = now;
    r->state = kDone;
    n_done += 1;
    n_failed += r->failed ? 1 : 0;
    cv_done.notify_all();
  }
  return any;
}

void Server::loop() {
  HIP_OK(hipSetDevice(dev));
  for (;;) {
    {
      std::unique_lock<std::mutex> g(mu);
      while (!inbox.empty()) {
        pending.push_back(inbox.front());
        inbox.pop_front();
      }
      bool idle = pending.empty() && ready.empty() && copying.empty();
      for (Lane& L : lanes) idle = idle && !L.busy;
      if (idle) {
        if (stop) break;
        cv_work.wait_for(g, std::chrono::milliseconds(2));
        continue;
      }
    }
    bool progress = poll_lanes();
    progress |= poll_copies();
    progress |= admit();
    progress |= launch();
    if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

Server* SV(void* p) { return static_cast<Server*>(p); }

}  // namespace
}  // namespace camp

using namespace camp;

extern "C" void* camp_server_new(int n_streams, long max_batch_cells) {
  if (n_streams < 1 || n_streams > 8) die("camp_server_new: n_streams must be 1..8");
  if (max_batch_cells < 1) die("camp_server_new: max_batch_cells must be positive");
  Server* s = new Server();
  HIP_OK(hipGetDevice(&s->dev));
  s->max_batch = max_batch_cells;
  s->lanes.resize(n_streams);
  for (Lane& L : s->lanes) {
    HIP_OK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
    HIP_OK(hipMalloc(&L.d_counter, sizeof(int)));
  }
  s->thr = std::thread([s] { s->loop(); });
  return s;
}

extern "C" int camp_solver_rate_layout(void* solver_data, int* types, int* uids) {
  Model& m = static_cast<Solver*>(solver_data)->m;
  int n = 0;
  for (int r = 0; r < m.n_rxn; ++r) {
    int* id = m.rint(r);
    int t = id[0], uid;
    if (t == kRxnPhotolysis) uid = id[3];
    else if (t == kRxnEmission || t == kRxnFirstOrderLoss) uid = id[1];
    else continue;
    if (types) types[n] = t;
    if (uids) uids[n] = uid;
    ++n;
  }
  return n;
}

extern "C" int camp_server_add_mechanism(void* server, void* solver_data, long max_cells) {
  Server* s = SV(server);
  Solver* S = static_cast<Solver*>(solver_data);
  if (!S->m.initialized) die("camp_server_add_mechanism: solver is not initialised");
  S->ensure_device();
  std::unique_ptr<Arena> A(new Arena());
  A->S = S;
  A->ns = S->m.prog.ns;
  A->n_env = S->m.prog.n_env;
  A->nnz = S->m.prog.nnz;
  A->nv = S->m.prog.nv;
  std::vector<int> off;
  for (int r = 0; r < S->m.n_rxn; ++r) {
    int t = S->m.rint(r)[0];
    if (t == kRxnPhotolysis || t == kRxnEmission || t == kRxnFirstOrderLoss)
      off.push_back(S->m.rxn_env_idx[r] + 1);  // BASE_RATE slot
  }
  A->n_upd = (int)off.size();
  size_t per_cell = sizeof(double) * ((size_t)A->ns + 2 + A->n_env + A->nnz + 2 * A->nv +
                                      kCellStats + 1 + A->n_upd) +
                    2 * sizeof(int) + s->lanes.size() * sizeof(int);
  if (max_cells <= 0) {
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    max_cells = (long)std::min<size_t>(fr / 2 / per_cell, (size_t)1 << 24);
  }
  if (max_cells < 1) die("camp_server_add_mechanism: no device memory for a cell arena");
  long n = max_cells;
  A->cap = n;
  A->state = dev_alloc<double>((size_t)n * A->ns, A->owned);
  A->env = dev_alloc<double>((size_t)n * 2, A->owned);
  A->renv = dev_alloc<double>((size_t)n * A->n_env, A->owned);
  A->Js = dev_alloc<double>((size_t)n * A->nnz, A->owned);
  A->Jy = dev_alloc<double>((size_t)n * A->nv, A->owned);
  A->Jf = dev_alloc<double>((size_t)n * A->nv, A->owned);
  A->stats = dev_alloc<double>((size_t)n * kCellStats, A->owned);
  A->t1 = dev_alloc<double>((size_t)n, A->owned);
  A->nsteps = dev_alloc<int>((size_t)n, A->owned);
  A->rates = dev_alloc<double>((size_t)n * std::max(A->n_upd, 1), A->owned);
  A->fail = dev_alloc<int>((size_t)n, A->owned);
  A->upd_off = dev_upload(off, A->owned);
  HIP_OK(hipHostMalloc((void**)&A->h_fail, (size_t)n * sizeof(int), hipHostMallocDefault));
  A->free_ranges[0] = n;
  // a launch takes at most twice the cells the GPU holds at once; its cell
  // list and wavefront scratch are allocated here, per stream, at that size
  int grid = 0, block = 0;
  size_t lds = 0;
  A->place = cell_placement(S->P, S->place);
  S->cell_launch_shape(1L << 30, &grid, &block, &lds, A->place);
  A->chunk = 2L * grid * (block / kWave);
  size_t waves = (size_t)grid * (block / kWave);
  for (size_t li = 0; li < s->lanes.size(); ++li) {
    A->d_list.push_back(dev_alloc<int>((size_t)A->chunk, A->owned));
    int* h = nullptr;
    HIP_OK(hipHostMalloc((void**)&h, (size_t)A->chunk * sizeof(int), hipHostMallocDefault));
    A->h_list.push_back(h);
    size_t per_wave = cell_bdf_wave_scratch(S->P, A->place);
    A->scratch.push_back(per_wave ? dev_alloc<double>(waves * per_wave, A->owned) : nullptr);
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (s->n_mechs == kMaxMechs) die("camp_server_add_mechanism: too many mechanisms (16)");
  s->mechs_[s->n_mechs] = std::move(A);
  return s->n_mechs++;
}

extern "C" long camp_server_capacity(void* server, int mech) {
  Server* s = SV(server);
  std::lock_guard<std::mutex> g(s->mu);
  if (mech < 0 || mech >= s->n_mechs) return -1;
  return s->mech(mech)->cap;
}

extern "C" long camp_server_submit(void* server, int mech, int n_cells, const double* state,
                                   const double* env, const double* base_rates, double dt,
                                   int n_steps, double* out_state, double* out_stats) {
  Server* s = SV(server);
  std::lock_guard<std::mutex> g(s->mu);
  if (mech < 0 || mech >= s->n_mechs) return -1;
  Arena& A = *s->mech(mech);
  if (n_cells < 1 || n_cells > A.cap || n_steps < 1 || !(dt > 0.0)) return -1;
  if (A.n_upd > 0 && !base_rates) return -1;
  std::unique_ptr<Request> r(new Request());
  r->id = s->next_id++;
  r->mech = mech;
  r->n = n_cells;
  r->dt = dt;
  // every timestep of a cell runs in one launch (the kernel restages the
  // cell between them, as separate launches would)
  r->n_steps = n_steps;
  r->in_state = state;
  r->in_env = env;
  r->in_rates = base_rates;
  r->out_state = out_state;
  r->out_stats = out_stats;
  r->t_submit = Clock::now();
  s->inbox.push_back(r.get());
  long id = r->id;
  s->reqs[id] = std::move(r);
  s->cv_work.notify_one();
  return id;
}

extern "C" int camp_server_wait(void* server, long request, double* latency_s) {
  Server* s = SV(server);
  std::unique_lock<std::mutex> g(s->mu);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) return 1;
  Request* r = it->second.get();
  s->cv_done.wait(g, [r] { return r->state == kDone; });
  if (latency_s) latency_s[0] = std::chrono::duration<double>(r->t_done - r->t_submit).count();
  int rc = r->failed ? 1 : 0;
  s->reqs.erase(it);
  return rc;
}

extern "C" int camp_server_set_pool(void* server, int mech, long n_cells, const double* state,
                                    const double* env, const double* base_rates) {
  Server* s = SV(server);
  Arena* A;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (mech < 0 || mech >= s->n_mechs || n_cells < 1) return 1;
    A = s->mech(mech);
  }
  if (A->pool_n) return 1;
  size_t d = sizeof(double);
  HIP_OK(hipHostMalloc((void**)&A->pool_state, n_cells * A->ns * d, hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&A->pool_out, n_cells * A->ns * d, hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&A->pool_env, n_cells * 2 * d, hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&A->pool_rates, n_cells * std::max(A->n_upd, 1) * d,
                       hipHostMallocDefault));
  std::memcpy(A->pool_state, state, n_cells * A->ns * d);
  std::memcpy(A->pool_env, env, n_cells * 2 * d);
  if (A->n_upd > 0) std::memcpy(A->pool_rates, base_rates, n_cells * A->n_upd * d);
  A->pool_n = n_cells;
  return 0;
}

extern "C" int camp_server_replay(void* server, int n_requests, const double* t_arrive,
                                  const int* mech, const int* n_cells, const long* offset,
                                  const double* dt, const int* n_steps, double* latency_s,
                                  double* makespan_s) {
  Server* s = SV(server);
  std::vector<long> ids(n_requests, -1);
  auto t0 = Clock::now();
  // a progress line on stderr every 20 s of a long replay
  Clock::time_point said = t0;
  double done0 = 0.0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    done0 = s->n_done;
  }
  auto progress = [&](int submitted) {
    auto now = Clock::now();
    if (now - said < std::chrono::seconds(20)) return;
    said = now;
    double done = 0.0;
    {
      std::lock_guard<std::mutex> g(s->mu);
      done = s->n_done - done0;
    }
    std::fprintf(stderr, "[server] replay: %d of %d requests submitted, %.0f done, %.0f s\n",
                 submitted, n_requests, done,
                 std::chrono::duration<double>(now - t0).count());
    std::fflush(stderr);
  };
  for (int i = 0; i < n_requests; ++i) {
    Arena* A;
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (mech[i] < 0 || mech[i] >= s->n_mechs) return -1;
      A = s->mech(mech[i]);
    }
    if (!A->pool_n || offset[i] < 0 || offset[i] + n_cells[i] > A->pool_n) return -1;
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(
                              std::chrono::duration<double>(t_arrive[i]));
    while (Clock::now() < due) {
      std::this_thread::sleep_until(std::min(due, Clock::now() + std::chrono::seconds(5)));
      progress(i);
    }
    long o = offset[i];
    ids[i] = camp_server_submit(server, mech[i], n_cells[i], A->pool_state + o * A->ns,
                                A->pool_env + o * 2, A->pool_rates + o * std::max(A->n_upd, 1),
                                dt[i], n_steps[i], A->pool_out + o * A->ns, nullptr);
    if (ids[i] < 0) return -1;
  }
  int failed = 0;
  Clock::time_point last = t0;
  for (int i = 0; i < n_requests; ++i) {
    std::unique_lock<std::mutex> g(s->mu);
    Request* r = s->reqs[ids[i]].get();
    while (!s->cv_done.wait_for(g, std::chrono::seconds(5), [r] { return r->state == kDone; })) {
      g.unlock();
      progress(n_requests);
      g.lock();
    }
    latency_s[i] = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    last = std::max(last, r->t_done);
    failed += r->failed ? 1 : 0;
    s->reqs.erase(ids[i]);
  }
  if (makespan_s) makespan_s[0] = std::chrono::duration<double>(last - t0).count();
  return failed;
}

extern "C" void camp_server_counters(void* server, double* out6) {
  Server* s = SV(server);
  std::lock_guard<std::mutex> g(s->mu);
  out6[0] = s->n_done;
  out6[1] = s->n_failed;
  out6[2] = s->cell_steps;
  out6[3] = s->batches;
  out6[4] = s->batches > 0 ? s->batch_cells / s->batches : 0.0;
  out6[5] = s->peak;
}

extern "C" void camp_server_free(void* server) {
  Server* s = SV(server);
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->thr.join();
  for (Lane& L : s->lanes) {
    (void)hipStreamSynchronize(L.st);
    (void)hipStreamDestroy(L.st);
    (void)hipEventDestroy(L.ev);
    (void)hipFree(L.d_counter);
  }
  for (int m = 0; m < s->n_mechs; ++m) {
    Arena* A = s->mech(m);
    for (void* p : A->owned) (void)hipFree(p);
    for (int* h : A->h_list) (void)hipHostFree(h);
    (void)hipHostFree(A->h_fail);
    if (A->pool_n) {
      (void)hipHostFree(A->pool_state);
      (void)hipHostFree(A->pool_out);
      (void)hipHostFree(A->pool_env);
      (void)hipHostFree(A->pool_rates);
    }
  }
  delete s;
}
