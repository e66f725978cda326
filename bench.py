"""Benchmark: cell-timesteps/s of the batched per-cell BDF solver on one MI355X.

Workload (default): the CB05-scale synthetic gas mechanism (72 species, 190
reactions: 133 Arrhenius, 28 Troe, 19 photolysis, 5 emission, 5 first-order
loss; mechanism seed 1) over 100,000 cells (cell seed 2) of uneven
stiffness: 50% night cells (photolysis J = 0) and 10% with concentrations
scaled by 1e3; FP64, rtol 1e-4, atol 1e-8 ppm.  One step = one 180 s
solver_run over all cells, each cell with its own BDF order, step size and
Newton control, on state already resident in HBM.  W warmup steps from the
seeded initial state (the first carries the stiff start-up transient) run
untimed on the measured solver, then K consecutive steps are timed.

Serving leg: the request server with the gas mechanism and the
multi-phase mechanism registered.  Requests are gas (1..50,000 cells drawn
from the spun-up 100k-cell pool) or multi-phase (1..2,000 cells of a seeded
multi-phase pool) with equal odds, log-uniform sizes, dt uniform 60..600 s
and 1..4 timesteps.  Capacity is measured with half the requests arriving at
once; then Poisson arrivals at 70% of it give p50/p99 request latency
(submission to results on the host).

Further legs: the plain one-launch-per-RHS path against the persistent
kernel on the 10,000-cell gas workload, the multi-phase workload (20,000
cells), the autotuner's choice of cells per workgroup.

Prints one JSON line.  `roofline` is for the dominant kernel (k_cell_bdf,
the persistent per-cell kernel): FP64 algorithmic flops (per-evaluation
counts of f, J, LU and triangular solves, times the evaluation counts the
kernel itself reports) over the kernel's HIP-event duration.
`roofline_fj` gives the plain-path f(y) and J(y) kernels against HBM.
`cpu_baseline` is the CPU oracle (a C restatement of the reference path)
on a bounded sample of the same cells and steps, on this host's cores.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def make_inputs(n_cells):
    from camp_amd.mechanisms import cb05_scale_mechanism, synthetic_cells
    mech = cb05_scale_mechanism(1)
    cells = synthetic_cells(mech, n_cells, seed=2, night_fraction=0.5, stressed_fraction=0.1)
    return mech, cells


def host_info():
    """The host the CPU baseline ran on: usable cores, the CPU model, and how
    the oracle was compiled."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    from oracle import oracle
    return {"nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(),
            "cpu_model": model, "oracle_build": "gcc " + oracle.NATIVE_FLAGS}


def _pool_map(fn, jobs):
    """Oracle work in spawned processes (one per job), which load the
    -O3 -march=native build of the oracle."""
    os.environ["CAMP_ORACLE_NATIVE"] = "1"
    with mp.get_context("spawn").Pool(len(jobs)) as pool:
        return pool.map(fn, jobs)


def _cpu_worker(args):
    n_cells, c0, c1, warm, steps, dt = args
    from oracle.oracle import Oracle
    from camp_amd.mechanisms import cb05_scale_mechanism, synthetic_cells
    mech = cb05_scale_mechanism(1)
    # cells c0..c1-1 of the workload (the generator is batch-independent)
    cells = synthetic_cells(mech, c1 - c0, seed=2, cell_offset=c0, night_fraction=0.5,
                            stressed_fraction=0.1)
    o = Oracle(mech, c1 - c0)
    o.set_rates(cells.rates)
    st = np.ascontiguousarray(cells.state)
    env = np.ascontiguousarray(cells.env)
    for _ in range(warm):
        o.run(st, env, 0.0, dt, mode=1)
    t = time.perf_counter()
    for _ in range(steps):
        o.run(st, env, 0.0, dt, mode=1)
    return time.perf_counter() - t


def cpu_baseline(n_cells, warm, steps, dt, cores, per_core):
    """Oracle per-cell solve of cells spread over the workload: the same W
    untimed and K timed steps as the GPU; rate = cells x K / slowest worker."""
    from oracle import oracle
    oracle.build_native()
    jobs = []
    for i in range(cores):
        c0 = i * (n_cells // cores)
        jobs.append((n_cells, c0, c0 + per_core, warm, steps, dt))
    times = _pool_map(_cpu_worker, jobs)
    wall = max(times)
    return {"value": cores * per_core * steps / wall, "unit": "cell-timesteps/s", "cores": cores,
            "kind": "port", "host": host_info(),
            "sample": "%d cells (%d blocks of %d consecutive cells spread over the %d-cell "
                      "workload) x %d timed steps of %g s after %d untimed, per-cell "
                      "CVODE-algorithm BDF, one process per core" % (
                          cores * per_core, cores, per_core, n_cells, steps, dt, warm)}


def _cpu_block_worker(args):
    c0, c1, warm, steps, dt = args
    from oracle.oracle import Oracle
    from camp_amd.mechanisms import cb05_scale_mechanism, synthetic_cells
    mech = cb05_scale_mechanism(1)
    cells = synthetic_cells(mech, c1 - c0, seed=2, cell_offset=c0)
    o = Oracle(mech, c1 - c0)
    o.set_rates(cells.rates)
    st = np.ascontiguousarray(cells.state)
    env = np.ascontiguousarray(cells.env)
    for _ in range(warm):
        o.run(st, env, 0.0, dt, mode=0)
    t = time.perf_counter()
    for _ in range(steps):
        o.run(st, env, 0.0, dt, mode=0)
    return time.perf_counter() - t


def cpu_baseline_block(n_cells, warm, steps, dt, cores, per_core):
    """The reference's own semantics on the CPU -- one CVODE over all cells
    of a solver_run call -- on a bounded sample of the plain-path workload:
    one block of `per_core` consecutive cells per core, the blocks spread
    over the workload (each block its own BDF with shared step control, as
    one MPI rank per core would run it); the same W untimed and K timed
    steps as the GPU's plain path.  Larger blocks cost more per cell (the
    stiffest cell sets every cell's step), so this favours the CPU over one
    block of all cells."""
    from oracle import oracle
    oracle.build_native()
    jobs = []
    for i in range(cores):
        c0 = i * (n_cells // cores)
        jobs.append((c0, c0 + per_core, warm, steps, dt))
    wall = max(_pool_map(_cpu_block_worker, jobs))
    return {"value": cores * per_core * steps / wall, "unit": "cell-timesteps/s",
            "cores": cores, "kind": "port",
            "sample": "%d cells of the %d-cell plain-path workload: %d blocks of %d consecutive "
                      "cells spread over it, one BDF with shared step control per block, one "
                      "process per core, %d timed steps of %g s after %d untimed" % (
                          cores * per_core, n_cells, cores, per_core, steps, dt, warm)}


def _box_worker(args):
    reps, = args
    from oracle.oracle import Oracle
    from camp_amd.mechanisms import box_mechanism, synthetic_cells
    mech = box_mechanism(1)
    cells = synthetic_cells(mech, 1, seed=2, T=298.0, P=101325.0)
    best = None
    for _ in range(reps):
        o = Oracle(mech, 1)
        o.set_rates(cells.rates)
        st = np.ascontiguousarray(cells.state).copy()
        env = np.ascontiguousarray(cells.env)
        t = time.perf_counter()
        for _ in range(60):
            assert o.run(st, env, 0.0, 60.0, mode=1) == 0
        el = time.perf_counter() - t
        best = el if best is None else min(best, el)
    return best


def box_model_leg(reps=20):
    """The reference's CPU-runnable configuration: the 10-species,
    15-reaction box model, one cell at 298 K, one hour in 60 s steps at
    rtol 1e-4 -- the oracle on one core, and the per-cell GPU kernel (one
    launch per step, host arrays, as solver_run is called)."""
    from oracle import oracle
    oracle.build_native()
    cpu_s = _pool_map(_box_worker, [(reps,)])[0]
    from camp_amd.core import CampSolver
    from camp_amd.mechanisms import box_mechanism, synthetic_cells
    mech = box_mechanism(1)
    cells = synthetic_cells(mech, 1, seed=2, T=298.0, P=101325.0)
    s = CampSolver(mech, 1)
    s.set_rates(cells.rates)
    st = cells.state.copy()
    assert s.solve(st, cells.env, 0.0, 60.0) == 0
    st = cells.state.copy()
    t = time.perf_counter()
    for _ in range(60):
        assert s.solve(st, cells.env, 0.0, 60.0) == 0
    gpu_s = time.perf_counter() - t
    s.close()
    return {"what": "box model (10 species, 15 Arrhenius), 1 cell, 60 x 60 s steps, rtol 1e-4",
            "unit": "cell-timesteps/s", "cpu_one_core": 60.0 / cpu_s,
            "gpu_solver_run": 60.0 / gpu_s,
            "note": "one cell cannot use a GPU; the per-call launch and copies dominate"}


def serving(mech, solver, pool_state, env, rates, n_requests, mp_max_cells, mp_spun=None,
            dt=180.0, max_batch=1 << 17, seed=3):
    """Request server with both mechanisms: capacity at saturation, then
    Poisson arrivals at 70% of it.  Each request is a gas request (1..50,000
    cells of the spun-up 100k-cell pool) or a multi-phase request (1..
    mp_max_cells cells of a multi-phase pool, spun up by one dt step like
    the gas pool: the multi-phase leg's cells when it ran) with equal
    probability, dt U(60, 600) s and 1..4 timesteps."""
    import torch
    from camp_amd.core import CampSolver
    from camp_amd.multiphase import multiphase_cells, multiphase_mechanism
    from camp_amd.server import CampServer, base_rate_matrix
    n_pool = pool_state.shape[0]
    rng = np.random.default_rng(seed)
    srv = CampServer(n_streams=4, max_batch_cells=max_batch)
    gas = srv.add_mechanism(CampSolver(mech, 1))
    srv.set_pool(gas, pool_state, env, base_rate_matrix(solver, rates, n_pool))
    mp_pool = 0
    if mp_max_cells > 0:
        mpm = multiphase_mechanism(1)
        mp_pool = max(mp_max_cells, 2000)
        if mp_spun is not None and mp_spun[1].shape[0] >= mp_pool:
            mcells, mstate = mp_spun
        else:
            mcells = multiphase_cells(mpm, mp_pool, seed=2)
            sp = CampSolver(mpm, mp_pool)
            sp.set_rates(mcells.rates)
            dev = torch.device("cuda:0")
            st_d = torch.from_numpy(mcells.state).to(dev)
            env_d = torch.from_numpy(mcells.env).to(dev)
            progress("serving: spinning up %d multi-phase pool cells" % mp_pool)
            sp.solve_device(st_d.data_ptr(), env_d.data_ptr(), 0.0, dt)
            torch.cuda.synchronize()
            mstate = st_d.cpu().numpy()
            progress("serving: pool ready")
        msol = CampSolver(mpm, 1)
        mpid = srv.add_mechanism(msol)
        srv.set_pool(mpid, np.ascontiguousarray(mstate[:mp_pool]), mcells.env[:mp_pool],
                     base_rate_matrix(msol, mcells.rates, mp_pool))

    def draw(k):
        is_mp = (rng.random(k) < 0.5) if mp_pool else np.zeros(k, bool)
        top = np.where(is_mp, float(mp_max_cells), 50000.0)
        sizes = np.exp(rng.uniform(0.0, 1.0, size=k) * np.log(top))
        sizes = np.maximum(np.minimum(sizes, np.where(is_mp, mp_pool, n_pool)).astype(np.int32), 1)
        pool = np.where(is_mp, mp_pool, n_pool)
        off = (rng.random(k) * (pool - sizes + 1)).astype(np.int64)
        dts = rng.uniform(60.0, 600.0, size=k)
        steps = rng.integers(1, 5, size=k).astype(np.int32)
        return is_mp.astype(np.int32), sizes, off, dts, steps

    k_cap = max(16, n_requests // 2)
    m, sizes, off, dts, steps = draw(k_cap)
    progress("serving: capacity pass, %d requests at once" % k_cap)
    lat, span, failed_cap = srv.replay(np.zeros(k_cap), m, sizes, off, dts, steps)
    progress("serving: capacity %.2f requests/s; %d Poisson arrivals at 70%%"
             % (k_cap / span, n_requests))
    cap_rps = k_cap / span
    capacity = float(np.sum(sizes * steps)) / span
    m, sizes, off, dts, steps = draw(n_requests)
    t_arr = np.cumsum(rng.exponential(1.0 / (0.7 * cap_rps), size=n_requests))
    lat, span, failed = srv.replay(t_arr, m, sizes, off, dts, steps)
    c = srv.counters()
    srv.close()
    work = sizes.astype(np.float64) * steps
    by_size = {}
    for lo, hi in ((1, 100), (100, 3000), (3000, 50001)):
        sel = (sizes >= lo) & (sizes < hi)
        if np.count_nonzero(sel) >= 5:
            l50, l99 = np.percentile(lat[sel], 50), np.percentile(lat[sel], 99)
            by_size["%d..%d cells" % (lo, hi - 1)] = {
                "requests": int(np.count_nonzero(sel)), "p50_ms": float(l50 * 1e3),
                "p99_ms": float(l99 * 1e3), "p99_over_p50": float(l99 / l50)}
    by_mech = {}
    for name, k in (("gas", 0), ("multiphase", 1)):
        sel = m == k
        if np.count_nonzero(sel) >= 5:
            by_mech[name] = {"requests": int(np.count_nonzero(sel)),
                             "p50_ms": float(np.percentile(lat[sel], 50) * 1e3),
                             "p99_ms": float(np.percentile(lat[sel], 99) * 1e3)}
    return {"requests": int(n_requests), "load": 0.7,
            "capacity_requests_per_s": cap_rps,
            "capacity_cell_timesteps_per_s": capacity,
            "throughput_cell_timesteps_per_s": float(np.sum(work)) / span,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p90_ms": float(np.percentile(lat, 90) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
            "by_request_size": by_size, "by_mechanism": by_mech,
            "failed_requests": int(failed + failed_cap),
            "mean_batch_cells": c["mean_batch_cells"], "batches": c["batches"],
            "request_mix": "gas 1..50000 cells or multi-phase 1..%d cells (equal odds), "
                           "log-uniform, dt U(60,600) s, 1..4 steps" % mp_max_cells}


def plain_path(mech, n_cells, steps, dt):
    """The plain GPU path on the 10,000-cell gas workload (T 220..310 K,
    P 2e4..1.013e5 Pa, no night or stressed cells): one BDF over all cells
    (shared step size and order, block-diagonal Newton) driven from the host
    with one launch per f(y), J(y), LU, solve and vector operation; one
    untimed step from the seeded state, then `steps` timed steps.  The
    persistent per-cell kernel runs the same cells and steps for the
    ratio."""
    import torch
    from camp_amd.core import CampSolver, MODE_BLOCK, MODE_PER_CELL
    from camp_amd.mechanisms import synthetic_cells
    cells = synthetic_cells(mech, n_cells, seed=2)
    dev = torch.device("cuda:0")
    env = torch.from_numpy(cells.env).to(dev)
    out = {}
    for name, mode in (("plain", MODE_BLOCK), ("persistent", MODE_PER_CELL)):
        st = torch.from_numpy(cells.state).to(dev)
        s = CampSolver(mech, n_cells, mode=mode)
        s.set_rates(cells.rates)
        progress("%s path: %d cells" % (name, n_cells))
        s.solve_device(st.data_ptr(), env.data_ptr(), 0.0, dt)
        torch.cuda.synchronize()
        t = time.perf_counter()
        failed = 0
        for _ in range(steps):
            failed += int(s.solve_device(st.data_ptr(), env.data_ptr(), 0.0, dt) != 0)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        progress("%s path done (%.0f ms/step)" % (name, 1e3 * el / steps))
        out[name] = {"value": n_cells * steps / el, "ms_per_step": 1e3 * el / steps,
                     "failed_steps": failed}
    return {"value": out["plain"]["value"], "unit": "cell-timesteps/s", "n_cells": n_cells,
            "steps": steps, "ms_per_step": out["plain"]["ms_per_step"],
            "failed_steps": out["plain"]["failed_steps"],
            "persistent_same_cells": out["persistent"],
            "persistent_over_plain": out["persistent"]["value"] / out["plain"]["value"],
            "what": "block mode: one BDF over all cells, one launch per f/J/LU/solve"}


def multiphase_leg(n_cells, steps, dt):
    """The multi-phase workload: the CB05-scale gas mechanism plus 8 aerosol
    bins x 30 species (20 SIMPOL partitioning species, 10 aqueous
    equilibria, ZSR water), n_cells cells, one untimed 180 s step from the
    seeded state, then `steps` timed steps, per-cell kernel."""
    import torch
    from camp_amd.core import CampSolver
    from camp_amd.multiphase import multiphase_cells, multiphase_mechanism
    mpm = multiphase_mechanism(1)
    cells = multiphase_cells(mpm, n_cells, seed=2)
    dev = torch.device("cuda:0")
    st = torch.from_numpy(cells.state).to(dev)
    env = torch.from_numpy(cells.env).to(dev)
    s = CampSolver(mpm, n_cells)
    s.set_rates(cells.rates)
    progress("multi-phase: %d cells" % n_cells)
    failed = int(s.solve_device(st.data_ptr(), env.data_ptr(), 0.0, dt) != 0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    kern = 0.0
    tot = np.zeros(7)
    for _ in range(steps):
        failed += int(s.solve_device(st.data_ptr(), env.data_ptr(), 0.0, dt) != 0)
        kern += s.last_kernel_ms()
        tot += s.cell_stats()[:, :7].sum(axis=0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    progress("multi-phase done (%.0f ms/step)" % (1e3 * el / steps))
    spun = (cells, st.cpu().numpy())
    return spun, {"value": n_cells * steps / el, "unit": "cell-timesteps/s", "n_cells": n_cells,
            "steps": steps, "ms_per_step": 1e3 * el / steps, "kernel_ms_per_step": kern / steps,
            "failed_steps": failed, "n_state": mpm.n_state, "n_dep": s.nv, "nnz": s.nnz,
            "bdf_steps_per_cell_step": float(tot[0] / n_cells / steps),
            "what": "cb05scale + 8 bins x 30 aerosol species (SIMPOL x20, aqueous "
                    "equilibrium x10, ZSR water), per-cell BDF, rtol 1e-4"}


def recorded_traffic(n_cells):
    """HBM bytes per launch of k_cell_bdf, k_deriv and k_jac from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes on this workload
    (calibrated on a device copy), or None when none was recorded for it."""
    path = os.path.join(HERE, "profiles", "hbm_traffic_cb05scale_100k.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    if rec["workload"]["n_cells"] != n_cells:
        return None
    return {k: v["bytes_per_launch"] for k, v in rec["kernels"].items()}


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100000)
    ap.add_argument("--dt", type=float, default=180.0)
    ap.add_argument("--cells-per-block", type=int, default=0)
    ap.add_argument("--placement", type=int, default=-1,
                    help="per-cell kernel cell placement (0: history in LDS, 1: in global "
                         "memory); with --cells-per-block skips the autotuner")
    ap.add_argument("--cpu-cores", type=int, default=min(16, len(os.sched_getaffinity(0))))
    ap.add_argument("--cpu-cells-per-core", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-cells", type=int, default=10000,
                    help="cells of the plain one-launch-per-RHS leg (0: skip)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--mp-cells", type=int, default=20000,
                    help="cells of the multi-phase leg (0: skip)")
    ap.add_argument("--mp-steps", type=int, default=2)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many MiB device-to-device three times (the "
                         "known-size kernel that calibrates rocprofv3's FETCH_SIZE and "
                         "WRITE_SIZE counters) and exit")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="requests in the serving leg at 70%% load (0: skip); the "
                         "capacity pass before it submits half as many at once")
    ap.add_argument("--serve-batch-cells", type=int, default=1 << 17,
                    help="cap on the cells of one server launch (the server's own chunk, "
                         "twice the cells the GPU holds at once, applies below it)")
    ap.add_argument("--serve-mp-cells", type=int, default=2000,
                    help="largest multi-phase request (0: gas requests only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (cells would shard "
              "across GPUs with no exchange step). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from camp_amd.core import CampSolver, MODE_PER_CELL

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n, "copies": 3}))
        return

    mech, cells = make_inputs(args.cells)
    dev = torch.device("cuda:0")
    state0 = torch.from_numpy(cells.state).to(dev)
    env = torch.from_numpy(cells.env).to(dev)

    def fresh():
        s = CampSolver(mech, args.cells, mode=MODE_PER_CELL)
        s.set_rates(cells.rates)
        if args.placement >= 0:
            s.set_placement(args.placement)
        if args.cells_per_block:
            s.set_cells_per_block(args.cells_per_block)
        return s

    solver = fresh()
    st = state0.clone()
    torch.cuda.synchronize()
    # warmup: the first W steps from the seeded state, untimed, same solver
    warm_failed = 0
    for _ in range(args.warmup):
        rc = solver.solve_device(st.data_ptr(), env.data_ptr(), 0.0, args.dt)
        warm_failed += int(np.sum(solver.cell_stats()[:, 7] < 0))
        progress("warmup step done (%.0f ms kernel)" % solver.last_kernel_ms())
    torch.cuda.synchronize()
    tune = None
    if not args.cells_per_block:
        # cells per workgroup picked by the native autotuner on a sample of
        # the spun-up cells (copies; the workload's cells are untouched)
        chosen, cand_ms = solver.autotune(st.data_ptr(), env.data_ptr(), 0.0, args.dt,
                                          n_sample=min(args.cells, 16384))
        tune = {"placement": ("history in LDS", "history in global memory",
                              "history and LU in global memory")[chosen[0]],
                "cells_per_workgroup": chosen[1],
                "candidate_ms": {"%s/%d" % (("lds", "global", "global_lu")[p], c): v
                                 for (p, c), v in cand_ms.items() if v is not None},
                "sample_cells": min(args.cells, 16384)}
        progress("autotuned: %s %s" % (chosen, tune["candidate_ms"]))
    torch.cuda.synchronize()
    step_ms, kern_ms, stats = [], [], []
    failed_cell_steps = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rc = solver.solve_device(st.data_ptr(), env.data_ptr(), 0.0, args.dt)
        step_ms.append(1e3 * (time.perf_counter() - ts))
        kern_ms.append(solver.last_kernel_ms())
        cs = solver.cell_stats()
        stats.append(cs[:, :7].sum(axis=0))
        failed_cell_steps += int(np.sum(cs[:, 7] < 0))
        progress("timed step done (%.0f ms)" % step_ms[-1])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.cells * args.steps / elapsed

    # the same solver through solver_run on host arrays (state and env
    # copied over PCIe inside each call), two steps after the timed ones
    host_st = st.cpu().numpy().copy()
    host_env = np.ascontiguousarray(cells.env).copy()
    t_h = time.perf_counter()
    k_h = 0.0
    for _ in range(2):
        solver.solve(host_st, host_env, 0.0, args.dt)
        k_h += solver.last_kernel_ms()
    el_h = time.perf_counter() - t_h
    host_leg = {"value": args.cells * 2 / el_h, "unit": "cell-timesteps/s",
                "ms_per_step": 1e3 * el_h / 2, "kernel_ms_per_step": k_h / 2,
                "what": "solver_run on host arrays (state and env copied to and from HBM in "
                        "each call), two steps after the timed ones"}
    progress("host-array solver_run: %.0f ms/step (kernel %.0f)" % (1e3 * el_h / 2, k_h / 2))

    flops, bytes_ = solver.work_model()
    tot = np.sum(stats, axis=0)  # nst nfe nsetups netf nni ncfn nje
    nfe, nsetups, nni, nje = tot[1], tot[2], tot[4], tot[6]
    total_flops = (nfe * flops["f"] + nje * (flops["jac"] + flops["f"]) + nsetups * flops["lu"]
                   + nni * flops["solve"])
    kernel_s = sum(kern_ms) / 1e3
    achieved_tf = total_flops / kernel_s / 1e12
    from camp_amd.mechanisms import pack
    n_env = sum(p[3] for p in pack(mech))
    # per cell: state in+out, env in, rate data in+out, J history in+out
    hbm_bytes_per_launch = 8.0 * args.cells * (2 * mech.n_state + 2 + 2 * n_env
                                               + 2 * (solver.nnz + 2 * solver.nv))
    traffic = recorded_traffic(args.cells)
    traffic = traffic or {}
    roofline = {"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS,
                "traffic": traffic.get("k_cell_bdf"),
                "kernel": "k_cell_bdf", "avg_launch_ms": kernel_s * 1e3 / args.steps,
                "flops_per_launch": total_flops / args.steps,
                "algorithmic_hbm_bytes_per_launch": hbm_bytes_per_launch,
                "placement": ("history in LDS", "history in global memory",
                              "history and LU in global memory")[solver.set_placement(-1)],
                "note": "FP64 VALU work (peak 78.6 TFLOP/s is the FP64 figure); algorithmic "
                        "HBM bytes are state, env, rates and J history in and out; with the "
                        "history in global memory the cells' Nordsieck history, work vectors "
                        "and J history also stream through L2 ('traffic')"}

    # plain-path f/J kernels against HBM (lanes per cell tuned first)
    fj_lanes, fj_cand = solver.tune_fj(st.data_ptr(), env.data_ptr(), reps=10)
    ms_f, ms_j = solver.time_fj(st.data_ptr(), env.data_ptr(), reps=10)
    gbs_f = args.cells * bytes_["f"] / (ms_f / 1e3) / 1e9
    gbs_j = args.cells * bytes_["jac"] / (ms_j / 1e3) / 1e9
    roofline_fj = {
        "deriv": {"bound": "hbm", "achieved": gbs_f, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": gbs_f / HBM_PEAK_GBS, "avg_launch_ms": ms_f,
                  "bytes_per_cell": bytes_["f"], "lanes_per_cell": fj_lanes[0],
                  "candidate_ms": fj_cand["f"],
                  "traffic": traffic.get("k_deriv_g") if traffic else None},
        "jac": {"bound": "hbm", "achieved": gbs_j, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs_j / HBM_PEAK_GBS, "avg_launch_ms": ms_j,
                "bytes_per_cell": bytes_["jac"], "lanes_per_cell": fj_lanes[1],
                "candidate_ms": fj_cand["jac"],
                "traffic": traffic.get("k_jac_g") if traffic else None},
        "what": "the block path's lane-group f(y) and J(y) kernels (no blend), rate "
                "constants already computed, on the 100k-cell workload's spun-up state"}

    plain = None
    if args.plain_cells > 0:
        plain = plain_path(mech, args.plain_cells, args.plain_steps, args.dt)
        if not args.no_cpu:
            progress("cpu block-mode baseline on %d cores" % args.cpu_cores)
            plain["cpu_baseline_block"] = cpu_baseline_block(args.plain_cells, 1,
                                                             args.plain_steps, args.dt,
                                                             args.cpu_cores, 100)

    mpleg = mp_spun = None
    if args.mp_cells > 0:
        mp_spun, mpleg = multiphase_leg(args.mp_cells, args.mp_steps, args.dt)

    serve = None
    if args.serve_requests > 0:
        progress("serving leg: %d requests" % args.serve_requests)
        serve = serving(mech, solver, st.cpu().numpy(), cells.env, cells.rates,
                        args.serve_requests, args.serve_mp_cells, mp_spun, args.dt,
                        args.serve_batch_cells)
        progress("serving leg done")

    cpu = box = None
    if not args.no_cpu:
        progress("cpu baseline on %d cores" % args.cpu_cores)
        cpu = cpu_baseline(args.cells, args.warmup, args.steps, args.dt, args.cpu_cores,
                           args.cpu_cells_per_core)
        progress("box model leg")
        box = box_model_leg()
    half = max(1, args.steps // 2)

    out = {
        "metric": "cell-timesteps/s at 100k cells, CB05-scale FP64; p50/p99 request latency",
        "value": value, "unit": "cell-timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cb05scale72x190 batched per-cell BDF: %d cells (50%% night, "
                               "10%% x1e3), %d x %g s steps after %d untimed from the seeded "
                               "initial state, rtol 1e-4, atol 1e-8" % (
                                   args.cells, args.steps, args.dt, args.warmup),
                   "n_cells": args.cells, "dt_s": args.dt, "mechanism_seed": 1, "cell_seed": 2,
                   "parallelism": "1 GPU"},
        "latency_ms": ({"p50": serve["p50_ms"], "p99": serve["p99_ms"],
                        "what": "serving request latency at 70% load"} if serve else None),
        "step_ms": {"p50": float(np.percentile(step_ms, 50)),
                    "p99": float(np.percentile(step_ms, 99)),
                    "each": [round(x, 2) for x in step_ms],
                    "what": "one 180 s step over all cells"},
        "steady_state": {"value": args.cells * half / (sum(step_ms[-half:]) / 1e3),
                         "unit": "cell-timesteps/s",
                         "what": "the last %d timed steps only (the cells are still relaxing "
                                 "from the seeded state: earlier steps take longer)" % half},
        "failed_cell_steps": {"timed": failed_cell_steps, "warmup": warm_failed},
        "serving": serve,
        "solver_stats_per_cell_step": {"steps": float(tot[0] / args.cells / args.steps),
                                       "rhs_evals": float(nfe / args.cells / args.steps),
                                       "lu_setups": float(nsetups / args.cells / args.steps),
                                       "newton_iters": float(nni / args.cells / args.steps),
                                       "jac_evals": float(nje / args.cells / args.steps)},
        "plain_path": plain,
        "multiphase": mpleg,
        "autotune": tune,
        "roofline": roofline,
        "roofline_fj": roofline_fj,
        "cpu_baseline": cpu,
        "box_model": box,
        "host_arrays": host_leg,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
