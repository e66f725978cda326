"""Benchmark: cell-timesteps/s of the batched per-cell BDF solver on one MI355X.

Workload (default): the CB05-scale synthetic gas mechanism (72 species, 190
reactions: 133 Arrhenius, 28 Troe, 19 photolysis, 5 emission, 5 first-order
loss; mechanism seed 1) over 100,000 cells (cell seed 2) of uneven
stiffness: 50% night cells (photolysis J = 0) and 10% with concentrations
scaled by 1e3; FP64, rtol 1e-4, atol 1e-8 ppm.  One step = one 180 s
solver_run over all cells, each cell with its own BDF order, step size and
Newton control, on state already resident in HBM.  K consecutive steps are
timed from the seeded initial state (the first step carries the stiff
start-up transient); warmup runs the same steps on a separate solver.

Prints one JSON line.  `roofline` is for the dominant kernel (k_cell_bdf,
the persistent per-cell kernel): FP64 algorithmic flops (per-evaluation
counts of f, J, LU and triangular solves, times the evaluation counts the
kernel itself reports) over the kernel's HIP-event duration.
`roofline_fj` gives the plain-path f(y) and J(y) kernels against HBM.
`cpu_baseline` is the CPU oracle (a C restatement of the reference path)
on a bounded sample of the same cells, on this host's cores.
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


def _cpu_worker(args):
    n_cells, c0, c1, steps, dt = args
    from oracle.oracle import Oracle
    mech, cells = make_inputs(n_cells)
    sub = {k: v[c0:c1] for k, v in cells.rates.items()}
    o = Oracle(mech, c1 - c0)
    o.set_rates(sub)
    st = np.ascontiguousarray(cells.state[c0:c1])
    env = np.ascontiguousarray(cells.env[c0:c1])
    t = time.perf_counter()
    for _ in range(steps):
        o.run(st, env, 0.0, dt, mode=1)
    return time.perf_counter() - t


def cpu_baseline(n_cells, steps, dt, cores, per_core):
    """Oracle per-cell solve of the first cores*per_core cells, K steps."""
    from oracle import oracle
    oracle.build(ref=False)
    jobs = [(n_cells, i * per_core, (i + 1) * per_core, steps, dt) for i in range(cores)]
    t = time.perf_counter()
    if cores == 1:
        _cpu_worker(jobs[0])
    else:
        with mp.get_context("spawn").Pool(cores) as pool:
            pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t
    return {"value": cores * per_core * steps / wall, "unit": "cell-timesteps/s", "cores": cores,
            "kind": "port",
            "sample": "%d cells (cells 0..%d of the same workload) x %d steps of %g s, "
                      "per-cell CVODE-algorithm BDF, %d processes" % (
                          cores * per_core, cores * per_core - 1, steps, dt, cores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100000)
    ap.add_argument("--dt", type=float, default=180.0)
    ap.add_argument("--cells-per-block", type=int, default=0)
    ap.add_argument("--cpu-cores", type=int, default=8)
    ap.add_argument("--cpu-cells-per-core", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (cells shard "
              "across GPUs with no exchange; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from camp_amd.core import CampSolver, MODE_PER_CELL

    mech, cells = make_inputs(args.cells)
    dev = torch.device("cuda:0")
    state0 = torch.from_numpy(cells.state).to(dev)
    env = torch.from_numpy(cells.env).to(dev)

    def fresh():
        s = CampSolver(mech, args.cells, mode=MODE_PER_CELL)
        s.set_rates(cells.rates)
        if args.cells_per_block:
            s.set_cells_per_block(args.cells_per_block)
        return s

    # warmup: same steps on a separate solver instance (kernels loaded, HBM paged)
    if args.warmup > 0:
        w = fresh()
        st = state0.clone()
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            rc = w.solve_device(st.data_ptr(), env.data_ptr(), 0.0, args.dt)
            assert rc == 0, "warmup solve failed"
        w.close()

    solver = fresh()
    st = state0.clone()
    torch.cuda.synchronize()
    step_ms, kern_ms, stats = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rc = solver.solve_device(st.data_ptr(), env.data_ptr(), 0.0, args.dt)
        step_ms.append(1e3 * (time.perf_counter() - ts))
        kern_ms.append(solver.last_kernel_ms())
        stats.append(solver.cell_stats()[:, :7].sum(axis=0))
        if rc != 0:
            print("solve failed", file=sys.stderr)
            sys.exit(1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.cells * args.steps / elapsed

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
    roofline = {"bound": "mfma", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_cell_bdf", "avg_launch_ms": kernel_s * 1e3 / args.steps,
                "flops_per_launch": total_flops / args.steps,
                "algorithmic_hbm_bytes_per_launch": hbm_bytes_per_launch,
                "note": "FP64 VALU work (peak 78.6 TFLOP/s is the FP64 figure); the cell "
                        "stays in LDS, HBM moves only state, rates and J history"}

    # plain-path f/J kernels against HBM
    ms_f, ms_j = solver.time_fj(st.data_ptr(), env.data_ptr(), reps=10)
    gbs_f = args.cells * bytes_["f"] / (ms_f / 1e3) / 1e9
    gbs_j = args.cells * bytes_["jac"] / (ms_j / 1e3) / 1e9
    roofline_fj = {
        "deriv": {"bound": "hbm", "achieved": gbs_f, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": gbs_f / HBM_PEAK_GBS, "avg_launch_ms": ms_f,
                  "bytes_per_cell": bytes_["f"]},
        "jac": {"bound": "hbm", "achieved": gbs_j, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs_j / HBM_PEAK_GBS, "avg_launch_ms": ms_j,
                "bytes_per_cell": bytes_["jac"]}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cells, args.steps, args.dt, args.cpu_cores, args.cpu_cells_per_core)

    out = {
        "metric": "cell-timesteps/s at 100k cells, CB05-scale FP64; p50/p99 request latency",
        "value": value, "unit": "cell-timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cb05scale72x190 batched per-cell BDF: %d cells (50%% night, "
                               "10%% x1e3), %d x %g s steps from the seeded initial state, "
                               "rtol 1e-4, atol 1e-8" % (args.cells, args.steps, args.dt),
                   "n_cells": args.cells, "dt_s": args.dt, "mechanism_seed": 1, "cell_seed": 2,
                   "parallelism": "1 GPU"},
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99)),
                       "what": "one 180 s step over all cells"},
        "solver_stats_per_cell_step": {"steps": float(tot[0] / args.cells / args.steps),
                                       "rhs_evals": float(nfe / args.cells / args.steps),
                                       "lu_setups": float(nsetups / args.cells / args.steps),
                                       "newton_iters": float(nni / args.cells / args.steps),
                                       "jac_evals": float(nje / args.cells / args.steps)},
        "roofline": roofline,
        "roofline_fj": roofline_fj,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
