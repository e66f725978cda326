"""Accuracy of the benchmark-size runs, checked on samples of their cells.

The 1e-4 / 1e-8 per-species bar lies below the integration error of any
BDF at rtol 1e-4 (the oracle against itself under a 1e-15 perturbation of
T moves by tens of bar units on these cells), so each run is held to a
converged solution -- the oracle per cell at rtol 1e-10, atol 1e-14 -- and
its error there, in bar units, to the oracle's own at the run's tolerance:
the median and 90th percentile of the cells' worst errors may not exceed
twice the oracle's.

* the per-cell path on the 100,000-cell benchmark workload (50 % night
  cells, 10 % scaled x1e3): the start-up step and one more 180 s step of
  all cells on the GPU, 512 cells spread over the workload checked;
* the block path (one BDF over all cells) on the 10,000-cell plain-path
  workload: two 180 s steps, 256 cells checked.  Its shared step is set by
  the stiffest cell, so its cells are at least as resolved as per-cell
  integration at the same tolerance.
"""
import numpy as np
import pytest

from camp_amd.core import CampSolver, MODE_BLOCK, MODE_PER_CELL
from tests.helpers import bar_units, cb05_inputs, oracle_per_cell

pytestmark = pytest.mark.gpu

DT = 180.0


def _gpu_steps(mech, cells, n, mode, steps):
    import torch
    dev = torch.device("cuda:0")
    s = CampSolver(mech, n, mode=mode)
    s.set_rates(cells.rates)
    st = torch.from_numpy(cells.state.copy()).to(dev)
    env = torch.from_numpy(np.ascontiguousarray(cells.env)).to(dev)
    for _ in range(steps):
        assert s.solve_device(st.data_ptr(), env.data_ptr(), 0.0, DT) == 0
    out = st.cpu().numpy()
    s.close()
    return out


def _check(mech, cells, a, idx, steps, label):
    rates = {k: v[idx] for k, v in cells.rates.items()}
    st, env = cells.state[idx], cells.env[idx]
    t = oracle_per_cell(mech, st, env, rates, 1e-10, 1e-14, steps, DT)
    b = oracle_per_cell(mech, st, env, rates, 1e-4, 1e-8, steps, DT)
    assert np.all(np.isfinite(a[idx]))
    e_hip, e_orc = bar_units(a[idx], t), bar_units(b, t)
    g50, g90 = np.percentile(e_hip, 50), np.percentile(e_hip, 90)
    o50, o90 = np.percentile(e_orc, 50), np.percentile(e_orc, 90)
    print("%s: error vs converged in bar units (median, p90): hip %.3g %.3g, oracle %.3g %.3g"
          % (label, g50, g90, o50, o90))
    assert g50 <= 2.0 * o50
    assert g90 <= 2.0 * o90


def test_bench_workload_sample_per_cell():
    n = 100000
    mech, cells = cb05_inputs(n, night_fraction=0.5, stressed_fraction=0.1)
    a = _gpu_steps(mech, cells, n, MODE_PER_CELL, 2)
    idx = np.arange(0, n, n // 512)[:512]
    _check(mech, cells, a, idx, 2, "per-cell, 1e5 cells")


def test_plain_path_workload_sample_block():
    n = 10000
    mech, cells = cb05_inputs(n)
    a = _gpu_steps(mech, cells, n, MODE_BLOCK, 2)
    idx = np.arange(0, n, n // 256)[:256]
    _check(mech, cells, a, idx, 2, "block, 1e4 cells")
