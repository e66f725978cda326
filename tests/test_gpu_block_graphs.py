"""Block mode with each Newton iteration replayed as HIP graphs against the
same solver launching kernel by kernel: results and statistics identical
bit for bit.

The graphs are replayed hundreds of times per solver_run, with the
solver's own eager work (error tests, Jacobian rebuilds, LU) and
synchronisations between the replays; the check covers the host-array
path (the solver's internal state buffer), the device-pointer path (graph
B recaptured for the caller's state pointer), a change of the f kernel's
width between calls (graph B recaptured), and several solver_run calls in
a row.
"""
import numpy as np
import pytest

from camp_amd.core import CampSolver, MODE_BLOCK
from tests.helpers import cb05_inputs

pytestmark = pytest.mark.gpu

KEYS = ("solver_flag", "num_steps", "RHS_evals", "LS_setups", "error_test_fails", "NLS_iters",
        "NLS_convergence_fails", "DLS_Jac_evals", "last_time_step__s", "next_time_step__s")


def _run(graphs, n=500, calls=3, device=False, lanes=None):
    import torch
    mech, cells = cb05_inputs(n)
    s = CampSolver(mech, n, mode=MODE_BLOCK)
    s.set_rates(cells.rates)
    assert s.set_graphs(graphs) in (0, 1)
    st = cells.state.copy()
    out = []
    if device:
        dev = torch.device("cuda:0")
        d_st = torch.from_numpy(st).to(dev)
        d_env = torch.from_numpy(np.ascontiguousarray(cells.env)).to(dev)
    for call in range(calls):
        if lanes is not None:
            s.set_fj_lanes(lanes[call % len(lanes)], 16)
        if device:
            assert s.solve_device(d_st.data_ptr(), d_env.data_ptr(), 0.0, 180.0) == 0
            out.append(d_st.cpu().numpy().copy())
        else:
            assert s.solve(st, cells.env, 0.0, 180.0) == 0
            out.append(st.copy())
        d = s.stats()
        out.append(np.array([d[k] for k in KEYS], dtype=float))
    s.close()
    return out


@pytest.mark.parametrize("device", [False, True])
def test_graphs_equal_eager_bitwise(device):
    a = _run(True, device=device)
    b = _run(False, device=device)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_graphs_recapture_on_width_change():
    a = _run(True, calls=3, lanes=[16, 8, 32])
    b = _run(False, calls=3, lanes=[16, 8, 32])
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
