"""Per-cell results do not depend on which wavefront integrates a cell, in
what order, or what the wavefront did before.

More cells than the GPU holds at once (4 096 gas cells of the CB05-scale
workload with night and x1e3 cells, 3 072 multi-phase cells), two
solver_run calls each (the second uses the J history the first left), are
solved in index order, in a random permutation with every wavefront's
working memory filled with NaNs before each cell, in reverse order with the
cell history in global memory, permuted and poisoned with the history and
LU factors in global memory, and longest-first by the first call's step
counts; then the same cells go through the request server as interleaved
requests of both mechanisms.  Every result is compared bit for bit.  This is
the guarantee of the reference's multi- versus single-cell test
(unit_test_driver.F90:262-285) taken to its limit: a cell's result is the
same alone, in any batch and in any order.
"""
import numpy as np
import pytest

from camp_amd.core import CampSolver
from camp_amd.mechanisms import cb05_scale_mechanism, synthetic_cells
from camp_amd.multiphase import multiphase_cells, multiphase_mechanism
from camp_amd.server import CampServer, base_rate_matrix

pytestmark = pytest.mark.gpu

N_GAS, N_MP, DT = 4096, 3072, 180.0


def _inputs(kind):
    if kind == "gas":
        mech = cb05_scale_mechanism(1)
        return mech, synthetic_cells(mech, N_GAS, seed=2, night_fraction=0.5,
                                     stressed_fraction=0.1)
    mech = multiphase_mechanism(1)
    return mech, multiphase_cells(mech, N_MP, seed=2)


def _two_calls(mech, cells, order=None, poison=False, place=0, longest=False):
    import torch
    n = cells.state.shape[0]
    s = CampSolver(mech, n)
    s.set_rates(cells.rates)
    s.set_placement(place)
    s.set_poison(poison)
    dev = torch.device("cuda:0")
    st = torch.from_numpy(cells.state.copy()).to(dev)
    env = torch.from_numpy(np.ascontiguousarray(cells.env)).to(dev)
    out = []
    for call in range(2):
        if longest and call == 1:
            s.set_cell_order(np.argsort(-stats[:, 0], kind="stable"))
        else:
            s.set_cell_order(order)
        assert s.solve_device(st.data_ptr(), env.data_ptr(), 0.0, DT) == 0
        stats = s.cell_stats()
        out.append((st.cpu().numpy().copy(), stats.copy()))
    s.close()
    return out


@pytest.fixture(scope="module", params=["gas", "mp"])
def reference_run(request):
    mech, cells = _inputs(request.param)
    return request.param, mech, cells, _two_calls(mech, cells)


def _same(a, b):
    for (sa, ta), (sb, tb) in zip(a, b):
        assert np.array_equal(sa, sb)
        assert np.array_equal(ta, tb)


def test_results_do_not_depend_on_order_placement_or_leftover_memory(reference_run):
    kind, mech, cells, ref = reference_run
    n = cells.state.shape[0]
    perm = np.random.default_rng(7).permutation(n)
    _same(ref, _two_calls(mech, cells, order=perm, poison=True))
    _same(ref, _two_calls(mech, cells, order=np.arange(n)[::-1], place=1))
    _same(ref, _two_calls(mech, cells, order=perm, poison=True, place=2))
    _same(ref, _two_calls(mech, cells, longest=True, poison=True))


def test_server_requests_interleaved_equal_solver_run():
    """The same cells as interleaved requests of both mechanisms, each
    request its cells for two timesteps: every request equals the rows of
    the two-call solver_run above."""
    gm, gc = _inputs("gas")
    mm, mc = _inputs("mp")
    gref = _two_calls(gm, gc)[1][0]
    mref = _two_calls(mm, mc)[1][0]
    srv = CampServer(n_streams=4)
    gs, ms = CampSolver(gm, 1), CampSolver(mm, 1)
    g = srv.add_mechanism(gs, max_cells=N_GAS)
    m = srv.add_mechanism(ms, max_cells=N_MP)
    gbr = base_rate_matrix(gs, gc.rates, N_GAS)
    mbr = base_rate_matrix(ms, mc.rates, N_MP)
    rng = np.random.default_rng(11)
    cuts_g = np.sort(rng.choice(np.arange(1, N_GAS), 7, replace=False))
    cuts_m = np.sort(rng.choice(np.arange(1, N_MP), 5, replace=False))
    pieces = ([(g, a, b) for a, b in zip(np.r_[0, cuts_g], np.r_[cuts_g, N_GAS])] +
              [(m, a, b) for a, b in zip(np.r_[0, cuts_m], np.r_[cuts_m, N_MP])])
    rng.shuffle(pieces)
    reqs = []
    for mech, a, b in pieces:
        c, br = (gc, gbr) if mech == g else (mc, mbr)
        reqs.append((mech, a, b, srv.submit(mech, c.state[a:b], c.env[a:b], br[a:b], DT,
                                            n_steps=2)))
    for mech, a, b, r in reqs:
        assert r.wait()
        want = gref if mech == g else mref
        assert np.array_equal(r.state, want[a:b]), (mech, a, b)
    srv.close()
