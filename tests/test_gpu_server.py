"""GPU tests of the request server: a request's results equal the solver's
own per-cell run, are bit-identical alone and in a mix of requests (other
cells, other time steps, another mechanism), state stays resident across a
request's timesteps, and admission waits for arena slots to free up."""
import numpy as np
import pytest

from camp_amd.core import CampSolver
from camp_amd.server import CampServer, base_rate_matrix
from tests.helpers import box_inputs, cb05_inputs

pytestmark = pytest.mark.gpu


def _cb05(n, seed=2):
    mech, cells = cb05_inputs(n, night_fraction=0.5, stressed_fraction=0.1)
    return mech, cells


def test_request_equals_solver_run():
    n = 48
    mech, cells = _cb05(n)
    ref = CampSolver(mech, n)
    ref.set_rates(cells.rates)
    a = cells.state.copy()
    for _ in range(2):
        assert ref.solve(a, cells.env, 0.0, 180.0) == 0
    srv = CampServer(n_streams=2)
    m = srv.add_mechanism(CampSolver(mech, 1), max_cells=1024)
    br = base_rate_matrix(ref, cells.rates, n)
    r = srv.submit(m, cells.state, cells.env, br, 180.0, n_steps=2, want_stats=True)
    assert r.wait()
    assert np.array_equal(r.state, a)
    assert np.array_equal(r.stats, ref.cell_stats())
    assert r.latency_s > 0
    srv.close()


def test_request_larger_than_a_launch_chunk():
    # a launch takes at most twice the cells the GPU holds at once (2 048
    # gas cells on MI355X): this request spans launches on several streams,
    # and a small request submitted after it interleaves with its chunks
    n = 2600
    mech, cells = _cb05(n)
    ref = CampSolver(mech, n)
    ref.set_rates(cells.rates)
    a = cells.state.copy()
    assert ref.solve(a, cells.env, 0.0, 180.0) == 0
    srv = CampServer(n_streams=4)
    m = srv.add_mechanism(CampSolver(mech, 1), max_cells=8192)
    br = base_rate_matrix(ref, cells.rates, n)
    big = srv.submit(m, cells.state, cells.env, br, 180.0, n_steps=1, want_stats=True)
    small = srv.submit(m, cells.state[:5], cells.env[:5], br[:5], 180.0, n_steps=1)
    assert big.wait() and small.wait()
    assert np.array_equal(big.state, a)
    assert np.array_equal(big.stats, ref.cell_stats())
    assert np.array_equal(small.state, a[:5])
    c = srv.counters()
    assert c["batches"] >= 2 and c["done"] == 2 and c["failed"] == 0
    srv.close()


def test_alone_equals_mix_across_requests_and_mechanisms():
    n = 32
    mech, cells = _cb05(n)
    bmech, bcells = box_inputs(4)
    srv = CampServer(n_streams=4)
    s_cb05 = CampSolver(mech, 1)
    m = srv.add_mechanism(s_cb05, max_cells=4096)
    mb = srv.add_mechanism(CampSolver(bmech, 1), max_cells=256)
    br = base_rate_matrix(s_cb05, cells.rates, n)
    alone = srv.submit(m, cells.state[:8], cells.env[:8], br[:8], 120.0, n_steps=2)
    assert alone.wait()
    # the same cells again, submitted amid other work of both mechanisms
    others = [srv.submit(m, cells.state[8:], cells.env[8:], br[8:], 300.0, n_steps=1),
              srv.submit(mb, bcells.state, bcells.env, np.zeros((4, 0)), 60.0, n_steps=3)]
    mixed = srv.submit(m, cells.state[:8], cells.env[:8], br[:8], 120.0, n_steps=2)
    others.append(srv.submit(m, cells.state[4:20], cells.env[4:20], br[4:20], 60.0, n_steps=1))
    assert mixed.wait() and all(o.wait() for o in others)
    assert np.array_equal(alone.state, mixed.state)
    c = srv.counters()
    assert c["done"] == 5 and c["failed"] == 0
    srv.close()


def test_admission_waits_for_free_slots():
    n = 40
    mech, cells = _cb05(n)
    srv = CampServer(n_streams=2)
    s = CampSolver(mech, 1)
    m = srv.add_mechanism(s, max_cells=64)
    assert srv.capacity(m) == 64
    br = base_rate_matrix(s, cells.rates, n)
    with pytest.raises(ValueError):
        srv.submit(m, np.tile(cells.state, (2, 1)), np.tile(cells.env, (2, 1)),
                   np.tile(br, (2, 1)), 60.0)
    # 3 x 40 cells through a 64-cell arena: each waits for the previous one
    reqs = [srv.submit(m, cells.state, cells.env, br, 60.0) for _ in range(3)]
    assert all(r.wait() for r in reqs)
    assert np.array_equal(reqs[0].state, reqs[2].state)
    assert srv.counters()["peak_resident"] <= 64
    srv.close()


def test_replay_poisson_requests():
    n_pool = 2000
    mech, cells = _cb05(n_pool)
    srv = CampServer(n_streams=4)
    s = CampSolver(mech, 1)
    m = srv.add_mechanism(s, max_cells=8192)
    srv.set_pool(m, cells.state, cells.env, base_rate_matrix(s, cells.rates, n_pool))
    rng = np.random.default_rng(5)
    k = 40
    t = np.cumsum(rng.exponential(0.002, size=k))
    sizes = np.exp(rng.uniform(0, np.log(500), size=k)).astype(np.int32)
    off = rng.integers(0, n_pool - 500, size=k)
    lat, span, failed = srv.replay(t, np.zeros(k, np.int32), sizes, off,
                                   rng.uniform(60, 600, size=k), rng.integers(1, 3, size=k))
    assert failed == 0
    assert np.all(lat > 0) and span >= t[-1]
    c = srv.counters()
    assert c["done"] == k and c["cell_steps"] > 0
    srv.close()


def _mp(n):
    from camp_amd.multiphase import multiphase_cells, multiphase_mechanism
    mpm = multiphase_mechanism(1)
    return mpm, multiphase_cells(mpm, n, seed=2)


def test_multiphase_request_equals_solver_run_and_is_the_same_in_a_mix():
    """Multi-phase requests through the server: a request of more cells
    than one launch takes (1 100 cells x 2 timesteps > 2 048 cell-timesteps)
    equals solver_run bit for bit; a small request gives the same bits
    alone and amid gas and multi-phase traffic."""
    n = 1100
    mpm, mcells = _mp(n)
    ref = CampSolver(mpm, n)
    ref.set_rates(mcells.rates)
    a = mcells.state.copy()
    for _ in range(2):
        assert ref.solve(a, mcells.env, 0.0, 120.0) == 0
    gmech, gcells = _cb05(64)
    srv = CampServer(n_streams=4)
    msol = CampSolver(mpm, 1)
    mm = srv.add_mechanism(msol, max_cells=4096)
    gsol = CampSolver(gmech, 1)
    mg = srv.add_mechanism(gsol, max_cells=4096)
    mbr = base_rate_matrix(msol, mcells.rates, n)
    gbr = base_rate_matrix(gsol, gcells.rates, 64)
    alone = srv.submit(mm, mcells.state[:6], mcells.env[:6], mbr[:6], 300.0, n_steps=1)
    assert alone.wait()
    big = srv.submit(mm, mcells.state, mcells.env, mbr, 120.0, n_steps=2, want_stats=True)
    others = [srv.submit(mg, gcells.state, gcells.env, gbr, 180.0, n_steps=3)]
    mixed = srv.submit(mm, mcells.state[:6], mcells.env[:6], mbr[:6], 300.0, n_steps=1)
    others.append(srv.submit(mm, mcells.state[40:90], mcells.env[40:90], mbr[40:90], 60.0,
                             n_steps=2))
    assert big.wait() and mixed.wait() and all(o.wait() for o in others)
    assert np.array_equal(big.state, a)
    assert np.array_equal(big.stats, ref.cell_stats())
    assert np.array_equal(alone.state, mixed.state)
    c = srv.counters()
    assert c["done"] == 5 and c["failed"] == 0 and c["batches"] >= 4
    srv.close()
