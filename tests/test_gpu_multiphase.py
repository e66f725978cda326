"""GPU parity of the multi-phase path (aerosol bins, SIMPOL partitioning,
aqueous equilibria, ZSR water) against the CPU oracle: rate data, f(y) and
J(y) within 1e-12 relative, and per-cell solves within the 1e-4 / 1e-8 bar
at a tolerance where the problem is well conditioned."""
import numpy as np
import pytest

from camp_amd.core import CampSolver
from camp_amd.multiphase import multiphase_cells, multiphase_mechanism

pytestmark = pytest.mark.gpu

F_RTOL = 1e-12
# Below ~1e-290 the low word of the device's double-double accumulators is
# subnormal, so magnitudes there (physically zero) carry fewer digits than
# the 80-bit long double the solver accumulates in.
F_FLOOR = 1e-280


@pytest.fixture(scope="module", params=["small", "cfg4"])
def case(request):
    if request.param == "small":
        mp = multiphase_mechanism(1, n_bins=3, n_simpol=5, n_aqeq=4)
    else:
        mp = multiphase_mechanism(1)
    return mp, multiphase_cells(mp, 16, seed=2)


def _pair(oracle_lib, mp, cells, n, **kw):
    sol = CampSolver(mp, n, **kw)
    orc = oracle_lib.Oracle(mp, n, **kw)
    rates = {k: v[:n] for k, v in cells.rates.items()}
    sol.set_rates(rates)
    orc.set_rates(rates)
    return sol, orc


def _states(mp, cells, n):
    """Seeded states and the same with every aerosol solute scaled, so that
    both dilute and concentrated bins are exercised."""
    st = cells.state[:n].copy()
    st[n // 2:, mp.n_gas:] *= 30.0
    return st


def test_multiphase_rate_data_matches_oracle(oracle_lib, case):
    mp, cells = case
    n = 16
    sol, orc = _pair(oracle_lib, mp, cells, n)
    rc, _ = sol.eval_deriv(cells.state[:n], cells.env[:n])
    assert rc == 0
    orc.update_env(cells.env[:n])
    a, b = sol.rxn_env(), orc.rxn_env()
    nz = b != 0
    assert np.array_equal(a == 0, b == 0)
    assert np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz])) <= F_RTOL


@pytest.mark.parametrize("use_est", [False, True])
def test_multiphase_deriv_matches_oracle(oracle_lib, case, use_est):
    mp, cells = case
    n = 16
    sol, orc = _pair(oracle_lib, mp, cells, n)
    st = _states(mp, cells, n)
    orc.update_env(cells.env[:n])
    if use_est:
        st0 = st * 1.01
        assert sol.eval_jac(st0, cells.env[:n], 10.0)[0] == 0
        for c in range(n):
            assert orc.eval_jac(c, st0[c], 10.0)[0] == 0
    rc, f_gpu = sol.eval_deriv(st, cells.env[:n], 10.0, use_est)
    assert rc == 0
    f_ref = np.stack([orc.eval_f(c, st[c], 10.0, use_est)[1] for c in range(n)])
    err = np.abs(f_gpu - f_ref)
    worst = float(np.max(err / (np.abs(f_ref) + F_FLOOR / F_RTOL)))
    print("multi-phase f worst rel", worst)
    assert (err <= F_RTOL * np.abs(f_ref) + F_FLOOR).all(), worst


def test_multiphase_jac_matches_oracle(oracle_lib, case):
    mp, cells = case
    n = 16
    sol, orc = _pair(oracle_lib, mp, cells, n)
    st = _states(mp, cells, n)
    rc, j_gpu = sol.eval_jac(st, cells.env[:n], 10.0)
    assert rc == 0
    orc.update_env(cells.env[:n])
    j_ref = np.stack([orc.eval_jac(c, st[c], 10.0)[1] for c in range(n)])
    err = np.abs(j_gpu - j_ref)
    worst = float(np.max(err / (np.abs(j_ref) + F_FLOOR / F_RTOL)))
    print("multi-phase J worst rel", worst)
    assert (err <= F_RTOL * np.abs(j_ref) + F_FLOOR).all(), worst


def test_multiphase_per_cell_solve_matches_oracle(oracle_lib, case):
    """Two 180 s steps per cell at rtol 1e-8 / atol 1e-12: final states,
    aerosol water included, within the 1e-4 / 1e-8 bar."""
    mp, cells = case
    n = 8
    kw = dict(rel_tol=1e-8, abs_tol=np.full(mp.n_state, 1e-12), max_steps=100000)
    sol, orc = _pair(oracle_lib, mp, cells, n, **kw)
    a = cells.state[:n].copy()
    b = cells.state[:n].copy()
    for _ in range(2):
        assert sol.solve(a, cells.env[:n], 0.0, 180.0) == 0
        assert orc.run(b, cells.env[:n], 0.0, 180.0, mode=1) == 0
    ok = np.abs(a - b) <= 1e-4 * np.abs(b) + 1e-8
    worst = float(np.max(np.abs(a - b) / (1e-4 * np.abs(b) + 1e-8)))
    print("multi-phase solve worst (bar units)", worst, sol.stats()["num_steps"])
    assert ok.all(), worst


def test_multiphase_alone_equals_mix(case):
    mp, cells = case
    mix = CampSolver(mp, 8)
    mix.set_rates({k: v[:8] for k, v in cells.rates.items()})
    a = cells.state[:8].copy()
    assert mix.solve(a, cells.env[:8], 0.0, 180.0) == 0
    one = CampSolver(mp, 1)
    one.set_rates({k: v[5:6] for k, v in cells.rates.items()})
    s1 = cells.state[5:6].copy()
    assert one.solve(s1, cells.env[5:6], 0.0, 180.0) == 0
    assert np.array_equal(s1[0], a[5])


# --------------------------------------------------------------------------
# The HIP path against the closed forms of the reference's SIMPOL and
# aqueous-equilibrium unit tests (tests/mp_closed_forms.py)
# --------------------------------------------------------------------------
from tests import mp_closed_forms as CF  # noqa: E402


@pytest.mark.parametrize("kind", ["simpol", "aqeq"])
def test_multiphase_f_matches_reference_closed_forms_on_gpu(kind):
    mk, stf, ff = ((CF.simpol_mechanism, CF.simpol_state, CF.simpol_f) if kind == "simpol"
                   else (CF.aqeq_mechanism, CF.aqeq_state, CF.aqeq_f))
    mp = mk()
    n = 6
    rng = np.random.default_rng(3)
    env = np.array([[CF.T_TEST, CF.P_TEST]] + [[rng.uniform(250, 310), rng.uniform(5e4, 1.05e5)]
                                               for _ in range(n - 1)])
    st = np.stack([stf(mp)] + [stf(mp) * np.exp(rng.uniform(-1, 1, mp.n_state))
                               for _ in range(n - 1)])
    sol = CampSolver(mp, n)
    rc, f = sol.eval_deriv(st, env)
    assert rc == 0
    dep = [i for i in range(mp.n_state) if mp.var_type[i] == 1]
    for c in range(n):
        T, P = env[c]
        want = (ff(mp, st[c], T, P) if kind == "simpol" else ff(mp, st[c], T))[dep]
        scale = np.max(np.abs(want))
        assert np.all(np.abs(f[c] - want) <= 1e-12 * np.maximum(np.abs(want), 1e-3 * scale))


@pytest.mark.parametrize("kind", ["simpol", "aqeq"])
def test_multiphase_box_solve_reaches_closed_form_equilibrium_on_gpu(kind):
    """solver_run on the GPU (per-cell kernel) reaches the Raoult's-law
    (SIMPOL) and closed-form aqueous equilibria within 1e-6."""
    if kind == "simpol":
        mp, st0, eq, steps = CF.simpol_mechanism(), None, CF.simpol_equilibrium, 10
        st0 = CF.simpol_state(mp)
    else:
        mp, eq, steps = CF.aqeq_mechanism(), CF.aqeq_equilibrium, 20
        st0 = CF.aqeq_state(mp)
    sol = CampSolver(mp, 2, rel_tol=1e-10, abs_tol=np.full(mp.n_state, 1e-22),
                     max_steps=100000)
    st = np.stack([st0, st0])
    env = np.array([[CF.T_TEST, CF.P_TEST]] * 2)
    for _ in range(steps):
        assert sol.solve(st, env, 0.0, 3600.0) == 0
    want = eq(mp, st0)
    for c in range(2):
        assert np.all(np.abs(st[c] - want) <= 1e-6 * want)
