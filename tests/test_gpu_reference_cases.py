"""The reference's own test cases on the HIP path, through the C ABI.

* The analytic-solution reaction tests (test_rxn_troe.F90,
  test_rxn_photolysis.F90, test_rxn_emission.F90,
  test_rxn_first_order_loss.F90, unit_test_rxn_arrhenius.F90 x 5 initial
  states): closed forms checked every step at the reference's own pass rule
  (almost_equal at 1e-2, util.F90:379-415) and at 1e-5 with tight solver
  tolerances.  Photolysis, emission and loss rates go in per cell through
  rxn_*_create_rate_update_data / rxn_*_set_rate_update_data /
  rxn_update_data (rxn_solver.c:708-763), the way camp_core%update_data
  does.
* solver_run's failure return (camp_solver.c:636-668): max_steps exceeded
  gives CAMP_SOLVER_FAIL with the integrator flag CV_TOO_MUCH_WORK, as the
  oracle does.
* guess_helper (camp_solver.c:1321-1408) firing on negative predictions.
* solver_get_statistics (camp_solver.c:725-791) against the oracle's
  counters, and the fields a build without CAMP_DEBUG reports as constants.
"""
import numpy as np
import pytest

from camp_amd.core import CampSolver, MODE_BLOCK, MODE_PER_CELL
from camp_amd.mechanisms import (Mechanism, Reaction, RXN_FIRST_ORDER_LOSS,
                                 cb05_scale_mechanism, synthetic_cells)
from tests.analytic_cases import all_cases, reference_bar
from tests.helpers import cb05_inputs

pytestmark = pytest.mark.gpu

CASES = all_cases()


def _per_cell_rates(sol, rates, n_cells):
    """camp_core%update_data for every cell: one payload per call."""
    for uid, v in rates.items():
        for c in range(n_cells):
            sol.update_rxn_data(c, uid, v)


@pytest.mark.parametrize("mode", [MODE_PER_CELL, MODE_BLOCK])
@pytest.mark.parametrize("name,case", CASES, ids=[c[0] for c in CASES])
def test_analytic_solution_on_gpu(name, case, mode):
    """Three copies of the reference test's box, integrated step by step:
    every value within the reference's pass rule, and within 1e-5 of the
    closed form at rtol 1e-10."""
    mech = case["mech"]
    n = 3
    sol = CampSolver(mech, n, rel_tol=1e-10, abs_tol=[1e-14] * mech.n_state, mode=mode)
    _per_cell_rates(sol, case["rates"], n)
    st = np.tile(case["y0"], (n, 1))
    env = np.tile([case["T"], case["P"]], (n, 1))
    first = case["solution"](case["dt"])
    worst = 0.0
    for i in range(1, case["n_steps"] + 1):
        assert sol.solve(st, env, 0.0, case["dt"]) == 0, i
        true = case["solution"](i * case["dt"])
        for c in range(n):
            assert reference_bar(st[c], true, first).all(), (i, c, st[c], true)
        big = np.abs(true) > 1e-5 * np.abs(first)
        if big.any():
            worst = max(worst, float(np.max(np.abs(st[:, big] - true[big]) / np.abs(true[big]))))
    assert np.array_equal(st[0], st[1]) and np.array_equal(st[0], st[2])
    assert worst < 1e-5, worst


def test_rxn_update_data_matches_bulk_rates_and_oracle(oracle_lib):
    """Per-cell rxn_update_data calls (one payload per cell and reaction,
    the reference's update path) leave the same rate data as the bulk
    update, and as the oracle's rxn_update_data, on the CB05-scale
    mechanism with night cells (J = 0)."""
    n = 24
    mech, cells = cb05_inputs(n, night_fraction=0.5)
    a = CampSolver(mech, n)
    for uid, vals in cells.rates.items():
        vals = np.broadcast_to(np.asarray(vals, float), (n,))
        for c in range(n):
            a.update_rxn_data(c, uid, vals[c])
    b = CampSolver(mech, n)
    b.set_rates(cells.rates)
    orc = oracle_lib.Oracle(mech, n)
    orc.set_rates(cells.rates)
    assert a.eval_deriv(cells.state, cells.env)[0] == 0
    assert b.eval_deriv(cells.state, cells.env)[0] == 0
    orc.update_env(cells.env)
    ra, rb, ro = a.rxn_env(), b.rxn_env(), orc.rxn_env()
    assert np.array_equal(ra, rb)
    nz = ro != 0
    assert np.array_equal(ra == 0, ro == 0)
    assert np.max(np.abs(ra[nz] - ro[nz]) / np.abs(ro[nz])) <= 1e-12
    # and the solves that follow are the same
    sa, sb = cells.state.copy(), cells.state.copy()
    assert a.solve(sa, cells.env, 0.0, 180.0) == 0
    assert b.solve(sb, cells.env, 0.0, 180.0) == 0
    assert np.array_equal(sa, sb)


@pytest.mark.parametrize("mode", [MODE_PER_CELL, MODE_BLOCK])
def test_max_steps_exceeded_fails_like_oracle(oracle_lib, mode):
    """solver_run returns CAMP_SOLVER_FAIL (1) and solver_get_statistics
    reports CV_TOO_MUCH_WORK (-1) when max_steps is exceeded
    (camp_solver.c:434, 636-668); a call with room to finish succeeds."""
    n = 4
    mech, cells = cb05_inputs(n)
    for max_steps, want in ((3, 1), (100000, 0)):
        sol = CampSolver(mech, n, max_steps=max_steps, mode=mode)
        orc = oracle_lib.Oracle(mech, n, max_steps=max_steps)
        sol.set_rates(cells.rates)
        orc.set_rates(cells.rates)
        a, b = cells.state.copy(), cells.state.copy()
        assert sol.solve(a, cells.env, 0.0, 180.0) == want
        assert orc.run(b, cells.env, 0.0, 180.0, mode=0 if mode == MODE_BLOCK else 1) == want
        s, o = sol.stats(), orc.stats()
        if want:
            assert s["solver_flag"] == -1 and o["solver_flag"] == -1
            assert s["num_steps"] == o["num_steps"] == (3 if mode == MODE_BLOCK else 3 * n)
        else:
            assert s["solver_flag"] == 0
        sol.close()


def _loss_mechanism(k=0.954):
    """A -> (first-order loss), B -> C (Arrhenius, k2 = 0.1): a 10 s step
    with h0 = dt predicts A < 0, so guess_helper runs."""
    sp = ["A", "B", "C"]
    rx = [Reaction("FIRST_ORDER_LOSS", [0], [], {"scaling factor": 1.0}, rxn_id=1),
          Reaction("ARRHENIUS", [1], [(2, 1.0)], {"A": 0.1})]
    return Mechanism(sp, [1, 1, 1], [1e-14] * 3, rx, rel_tol=1e-8, name="loss"), k


@pytest.mark.parametrize("mode", [MODE_PER_CELL, MODE_BLOCK])
def test_guess_helper_fires_and_matches_oracle(oracle_lib, mode):
    """With h0 = dt the first prediction of a fast loss is negative and
    guess_helper takes its explicit sub-steps (their f() calls count as RHS
    evaluations, as in the CVODE fork).  The GPU and the oracle agree on the
    counters and on the result; switching the helper off changes the
    counters, so it did fire."""
    mech, k = _loss_mechanism()
    n = 2
    y0 = np.array([[1.0, 1.0, 0.0]] * n)
    env = np.array([[298.0, 101325.0]] * n)
    res = {}
    for on in (1, 0):
        sol = CampSolver(mech, n, rel_tol=1e-8, abs_tol=[1e-14] * 3, mode=mode)
        orc = oracle_lib.Oracle(mech, n, rel_tol=1e-8, abs_tol=[1e-14] * 3)
        sol.set_guess_helper(on)
        orc.set_guess_helper(on)
        for c in range(n):
            sol.update_rxn_data(c, 1, k)
        orc.set_rates({1: [k] * n})
        a, b = y0.copy(), y0.copy()
        assert sol.solve(a, env, 0.0, 10.0) == 0
        assert orc.run(b, env, 0.0, 10.0, mode=0 if mode == MODE_BLOCK else 1) == 0
        s, o = sol.stats(), orc.stats()
        for key in ("num_steps", "RHS_evals", "LS_setups", "error_test_fails", "NLS_iters",
                    "NLS_convergence_fails", "DLS_Jac_evals"):
            assert s[key] == o[key], (on, key, s[key], o[key])
        assert np.allclose(a, b, rtol=1e-4, atol=1e-8)
        assert abs(a[0, 0] - np.exp(-k * 10.0)) <= 1e-5 * np.exp(-k * 10.0)
        res[on] = s
        sol.close()
    assert res[1]["RHS_evals"] != res[0]["RHS_evals"]


@pytest.mark.parametrize("name,case", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_statistics_match_oracle(oracle_lib, name, case):
    """solver_get_statistics after one solver_run of one cell: the CVODE
    counters and step sizes equal the oracle's (the same algorithm on the
    same problem), and the fields a build without CAMP_DEBUG fills with
    constants (camp_solver.c:776-789) have those values."""
    mech = case["mech"]
    sol = CampSolver(mech, 1, rel_tol=1e-6, abs_tol=[1e-12] * mech.n_state)
    orc = oracle_lib.Oracle(mech, 1, rel_tol=1e-6, abs_tol=[1e-12] * mech.n_state)
    _per_cell_rates(sol, case["rates"], 1)
    orc.set_rates({k: [v] for k, v in case["rates"].items()})
    a = case["y0"][None, :].copy()
    b = a.copy()
    env = np.array([[case["T"], case["P"]]])
    assert sol.solve(a, env, 0.0, 10.0 * case["dt"]) == 0
    assert orc.run(b, env, 0.0, 10.0 * case["dt"], mode=1) == 0
    s, o = sol.full_stats(), orc.stats()
    for key in ("solver_flag", "num_steps", "RHS_evals", "LS_setups", "error_test_fails",
                "NLS_iters", "NLS_convergence_fails", "DLS_Jac_evals"):
        assert s[key] == o[key], (key, s[key], o[key])
    # step sizes come from error norms, in which rounding differences of the
    # order of 1e-16 show; the proposed next step of the emission case (f
    # constant, so its error norm is at rounding level) moves by ~1e-4
    assert abs(s["last_time_step__s"] - o["last_time_step__s"]) <= \
        1e-6 * abs(o["last_time_step__s"])
    assert abs(s["next_time_step__s"] - o["next_time_step__s"]) <= \
        1e-3 * abs(o["next_time_step__s"])
    assert s["DLS_RHS_evals"] == 0 and s["Jac_eval_fails"] == 0
    assert s["RHS_evals_total"] == -1 and s["Jac_evals_total"] == -1
    assert s["max_loss_precision"] == 0.0


def test_jacobian_check_passes_the_analytic_jacobian_and_flags_wrong_elements():
    """The device finite-difference check (check_Jac's role and pass rule,
    camp_solver.c:1146-1257): J(y) of the CB05-scale mechanism passes at
    random states; a Jacobian with one element off by 1% or one column
    zeroed fails in that cell only; with solver_set_eval_jac the solver's
    own last Jacobians pass (Jac_eval_fails == 0, the reference tests'
    assertion)."""
    n = 16
    mech, cells = cb05_inputs(n, night_fraction=0.5)
    sol = CampSolver(mech, n)
    sol.set_rates(cells.rates)
    rc, fails = sol.check_jac(cells.state, cells.env, 10.0)
    assert rc == 0 and not fails.any(), fails
    rc, J = sol.eval_jac(cells.state, cells.env, 10.0)
    assert rc == 0
    p = sol.pattern()
    bad = J.copy()
    c0, c1 = p["colptr"][7], p["colptr"][8]
    bad[3, c0 + int(np.argmax(np.abs(J[3, c0:c1])))] *= 1.01   # the column's largest
    c0, c1 = p["colptr"][20], p["colptr"][21]
    bad[9, c0:c1] = 0.0
    rc, fails = sol.check_jac(cells.state, cells.env, 10.0, jac=bad)
    assert fails[3] >= 1 and fails[9] >= 1, fails
    assert np.count_nonzero(fails) == 2 and rc == int(fails.sum())
    sol.set_eval_jac(True)
    st = cells.state.copy()
    assert sol.solve(st, cells.env, 0.0, 180.0) == 0
    assert sol.stats()["Jac_eval_fails"] == 0
