"""Pins the CPU oracle before it is trusted as the parity checker.

* the reference's own known-answer data (test_jacobian.c, the analytic
  reaction tests and their rate-constant closed forms);
* the reference's own compiled C (oracle/_ref: pattern registration, id
  update, rate constants, rate updates, time-derivative accumulation) --
  these run only where the reference tree is present;
* mathematics: J against finite differences of f, sparse LU against dense
  partial-pivoting LU, BDF convergence to a tight Radau solution.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from camp_amd.mechanisms import (box_mechanism, cb05_scale_mechanism, load_camp_json,
                                 load_camp_json_file, synthetic_cells, pack, TYPE_CODES)
from tests.analytic_cases import KATS, all_cases, reference_bar

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def ref_or_skip(O):
    if not O.ref_available():
        pytest.skip("reference sources not present (oracle/_ref not built)")


# ---------------------------------------------------------------- KATs ----
def test_pattern_builder_kat(O):
    k = KATS["unit_jacobian"]
    cp, rv = O.kat_pattern(k["n"], k["register"])
    assert cp.tolist() == k["col_ptrs"]
    assert rv.tolist() == k["row_ids"]
    for dep, ind, eid in k["element_ids"]:
        col = rv[cp[ind]:cp[ind + 1]].tolist()
        assert cp[ind] + col.index(dep) == eid
    # a column with more rows than the registration buffer
    b = k["big_column"]
    cp, rv = O.kat_pattern(b["n"], [(i, b["col"]) for i in range(b["n_rows"])])
    assert cp[:b["col"] + 1].tolist() == [0] * (b["col"] + 1)
    assert cp[b["col"] + 1:].tolist() == [b["n_rows"]] * (b["n"] - b["col"])
    assert rv.tolist() == list(range(b["n_rows"]))


def test_jacobian_accumulator_kat(O):
    k = KATS["unit_jacobian"]
    out = O.kat_jac_accum(4, k["adds"])
    assert out.tolist() == k["outputs"]
    adds = [(i, 0, 10.0 + i) for i in range(20)] + [(i, 1, 10.0 + 2.0 * i) for i in range(20)]
    out = O.kat_jac_accum(20, adds)
    assert out.tolist() == [-1.0 * i for i in range(20)]


def test_reference_unit_jacobian_program(O):
    exe = os.path.join(os.path.dirname(O.LIB_PATH), "_ref", "test_jacobian")
    if not os.path.exists(exe):
        pytest.skip("reference sources not present")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=30).stdout
    assert "PASS" in out


def test_troe_rate_constants_kat(O):
    from tests.analytic_cases import troe_case
    c = troe_case()
    o = O.Oracle(c["mech"], 1)
    o.update_env(np.array([[c["T"], c["P"]]]))
    k = o.rxn_env()[0]
    assert np.allclose(k, c["k"], rtol=1e-12, atol=0)


def test_arrhenius_rate_constants_kat(O):
    from tests.analytic_cases import arrhenius_cases
    for c in arrhenius_cases():
        o = O.Oracle(c["mech"], 1)
        o.update_env(np.array([[c["T"], c["P"]]]))
        k = o.rxn_env()[0]
        # k1 = D * k(A + D -> B) + k(A -> B); k2 = k(B -> C + D)
        k1 = c["y0"][1] * k[0] + k[1]
        assert abs(k1 - c["k"][0]) <= 1e-12 * c["k"][0]
        assert abs(k[2] - c["k"][1]) <= 1e-12 * c["k"][1]


@pytest.mark.parametrize("name,case", all_cases())
def test_oracle_analytic_solutions(O, name, case):
    """The reference's analytic tests, at their own pass rule and tighter."""
    mech = case["mech"]
    o = O.Oracle(mech, 1, rel_tol=1e-10, abs_tol=[1e-14] * mech.n_state)
    o.set_rates({k: [v] for k, v in case["rates"].items()})
    st = case["y0"][None, :].copy()
    env = np.array([[case["T"], case["P"]]])
    first = case["solution"](case["dt"])
    worst = 0.0
    for i in range(1, case["n_steps"] + 1):
        assert o.run(st, env, 0.0, case["dt"], mode=1) == 0
        true = case["solution"](i * case["dt"])
        assert reference_bar(st[0], true, first).all(), (i, st[0], true)
        big = np.abs(true) > 1e-5 * np.abs(first)
        if big.any():
            worst = max(worst, float(np.max(np.abs(st[0][big] - true[big]) / np.abs(true[big]))))
    assert worst < 1e-5, worst


# ------------------------------------------- the reference's own code ----
MECHS = [("box", box_mechanism(1)), ("cb05", cb05_scale_mechanism(1)),
         ("arrhenius_json", load_camp_json(json.load(open(os.path.join(HERE, "golden",
                                                                       "rxn_arrhenius.json"))))),
         ("cb05cl_ae5", load_camp_json_file(os.path.join(HERE, "golden", "cb05cl_ae5",
                                                         "config.json")))]


@pytest.mark.parametrize("name,mech", MECHS)
def test_pattern_and_ids_match_reference_code(O, name, mech):
    ref_or_skip(O)
    o = O.Oracle(mech, 2)
    r = O.RefModel(mech, 2)
    cp, rv = r.pattern()
    p = o.pattern()
    assert np.array_equal(cp, p["rxn_colptr"]) and np.array_equal(rv, p["rxn_rowval"])
    ids = r.update_ids(p["deriv_ids"])
    assert np.array_equal(ids, o.int_data())


@pytest.mark.parametrize("name", ["cb05", "cb05cl_ae5"])
def test_rate_constants_match_reference_code(O, name):
    """Rate constants and rate updates bit-equal to the reference's own
    compiled C, on the CB05-scale mechanism and on the real CB05cl_ae5
    (Arrhenius, Troe, CMAQ_H2O2, CMAQ_OH_HNO3, photolysis) over seeded
    temperatures and pressures."""
    ref_or_skip(O)
    mech = dict(MECHS)[name]
    n = 64
    cells = synthetic_cells(mech, n, seed=5)
    o = O.Oracle(mech, n)
    r = O.RefModel(mech, n)
    o.set_rates(cells.rates)
    types = {x.rxn_id: TYPE_CODES[x.type] for x in mech.reactions if x.rxn_id > 0}
    for uid, vals in cells.rates.items():
        for c in range(n):
            r.update_data(c, types[uid], uid, float(vals[c]))
    o.update_env(cells.env)
    for c in range(n):
        r.update_env(c, cells.env[c, 0], cells.env[c, 1])
    assert np.array_equal(o.rxn_env(), r.env_data())


def test_time_derivative_matches_reference_code(O):
    ref_or_skip(O)
    rng = np.random.default_rng(11)
    n_spec, n = 9, 200
    spec = rng.integers(0, n_spec, size=n)
    contrib = rng.normal(size=n) * 10.0 ** rng.uniform(-8, 4, size=n)
    contrib[::17] = 0.0
    est = rng.normal(size=n_spec)
    for e in (None, est):
        a = O.kat_time_derivative(n_spec, spec, contrib, e)
        b = O.ref_time_derivative(n_spec, spec, contrib, e)
        assert np.array_equal(a, b)


# -------------------------------------------------------- mathematics ----
def test_jacobian_matches_finite_differences(O):
    mech = cb05_scale_mechanism(1)
    cells = synthetic_cells(mech, 1, seed=3)
    o = O.Oracle(mech, 1)
    o.set_rates({k: v[:1] for k, v in cells.rates.items()})
    o.update_env(cells.env)
    y = cells.state[0]
    rc, J = o.eval_jac(0, y)
    assert rc == 0
    p = o.pattern()
    dense = np.zeros((o.nv, o.nv))
    for j in range(o.nv):
        for e in range(p["colptr"][j], p["colptr"][j + 1]):
            dense[p["rowval"][e], j] = J[e]
    # Each reaction has distinct reactants, so f is affine in every single
    # y_j and a central difference with a large step is exact up to roundoff
    # (a small step would drown in the cancellation of production and loss).
    for j in range(o.nv):
        h = 0.5 * y[j]
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        fp, fm = o.eval_f(0, yp)[1], o.eval_f(0, ym)[1]
        fd = (fp - fm) / (2 * h)
        err = np.abs(fd - dense[:, j])
        roundoff = 1e-12 * (np.abs(fp) + np.abs(fm)) / (2 * h)
        tol = 1e-6 * np.abs(dense[:, j]) + roundoff
        assert np.all(err <= tol), (j, np.argmax(err - tol), np.max(err - tol))
    # structure: every entry outside the pattern has a zero derivative
    assert np.count_nonzero(dense) <= o.nnz


def test_sparse_lu_matches_dense_partial_pivoting(O):
    mech = cb05_scale_mechanism(1)
    cells = synthetic_cells(mech, 4, seed=3)
    o = O.Oracle(mech, 4)
    o.set_rates(cells.rates)
    o.update_env(cells.env)
    rng = np.random.default_rng(1)
    for c in range(4):
        rc, J = o.eval_jac(c, cells.state[c])
        b = rng.normal(size=o.nv)
        for gamma in (1e-3, 1.0, 100.0):
            x1 = o.lu_solve(J, gamma, b)[1]
            x2 = o.lu_solve(J, gamma, b, dense=True)[1]
            assert np.linalg.norm(x1 - x2) <= 1e-9 * np.linalg.norm(x2)


def test_deriv_estimate_blend_without_history_is_exact_f(O):
    """With the J history at the evaluation state, the estimate equals f."""
    mech = cb05_scale_mechanism(1)
    cells = synthetic_cells(mech, 1, seed=4)
    o = O.Oracle(mech, 1)
    o.set_rates({k: v[:1] for k, v in cells.rates.items()})
    o.update_env(cells.env)
    o.eval_jac(0, cells.state[0])
    f0 = o.eval_f(0, cells.state[0], use_deriv_est=False)[1]
    f1 = o.eval_f(0, cells.state[0], use_deriv_est=True)[1]
    assert np.allclose(f0, f1, rtol=1e-13, atol=0)


def test_bdf_converges_to_radau_reference(O):
    """Golden: scipy Radau at rtol 1e-12 (tests/golden/make_golden.py)."""
    g = json.load(open(os.path.join(HERE, "golden", "box10x15_radau.json")))
    mech = box_mechanism(1)
    cells = synthetic_cells(mech, 1, seed=2, T=298.0, P=101325.0)
    assert np.allclose(cells.state[0], g["y0"], rtol=0, atol=0)
    errs = []
    for rtol in (1e-6, 1e-10):
        o = O.Oracle(mech, 1, rel_tol=rtol, abs_tol=[rtol * 1e-4] * 10)
        st = cells.state.copy()
        for _ in range(g["n_steps"]):
            assert o.run(st, cells.env, 0.0, g["dt"], mode=1) == 0
        ref = np.array(g["y_final"])
        big = ref > 1e-6
        errs.append(float(np.max(np.abs(st[0][big] - ref[big]) / ref[big])))
    assert errs[1] < 1e-6 and errs[1] < errs[0], errs


def test_block_and_per_cell_modes_agree_within_tolerance(O):
    """unit_test_driver.F90:262-285: one multi-cell core vs single cells."""
    mech = box_mechanism(1)
    cells = synthetic_cells(mech, 6, seed=2)
    o = O.Oracle(mech, 6, rel_tol=1e-8, abs_tol=[1e-12] * 10)
    a, b = cells.state.copy(), cells.state.copy()
    for i in range(5):
        assert o.run(a, cells.env, 0.0, 60.0, mode=0) == 0
        assert o.run(b, cells.env, 0.0, 60.0, mode=1) == 0
        tol_r, tol_a = 1e-8 * (i + 1) * 100, 1e-12 * (i + 1) * 100
        assert np.all(np.abs(a - b) <= tol_r * np.abs(b) + tol_a)


def test_deriv_estimate_is_exact_for_an_affine_mechanism(O):
    """camp_solver.c:877-879 / time_derivative.c:58-64: the estimate is
    J_deriv + J_solver (y - J_state), the first-order expansion around the
    state of the last Jac().  For a mechanism of first- and zero-order
    reactions f is affine, so away from J_state the estimate must equal
    f(y) to rounding.  A transposed or sign-flipped matvec is far off (the
    test checks that it would be)."""
    from tests.helpers import affine_mechanism, dense_from_csc
    mech, rates = affine_mechanism()
    o = O.Oracle(mech, 1)
    o.set_rates(rates)
    env = np.array([[285.0, 9.0e4]])
    o.update_env(env)
    rng = np.random.default_rng(8)
    y0 = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), size=mech.n_state))
    rc, jv = o.eval_jac(0, y0)
    assert rc == 0
    pat = o.pattern()
    J = dense_from_csc(pat["colptr"], pat["rowval"], jv, o.nv)
    f0 = o.eval_f(0, y0)[1]
    assert np.linalg.norm(J - J.T) > 1e-2 * np.linalg.norm(J)
    for trial in range(3):
        y1 = y0 * np.exp(rng.uniform(-1.0, 1.0, size=mech.n_state))
        est = o.eval_est(0, y1)
        f1 = o.eval_f(0, y1)[1]
        scale = np.abs(f0) + np.abs(J) @ np.abs(y1 - y0) + np.abs(f1)
        assert np.all(np.abs(est - f1) <= 1e-13 * scale), np.max(np.abs(est - f1) / scale)
        d = y1 - y0
        for wrong in (f0 + J.T @ d, f0 - J @ d, J @ d):
            assert np.max(np.abs(wrong - f1) / scale) > 1e-3


def test_deriv_blend_returns_the_estimate_where_production_and_loss_cancel(O):
    """time_derivative.c:58-64: a species whose production and loss cancel
    exactly gets weight 0 on its (zero) sum and 1 on the estimate, while the
    others keep weight ~1 on their own sums.  The emission of Z changes
    between the Jac() that stores the history and the f() call, so Z's
    estimate (old emission + k_Y [Y] - k_Z [Z], through the one-way Y -> Z
    coupling) is not zero; the transposed coupling would give another
    value."""
    from tests.helpers import affine_cancel_state, affine_mechanism
    mech, rates = affine_mechanism()
    o = O.Oracle(mech, 1)
    o.set_rates(rates)
    env = np.array([[285.0, 9.0e4]])
    o.update_env(env)
    rng = np.random.default_rng(9)
    y0 = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), size=mech.n_state))
    o.eval_jac(0, y0)
    o.set_rates({5: np.array([2.0 ** -21])})
    o.update_env(env)
    y1 = affine_cancel_state(y0 * np.exp(rng.uniform(-0.5, 0.5, size=mech.n_state)))
    pure = o.eval_f(0, y1)[1]
    blend = o.eval_f(0, y1, use_deriv_est=True)[1]
    est = o.eval_est(0, y1)
    assert np.all(pure[[-4, -3, -1]] == 0.0)
    assert np.array_equal(blend[[-4, -3, -1]], est[[-4, -3, -1]])
    want = 3e-6 + 2.0 ** -12 * 2.0 ** -8 - 2.0 ** -10 * (3 * 2.0 ** -11)
    assert abs(est[-1] - want) <= 1e-12 * 3e-6
    assert abs(est[-1] - (want - 2.0 ** -12 * (y1[-2] - y0[-2]))) > 1e-3 * abs(want)
    others = np.ones(mech.n_state, bool)
    others[[-4, -3, -1]] = False
    assert np.allclose(blend[others], pure[others], rtol=1e-12, atol=0)


def test_cmaq_rate_constants_closed_form(O):
    """The real CB05cl_ae5's CMAQ_H2O2 (R34: k = k1 + k2 [M]) and CMAQ_OH_HNO3
    (R29: k = k0 + k3 [M] / (1 + k3 [M] / k2)) rate constants against the
    closed forms of the CMAQ rate expressions CMAQ_9 and CMAQ_8 the
    mechanism's "orig params" give (rate constants in cm3 s-1 times [M] in
    molecules cm-3), converted to ppm-1 s-1, at the reference test's
    conditions (test_cb05cl_ae5.F90:198-206: T 272.5 K, P 0.8 atm)."""
    import math
    from camp_amd.mechanisms import AVOGADRO, UNIV_GAS_CONST
    mech = dict(MECHS)["cb05cl_ae5"]
    T = 272.5
    P = float(np.float32(0.8)) * 101325.0
    o = O.Oracle(mech, 1)
    o.update_env(np.array([[T, P]]))
    k = o.rxn_env()[0]
    M = AVOGADRO / UNIV_GAS_CONST * P / T * 1e-6    # air molecules per cm3
    ppm = M * 1e-6                                  # molecules per cm3 per ppm
    # each reaction's first env slot (photolysis reactions hold two)
    offs = np.concatenate([[0], np.cumsum([p[3] for p in pack(mech)])])
    idx = {r.params.get("rxn id"): offs[i] for i, r in enumerate(mech.reactions)}
    # R34: HO2 + HO2 -> H2O2, CMAQ_9(2.3E-13, -600, 1.7E-33, -1000)
    k34 = (2.3e-13 * math.exp(600.0 / T) + 1.7e-33 * math.exp(1000.0 / T) * M) * ppm
    # R29: OH + HNO3 -> NO3, CMAQ_8(2.4E-14, -460, 2.7E-17, -2199, 6.5E-34, -1335)
    k0, k2, k3 = (2.4e-14 * math.exp(460.0 / T), 2.7e-17 * math.exp(2199.0 / T),
                  6.5e-34 * math.exp(1335.0 / T) * M)
    k29 = (k0 + k3 / (1.0 + k3 / k2)) * ppm
    for rid, want in (("R34", k34), ("R29", k29)):
        got = k[idx[rid]]
        assert abs(got - want) <= 1e-12 * want, (rid, got, want)
