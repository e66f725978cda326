"""The multi-phase part of the oracle (aerosol phase, modal/binned-mass
representation, SIMPOL phase transfer, aqueous equilibrium, ZSR aerosol
water) pinned against the reference's own C sources compiled into
oracle/_ref (pattern registration, id updates, environment updates, the
representation's state update and getters with their partial derivatives,
the ZSR water calculation), and its Jacobian against finite differences of
f (the f/J contributions of these types are compiled only with SUNDIALS and
cannot be built here)."""
import numpy as np
import pytest

from camp_amd.multiphase import multiphase_cells, multiphase_mechanism

N_CELLS = 6


@pytest.fixture(scope="module")
def mp():
    return multiphase_mechanism(1, n_bins=3, n_simpol=5, n_aqeq=4)


@pytest.fixture(scope="module")
def cells(mp):
    return multiphase_cells(mp, N_CELLS, seed=2)


@pytest.fixture(scope="module")
def pair(oracle_lib, mp):
    if not oracle_lib.ref_available():
        pytest.skip("reference sources not built (oracle/_ref)")
    orc = oracle_lib.Oracle(mp, N_CELLS)
    ref = oracle_lib.RefModel(mp, N_CELLS)
    return orc, ref


def test_rxn_and_param_patterns_and_ids_match_reference(pair, mp):
    orc, ref = pair
    pat = orc.pattern()
    cp, rv = ref.pattern()
    assert np.array_equal(pat["rxn_colptr"], cp) and np.array_equal(pat["rxn_rowval"], rv)
    ids = ref.update_ids(pat["deriv_ids"])
    assert np.array_equal(orc.int_data(), ids)
    # rxn float data: SMALL_CONC / SMALL_WATER of the aqueous equilibria
    assert np.array_equal(orc.arrays(0, orc.n_float), ref.arrays(0, ref.n_float))
    # representation int data: PHASE_NUM_JAC_ELEM written by SIMPOL's registration
    n_rep_int = orc.sizes[4] + orc.sizes[3]
    assert np.array_equal(orc.arrays(1, n_rep_int), ref.arrays(1, n_rep_int))
    pcp, prv = orc.param_pattern()
    rcp, rrv = ref.param_pattern(pat["deriv_ids"])
    assert np.array_equal(pcp, rcp) and np.array_equal(prv, rrv)
    n_sub_int = orc.sizes[8] + orc.sizes[7]
    assert np.array_equal(orc.arrays(3, n_sub_int), ref.arrays(3, n_sub_int))
    # the solver pattern chains through the sub-model Jacobian: aerosol water
    # is CONSTANT, so reactions that read it gain columns of the ions and of
    # nothing else
    assert orc.nnz > 0 and len(prv) > 1


def test_environment_updates_match_reference(pair, mp, cells):
    orc, ref = pair
    orc.update_env(cells.env)
    for c in range(N_CELLS):
        ref.update_env_all(c, *cells.env[c])
    assert np.array_equal(orc.rxn_env(), ref.env_data())
    n_sub_env = orc.sizes[10] * N_CELLS
    assert np.array_equal(orc.arrays(4, n_sub_env), ref.arrays(4, n_sub_env))


def test_cell_update_and_representation_getters_match_reference(pair, mp, cells):
    orc, ref = pair
    orc.update_env(cells.env)
    for c in range(N_CELLS):
        ref.update_env_all(c, *cells.env[c])
    n_rep_flt = orc.sizes[5]
    nje = sum(1 for v in mp.rep.phases[0].var_type if v in (1, 2, 3))
    for c in range(N_CELLS):
        a = orc.cell_update(c, cells.state[c])
        b = ref.cell_update(c, cells.state[c])
        assert np.array_equal(a, b)                 # ZSR water
        assert not np.array_equal(a, cells.state[c])
        assert np.array_equal(orc.arrays(2, n_rep_flt), ref.arrays(2, n_rep_flt))
        for phase in range(mp.rep.n_bins):
            for which in range(4):
                va, pa = orc.rep_get(0, c, a, phase, which, nje)
                vb, pb = ref.rep_get(0, c, b, phase, which, nje)
                assert va == vb and np.array_equal(pa, pb), (c, phase, which)


def _fd_jacobian(orc, c, st, nv, dep, h_rel=1e-3):
    """Central differences of f over the solver variables of one cell.

    f() updates the aerosol representation before the ZSR sub model
    recomputes water (camp_solver.c:913-917), so within one call the
    representation sees the water already in the state.  The Jacobian
    assumes water follows the ions (the J_rxn * J_params chain), so each
    perturbed state first gets its ZSR water, then f is evaluated.  f sums
    large production and loss terms, so small steps are roundoff-bound:
    steps of 1e-3 and 5e-4 (relative) are combined by Richardson
    extrapolation, leaving an O(h^4) truncation error."""
    def central(h_r):
        D = np.zeros((nv, nv))
        for j in range(nv):
            i = dep[j]
            h = max(abs(st[i]) * h_r, 1e-12)
            sp, sm = st.copy(), st.copy()
            sp[i] += h
            sm[i] -= h
            _, fp = orc.eval_f(c, orc.cell_update(c, sp), 1.0)
            _, fm = orc.eval_f(c, orc.cell_update(c, sm), 1.0)
            D[:, j] = (fp - fm) / (2 * h)
        return D
    return (4.0 * central(h_rel / 2) - central(h_rel)) / 3.0


def test_multiphase_jacobian_matches_finite_differences(oracle_lib, mp, cells):
    """The solver Jacobian (reaction Jacobian chained through the ZSR
    parameter Jacobian) against central differences of f, relative 1e-4 /
    absolute 1e-9 as CAMP's own check_Jac (camp_solver.c:1146-1257)."""
    orc = oracle_lib.Oracle(mp, N_CELLS)
    orc.update_env(cells.env)
    pat = orc.pattern()
    dep = [i for i in range(mp.n_state) if mp.var_type[i] == 1]
    nv = len(dep)
    cp, rv = pat["colptr"], pat["rowval"]
    for c in range(2):
        st = orc.cell_update(c, cells.state[c])
        rc, jv = orc.eval_jac(c, st, 1.0)
        assert rc == 0
        J = np.zeros((nv, nv))
        for j in range(nv):
            for k in range(cp[j], cp[j + 1]):
                J[rv[k], j] = jv[k]
        F = _fd_jacobian(orc, c, st, nv, dep)
        # FD of the non-smooth ZSR branches (low-RH clamp) is excluded by
        # the choice of RH; everything else is smooth in y
        err = np.abs(J - F)
        tol = 1e-4 * np.maximum(np.abs(J), np.abs(F)) + 1e-9 * np.max(np.abs(F))
        assert (err <= tol).all(), np.max(err / np.maximum(tol, 1e-300))
        # no FD-visible dependency outside the pattern
        mask = np.zeros((nv, nv), bool)
        for j in range(nv):
            mask[rv[cp[j]:cp[j + 1]], j] = True
        assert np.all(np.abs(F[~mask]) <= 1e-9 * np.max(np.abs(F)))


def test_multiphase_box_solve_conserves_and_converges(oracle_lib, mp, cells):
    """Per-cell BDF on the multi-phase mechanism: solves succeed, every
    state stays non-negative, and the result converges as the tolerance
    tightens (rtol 1e-6 vs 1e-9 within the 1e-4 / 1e-8 bar)."""
    out = []
    for rtol, atol in ((1e-6, 1e-10), (1e-9, 1e-13)):
        orc = oracle_lib.Oracle(mp, 2, rel_tol=rtol, abs_tol=np.full(mp.n_state, atol),
                                max_steps=100000)
        orc.set_rates({k: v[:2] for k, v in cells.rates.items()})
        st = cells.state[:2].copy()
        for _ in range(2):
            assert orc.run(st, cells.env[:2], 0.0, 180.0, mode=1) == 0
        assert (st >= 0).all()
        out.append(st)
    ok = np.abs(out[0] - out[1]) <= 1e-4 * np.abs(out[1]) + 1e-8
    assert ok.all(), np.max(np.abs(out[0] - out[1]) / (np.abs(out[1]) + 1e-8))


# --------------------------------------------------------------------------
# Closed forms of the reference's SIMPOL and aqueous-equilibrium unit tests
# (test_rxn_SIMPOL_phase_transfer.F90:330-410, test_rxn_aqueous_
# equilibrium.F90:300-372) on BINNED mechanisms: see tests/mp_closed_forms.py
# --------------------------------------------------------------------------
from tests import mp_closed_forms as CF  # noqa: E402


def _dep(mp):
    return [i for i in range(mp.n_state) if mp.var_type[i] == 1]


@pytest.mark.parametrize("kind", ["simpol", "aqeq"])
def test_multiphase_f_matches_reference_closed_forms(oracle_lib, kind):
    """f at the initial state and at perturbed states and environments
    against the closed-form condensation/evaporation (SIMPOL) and forward/
    reverse molality rates (aqueous equilibrium), to 1e-12 of the largest
    term of each derivative."""
    mk, stf, ff = ((CF.simpol_mechanism, CF.simpol_state, CF.simpol_f) if kind == "simpol"
                   else (CF.aqeq_mechanism, CF.aqeq_state, CF.aqeq_f))
    mp = mk()
    rng = np.random.default_rng(3)
    st0 = stf(mp)
    for trial in range(6):
        T = CF.T_TEST if trial == 0 else rng.uniform(250.0, 310.0)
        P = CF.P_TEST if trial == 0 else rng.uniform(5e4, 1.05e5)
        st = st0 if trial == 0 else st0 * np.exp(rng.uniform(-1, 1, st0.size))
        o = oracle_lib.Oracle(mp, 1)
        o.update_env(np.array([[T, P]]))
        rc, f = o.eval_f(0, st)
        assert rc == 0
        want = (ff(mp, st, T, P) if kind == "simpol" else ff(mp, st, T))[_dep(mp)]
        scale = np.max(np.abs(want))
        assert np.all(np.abs(f - want) <= 1e-12 * np.maximum(np.abs(want), 1e-3 * scale)), \
            np.max(np.abs(f - want) / np.maximum(np.abs(want), 1e-300))


def test_simpol_box_solve_reaches_raoult_equilibrium(oracle_lib):
    """Ten hours of ethanol partitioning into two bins of aqueous aerosol end
    at Raoult's-law equilibrium in both bins with the total mass conserved
    (the equilibrium the reference test's fixture value 2.15877e-7 kg m-3
    iterates to for its MODAL/UNIFAC setup), within 1e-6."""
    mp = CF.simpol_mechanism()
    st = CF.simpol_state(mp)[None, :].copy()
    o = oracle_lib.Oracle(mp, 1, rel_tol=1e-10, abs_tol=np.full(mp.n_state, 1e-22), max_steps=100000)
    env = np.array([[CF.T_TEST, CF.P_TEST]])
    for _ in range(10):
        assert o.run(st, env, 0.0, 3600.0, mode=1) == 0
    want = CF.simpol_equilibrium(mp, CF.simpol_state(mp))
    assert np.all(np.abs(st[0] - want) <= 1e-6 * want)
    # the closed form really moved: gas depleted, both bins grew
    assert want[0] < 0.9 * CF.simpol_state(mp)[0]


def test_aqueous_equilibria_reach_closed_forms(oracle_lib):
    """2A <-> B + C, D + E <-> 2F and G <-> H in two bins end at their
    closed-form equilibria (test_rxn_aqueous_equilibrium.F90:327-348)
    within 1e-6; the linear G <-> H follows its exact exponential transient
    (:426-429) at t = 1000 s within 1e-8."""
    mp = CF.aqeq_mechanism()
    env = np.array([[CF.T_TEST, CF.P_TEST]])
    st = CF.aqeq_state(mp)[None, :].copy()
    o = oracle_lib.Oracle(mp, 1, rel_tol=1e-10, abs_tol=np.full(mp.n_state, 1e-22), max_steps=100000)
    for _ in range(20):
        assert o.run(st, env, 0.0, 3600.0, mode=1) == 0
    want = CF.aqeq_equilibrium(mp, CF.aqeq_state(mp))
    assert np.all(np.abs(st[0] - want) <= 1e-6 * want)
    st2 = CF.aqeq_state(mp)[None, :].copy()
    o2 = oracle_lib.Oracle(mp, 1, rel_tol=1e-10, abs_tol=np.full(mp.n_state, 1e-22), max_steps=100000)
    assert o2.run(st2, env, 0.0, 1000.0, mode=1) == 0
    for i, v in CF.relax_GH(mp, CF.aqeq_state(mp), 1000.0).items():
        assert abs(st2[0][i] - v) <= 1e-8 * v
