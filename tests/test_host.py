"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared symbol, solver_initialize builds CAMP's pattern and ids
bit-exactly on the host, the packer follows the reference's Fortran
packers, and setup errors follow the reference's convention."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from camp_amd.mechanisms import (CONV, box_mechanism, cb05_scale_mechanism, load_camp_json,
                                 load_camp_json_file, pack, synthetic_cells, RXN_TROE)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text))
    return sorted(names - {"void", "int", "double", "long"})


def test_library_exports_every_declared_symbol():
    from camp_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert "solver_run" in names and "rxn_add_condensed_data" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


CB05CL_AE5 = os.path.join(HERE, "golden", "cb05cl_ae5", "config.json")


@pytest.mark.parametrize("mech", [box_mechanism(1), cb05_scale_mechanism(1),
                                  load_camp_json(json.load(open(os.path.join(
                                      HERE, "golden", "rxn_arrhenius.json")))),
                                  load_camp_json_file(CB05CL_AE5)],
                         ids=["box", "cb05", "arrhenius_json", "cb05cl_ae5"])
def test_product_pattern_and_ids_match_oracle(oracle_lib, mech):
    from camp_amd.core import CampSolver
    s = CampSolver(mech, 3)          # solver_initialize runs on the host
    o = oracle_lib.Oracle(mech, 3)
    ps, po = s.pattern(), o.pattern()
    for k in ("rxn_colptr", "rxn_rowval", "colptr", "rowval", "deriv_ids"):
        assert np.array_equal(ps[k], po[k]), k
    assert np.array_equal(ps["rxn_int_data"], o.int_data())
    # invariants of the CAMP pattern: sorted rows per column, full diagonal
    cp, rv = ps["colptr"], ps["rowval"]
    for j in range(s.nv):
        col = rv[cp[j]:cp[j + 1]]
        assert np.all(np.diff(col) > 0) and j in col


def test_cb05_scale_dimensions():
    from camp_amd.core import CampSolver
    s = CampSolver(cb05_scale_mechanism(1), 2)
    assert (s.nv, s.ns) == (72, 72)
    assert 600 < s.nnz < 900
    assert s.lu_nnz >= s.nnz
    types = [r.type for r in cb05_scale_mechanism(1).reactions]
    assert {t: types.count(t) for t in set(types)} == {
        "ARRHENIUS": 133, "TROE": 28, "PHOTOLYSIS": 19, "EMISSION": 5, "FIRST_ORDER_LOSS": 5}


def test_troe_packing_follows_fortran_packer():
    """rxn_troe.F90:196-245: k0_A x 1e6, time unit MIN -> SCALING 1/60, CONV."""
    mech = load_camp_json(json.load(open(os.path.join(HERE, "golden", "test_troe.json"))))
    (t0, i0, f0, e0), (t1, i1, f1, e1) = pack(mech)
    assert t0 == RXN_TROE and e0 == 1
    assert i0 == [1, 1, 1, 2, 0, 0, 0, 0]
    assert f0[0] == 4.0e-18 * 1.0e6 and f0[6] == 0.6 and f0[7] == 1.0 and f0[8] == 1.0
    assert f0[9] == CONV
    assert f1[:9] == [1.2e-12 * 1.0e6, 167, 3, 136, 5, 24, 0.9, 0.8, 1.0 / 60.0]


def test_seeded_inputs_are_reproducible_and_batch_independent():
    mech = cb05_scale_mechanism(1)
    a = synthetic_cells(mech, 50, seed=2, night_fraction=0.5, stressed_fraction=0.1)
    b = synthetic_cells(mech, 10, seed=2, cell_offset=20, night_fraction=0.5,
                        stressed_fraction=0.1)
    assert np.array_equal(a.state[20:30], b.state)
    assert np.array_equal(a.env[20:30], b.env)
    for k in a.rates:
        assert np.array_equal(a.rates[k][20:30], b.rates[k])
    assert cb05_scale_mechanism(1).reactions[5].params == cb05_scale_mechanism(1).reactions[5].params


def test_setup_error_exits_like_the_reference():
    """camp_solver.c:93-96: setup errors print and exit(EXIT_FAILURE); here a
    reaction type outside the supported set (14, wet deposition)."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from camp_amd._lib import lib\n"
            "import numpy as np, ctypes as C\n"
            "vt = np.ones(3, np.int32)\n"
            "ip = C.POINTER(C.c_int); dp = C.POINTER(C.c_double)\n"
            "h = lib().solver_new(3, 1, vt.ctypes.data_as(ip), 1, 4, 1, 1,"
            " 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)\n"
            "ints = np.array([1, 1, 0, 0], np.int32); fl = np.zeros(1)\n"
            "lib().rxn_add_condensed_data(14, 4, 1, 1, ints.ctypes.data_as(ip),"
            " fl.ctypes.data_as(dp), h)\n"
            "at = np.full(3, 1e-8)\n"
            "lib().solver_initialize(h, at.ctypes.data_as(dp), 1e-4, 100, 10)\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "ERROR" in r.stderr


def test_server_rate_layout_follows_reaction_order():
    """The per-cell base rates a server request carries follow the order in
    which photolysis, emission and first-order-loss reactions were added."""
    from camp_amd.core import CampSolver
    from camp_amd.mechanisms import RXN_EMISSION, RXN_FIRST_ORDER_LOSS, RXN_PHOTOLYSIS
    from camp_amd.server import base_rate_matrix, rate_layout
    mech = cb05_scale_mechanism(1)
    s = CampSolver(mech, 1)
    lay = rate_layout(s)
    want = []
    for r in mech.reactions:
        if r.type in ("PHOTOLYSIS", "EMISSION", "FIRST_ORDER_LOSS"):
            want.append({"PHOTOLYSIS": RXN_PHOTOLYSIS, "EMISSION": RXN_EMISSION,
                         "FIRST_ORDER_LOSS": RXN_FIRST_ORDER_LOSS}[r.type])
    assert [t for t, _ in lay] == want and len(lay) == 29
    cells = synthetic_cells(mech, 5, seed=2)
    br = base_rate_matrix(s, cells.rates, 5)
    for j, (_, uid) in enumerate(lay):
        assert np.array_equal(br[:, j], cells.rates[uid])


def test_multiphase_pattern_ids_and_map_match_oracle(oracle_lib):
    """Product host model vs oracle on the multi-phase mechanism: reaction
    and solver patterns, derivative ids and the reaction int data after the
    id updates (SIMPOL phase Jacobian ids, aqueous-equilibrium ids)."""
    from camp_amd.core import CampSolver
    from camp_amd.multiphase import multiphase_mechanism
    mp = multiphase_mechanism(1, n_bins=3, n_simpol=5, n_aqeq=4)
    s = CampSolver(mp, 2)
    o = oracle_lib.Oracle(mp, 2)
    a, b = s.pattern(), o.pattern()
    for k in ("rxn_colptr", "rxn_rowval", "colptr", "rowval", "deriv_ids"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["rxn_int_data"], o.int_data())
    assert s.nv == o.nv and s.nnz == o.nnz


def test_elementary_functions_are_correctly_rounded():
    """The device's exp/log/log10/pow for the restated formulas (host build of
    the same code) against 250-bit values: every result is the double nearest
    the exact value, over the argument ranges the formulas meet."""
    import ctypes as C
    import mpmath
    from camp_amd import _lib
    L = _lib.lib()
    mpmath.mp.prec = 250
    rng = np.random.default_rng(11)
    n = 1500
    cases = {
        0: (np.concatenate([rng.uniform(-740, 709, n), rng.uniform(-5, 5, n)]), None,
            mpmath.exp),
        1: (np.concatenate([10 ** rng.uniform(-300, 300, n), rng.uniform(0.3, 3, n)]), None,
            mpmath.log),
        2: (10 ** rng.uniform(-40, 40, n), None, mpmath.log10),
        3: (np.concatenate([rng.uniform(0.01, 5, n), np.full(n, 10.0), rng.uniform(1e-9, 1e-5, n),
                            -rng.uniform(0.1, 30, n)]),
            np.concatenate([rng.uniform(-4, 4, n), rng.uniform(-12, 2, n),
                            rng.integers(0, 9, n).astype(float), rng.integers(-5, 9, n).astype(float)]),
            mpmath.power),
    }
    for fn, (x, y, exact) in cases.items():
        x = np.ascontiguousarray(x)
        yy = np.ascontiguousarray(y if y is not None else np.zeros_like(x))
        out = np.empty_like(x)
        dp = C.POINTER(C.c_double)
        L.camp_cr_math(fn, len(x), x.ctypes.data_as(dp), yy.ctypes.data_as(dp), out.ctypes.data_as(dp))
        if y is None:
            ref = [float(exact(mpmath.mpf(float(a)))) for a in x]
        else:
            ref = [float(exact(mpmath.mpf(float(a)), mpmath.mpf(float(b)))) for a, b in zip(x, yy)]
        bad = np.flatnonzero(out != np.array(ref))
        assert bad.size == 0, (fn, x[bad[:3]], yy[bad[:3]], out[bad[:3]])


def test_camp_files_config_loads_the_real_cb05():
    """camp_core's camp-files list (camp_core.F90:310-360): the real
    CB05cl_ae5 mechanism, species, tolerance and initial-condition files,
    with species objects of the same name merged across files.  The
    reference test counts 186 reactions (test_cb05cl_ae5.F90:367-370)."""
    from camp_amd.core import CampSolver
    m = load_camp_json_file(CB05CL_AE5)
    assert len(m.reactions) == 186
    types = {}
    for r in m.reactions:
        types[r.type] = types.get(r.type, 0) + 1
    assert types == {"ARRHENIUS": 142, "PHOTOLYSIS": 25, "TROE": 15, "CMAQ_H2O2": 3,
                     "CMAQ_OH_HNO3": 1}
    assert m.rel_tol == 1.0e-4
    const = sorted(s for s, v in zip(m.species, m.var_type) if v == 2)
    assert const == ["CH4", "H2", "H2O", "M", "N2", "O2"]
    assert m.init_conc[m.species.index("M")] == 1.0e6
    assert m.abs_tol[m.species.index("NO2")] == 1.0e-3
    s = CampSolver(m, 2)
    assert (s.nv, s.nnz) == (74, 671)


def test_cmaq_packing_follows_fortran_packers():
    """rxn_CMAQ_H2O2.F90 / rxn_CMAQ_OH_HNO3.F90 initialize: float layouts
    [k1_A, k1_B, k1_C, k2_A*1e6, k2_B, k2_C, CONV, yields] and [k0_A, k0_B,
    k0_C, k2_A, k2_B, k2_C, k3_A*1e6, k3_B, k3_C, SCALING, CONV, yields],
    defaults A = 1, B = C = 0, time unit MIN dividing the A's by 60 (H2O2)
    or giving SCALING = 1/60 (OH_HNO3)."""
    from camp_amd.mechanisms import Reaction, pack_reaction
    t, ints, fl, ne = pack_reaction(Reaction("CMAQ_H2O2", [0, 0], [(1, 0.5)],
                                             {"k1_A": 2.0, "k2_C": 5.0, "time unit": "MIN"}))
    assert (t, ne) == (3, 1) and ints[:6] == [2, 1, 1, 1, 2, 0]
    assert fl == [2.0 / 60.0, 0.0, 0.0, 1.0 / 60.0 * 1.0e6, 0.0, 5.0, CONV, 0.5]
    t, ints, fl, ne = pack_reaction(Reaction("CMAQ_OH_HNO3", [0, 1], [(2, 1.0)],
                                             {"k3_A": 3.0, "time unit": "MIN"}))
    assert (t, ne) == (4, 1) and ints[:5] == [2, 1, 1, 2, 3]
    assert fl == [1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 3.0e6, 0.0, 0.0, 1.0 / 60.0, CONV, 1.0]
