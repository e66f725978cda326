"""The reference's analytic-solution reaction tests, as (mechanism, inputs,
closed form) cases shared by the oracle tests and the GPU tests.

Each case reproduces one reference test program: the mechanism JSON it
loads (copied under tests/golden/), its conditions and rates, and its closed
form (file:line in each docstring).  Solutions are compared every
`time_step` for `n_steps` steps, as the reference does.
"""
import json
import math
import os

import numpy as np

from camp_amd.mechanisms import (BOLTZMANN, CONV, load_camp_json)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))


def _mech(name):
    return load_camp_json(json.load(open(os.path.join(GOLDEN, name))), name=name)


def troe_case():
    """test_rxn_troe.F90:98-112 (k1, k2) and :231-236 (A -> B -> C)."""
    k = KATS["troe"]
    T, P = k["temperature"], k["pressure"]
    mech = _mech("test_troe.json")
    conv = CONV * P / T
    k_0 = 1.0e6 * 4.0e-18 * conv
    k_inf = k_0 / 1.0
    k1 = k_0 / (1 + k_inf) * 0.6 ** (1.0 / (1.0 + (math.log10(k_inf) / 1.0) ** 2))
    k_0 = 1.0e6 * 1.2e-12 * math.exp(3.0 / T) * (T / 300.0) ** 167.0 * conv
    k_inf = 136.0 * math.exp(24.0 / T) * (T / 300.0) ** 5.0
    k_inf = k_0 / k_inf
    k2 = k_0 / (1 + k_inf) * 0.9 ** (1.0 / (1.0 + (math.log10(k_inf) / 0.8) ** 2)) / 60.0
    y0 = np.array([1.0, 0.0, 0.0])
    return dict(mech=mech, T=T, P=P, y0=y0, rates={}, k=(k1, k2), dt=k["time_step"],
                n_steps=k["n_steps"], solution=lambda t: _abc(y0[0], k1, k2, t))


def _abc(a0, k1, k2, t):
    return np.array([a0 * math.exp(-k1 * t),
                     a0 * (k1 / (k2 - k1)) * (math.exp(-k1 * t) - math.exp(-k2 * t)),
                     a0 * (1.0 + (k1 * math.exp(-k2 * t) - k2 * math.exp(-k1 * t)) / (k2 - k1))])


def photolysis_case():
    """test_rxn_photolysis.F90:113-116 (k1 = 0.954, k2 = 0.01 * 12.3), A->B->C."""
    mech = _mech("test_photolysis.json")
    ids = [r.rxn_id for r in mech.reactions]
    k1, k2 = 0.954, 1.0e-2 * 12.3
    y0 = np.array([1.0, 0.0, 0.0])
    return dict(mech=mech, T=272.5, P=101253.3, y0=y0, rates={ids[0]: 0.954, ids[1]: 1.0e-2},
                k=(k1, k2), dt=1.0, n_steps=100, solution=lambda t: _abc(1.0, k1, k2, t))


def emission_case():
    """test_rxn_emission.F90:113-118, 245-286: y = y0 + rate * scaling * t."""
    k = KATS["emission"]
    mech = _mech("test_emission.json")
    ids = [r.rxn_id for r in mech.reactions]
    k1, k2 = k["rate_A"], k["rate_B"] * k["scaling_B"]
    y0 = np.array([1.0, 1.0])
    return dict(mech=mech, T=k["temperature"], P=k["pressure"], y0=y0,
                rates={ids[0]: k["rate_A"], ids[1]: k["rate_B"]}, k=(k1, k2), dt=1.0,
                n_steps=100, solution=lambda t: np.array([1.0 + k1 * t, 1.0 + k2 * t]))


def first_order_loss_case():
    """test_rxn_first_order_loss.F90:115-118, 245-246: y = y0 exp(-k t)."""
    mech = _mech("test_first_order_loss.json")
    ids = [r.rxn_id for r in mech.reactions]
    k1, k2 = 0.954, 1.0e-2 * 12.3
    return dict(mech=mech, T=272.5, P=101253.3, y0=np.array([1.0, 1.0]),
                rates={ids[0]: 0.954, ids[1]: 1.0e-2}, k=(k1, k2), dt=1.0, n_steps=100,
                solution=lambda t: np.array([math.exp(-k1 * t), math.exp(-k2 * t)]))


def arrhenius_cases():
    """unit_test_rxn_arrhenius.F90:21-52 (5 initial states) and :335-364:
    A + D -> B and A -> B (k1), B -> C + D (k2), D constant."""
    k = KATS["arrhenius"]
    mech = _mech("rxn_arrhenius.json")   # state order: A, D, B, C
    out = []
    for s in range(5):
        T, P = k["temperature"][s], k["pressure"][s]
        conv = CONV * P / T
        d = k["conc_D_ppm_times_conv"][s] / conv
        k1 = d * conv + 1476.0 * math.exp(-5.5e-21 / (BOLTZMANN * T)) * (T / 300.0) ** 150.0 * \
            (1.0 + 0.15 * P) / 60.0
        k2 = 21.0 * math.exp(-4000.0 / T) * (T / 315.0) ** 11.0 * (1.0 + 0.05 * P)
        a0 = k["conc_A"][s]
        y0 = np.array([a0, d, 0.0, 0.0])

        def sol(t, a0=a0, k1=k1, k2=k2, d=d):
            abc = _abc(a0, k1, k2, t)
            return np.array([abc[0], d, abc[1], abc[2]])
        out.append(dict(mech=mech, T=T, P=P, y0=y0, rates={}, k=(k1, k2), dt=1.0,
                        n_steps=100, solution=sol))
    return out


def all_cases():
    return ([("troe", troe_case()), ("photolysis", photolysis_case()),
             ("emission", emission_case()), ("first_order_loss", first_order_loss_case())]
            + [("arrhenius%d" % i, c) for i, c in enumerate(arrhenius_cases())])


def almost_equal(a, b, rel_tol, abs_tol=1e-30):
    """util.F90:379-415 (elementwise)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.abs(a - b) / (np.abs(a) + np.abs(b)) < rel_tol
    return (a == b) | rel | (np.abs(a - b) <= abs_tol)


def reference_bar(model, true, first):
    """The reference tests' pass rule: almost_equal at 1e-2, or both values
    below 1e-5 of the reference magnitude (test_rxn_troe.F90:254-262,
    unit_test_rxn_arrhenius.F90:386-390)."""
    small = (model < 1e-5 * np.abs(first)) & (true < 1e-5 * np.abs(first))
    return almost_equal(model, true, 1e-2) | small
