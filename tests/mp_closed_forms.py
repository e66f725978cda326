"""Closed forms of the reference's own SIMPOL and aqueous-equilibrium unit
tests, restated for BINNED aerosol mechanisms this solver accepts.

test/unit_rxn_data/test_rxn_SIMPOL_phase_transfer.F90:330-410 builds the
transition-regime condensation rate from first principles -- the mass
accommodation coefficient from N* (del_H, del_S), the mean free path
3 D / sqrt(8 R T / (pi MW)), the Knudsen number and the Fuchs-Sutugin-type
correction -- and the SIMPOL.1 vapour pressure 10^(B1/T + B2 + B3 T + B4 ln T)
atm.  Its fixture runs a MODAL section with UNIFAC activity coefficients;
here the same formulas are written for BINNED bins (radius dp/2, number
concentration = phase volume / particle volume) without activity
coefficients, where the gas-particle equilibrium is Raoult's law in each bin:
[gas] (ppm) = p_sat (ppm) x mole fraction of the species in the bin's phase.

test/unit_rxn_data/test_rxn_aqueous_equilibrium.F90:300-372 writes the
forward and reverse rates in molality (mol per kg of aerosol water:
conc / water * 1000 / MW) with K_eq = A exp(C (1/T - 1/298)), k_f = K_eq k_r,
and the equilibria of 2A <-> B + C, D + E <-> 2F and G <-> H in closed form.
Its fixture uses single-particle representations and an ion-pair activity
coefficient; here the same reactions run in BINNED bins without activity
coefficients (so k_r of the first reaction is 0.32, not 0.32 x 0.5).

Everything below is plain numpy from those definitions, independent of the
oracle's and the product's code.  The aerosol-phase sums carry the
reference's minimum mass, moles and volume (aero_phase_solver.c:16-20,
104-105, 155).
"""
import math

import numpy as np

from camp_amd.mechanisms import Mechanism
from camp_amd.multiphase import AeroPhase, AqEqRxn, BinnedRep, MultiPhaseMechanism, SimpolRxn

R = 8.314472                    # constants.F90:39
PI = 3.14159265358979323846
MIN_MASS, MIN_MW, MIN_DENS = 1.0e-25, 0.1, 1800.0

T_TEST, P_TEST = 272.5, 101253.3   # both reference tests

# --------------------------------------------------------------------------
# SIMPOL: ethanol partitioning into an aqueous phase (the reference test's
# species and parameters) in two BINNED bins
# --------------------------------------------------------------------------
ETHANOL_B = (-1.97e3, 2.91e0, 1.96e-3, -4.96e-1)
ETHANOL_MW, WATER_MW = 0.04607, 0.01801
ETHANOL_D, N_STAR = 9.50e-6, 2.55


def simpol_mechanism(n_bins=2, dens=1000.0):
    gas = Mechanism(["ethanol"], [1], [1.0e-20], [], 1.0e-10, "simpol_gas")
    ph = AeroPhase("aqueous aerosol", ["ethanol_aq", "H2O_aq"], [ETHANOL_MW, WATER_MW],
                   [dens, dens], [1, 1])
    rep = BinnedRep(n_bins, 4.0e-7, 1.6e-6, [ph])
    sp = SimpolRxn(gas=0, aero_name="ethanol_aq", phase="aqueous aerosol", B=ETHANOL_B,
                   diff_coeff=ETHANOL_D, mw=ETHANOL_MW, gas_mw=ETHANOL_MW, n_star=N_STAR)
    mp = MultiPhaseMechanism(gas=gas, rep=rep, simpol=[sp], aqeq=[], zsr=None,
                             name="simpol_box")
    mp.abs_tol = [1.0e-20] * mp.n_state
    return mp


def simpol_state(mp, gas_ppm=0.1, eth_aq=(2.0e-10, 5.0e-11), water=(1.0e-8, 3.0e-8)):
    st = np.zeros(mp.n_state)
    st[0] = gas_ppm
    for b in range(mp.rep.n_bins):
        st[mp.aero_id("aqueous aerosol", b, "ethanol_aq")] = eth_aq[b]
        st[mp.aero_id("aqueous aerosol", b, "H2O_aq")] = water[b]
    return st


def _simpol_constants(T, P):
    ns = N_STAR
    del_h = -10.0 * (ns - 1.0) + 7.53 * (ns ** (2.0 / 3.0) - 1.0) - 1.0
    del_s = -32.0 * (ns - 1.0) + 9.21 * (ns ** (2.0 / 3.0) - 1.0) - 1.3
    del_g = (del_h - T * del_s / 1000.0) * 4184.0
    alpha = math.exp(-del_g / (R * T))
    alpha = alpha / (1.0 + alpha)
    mfp = 3.0 * ETHANOL_D / math.sqrt(8.0 * R * T / (PI * ETHANOL_MW))
    vp = 101325.0 * 10.0 ** (ETHANOL_B[0] / T + ETHANOL_B[1] + ETHANOL_B[2] * T
                             + ETHANOL_B[3] * math.log(T))
    kgm3_to_ppm = R * 1.0e6 * T / (ETHANOL_MW * P)
    psat_ppm = vp / P * 1.0e6
    return alpha, mfp, kgm3_to_ppm, psat_ppm


def _bin_terms(mp, st, b, T, P):
    alpha, mfp, kgm3_to_ppm, psat_ppm = _simpol_constants(T, P)
    dp = mp.rep.bin_dp()[b]
    r = dp / 2.0
    x = st[mp.aero_id("aqueous aerosol", b, "ethanol_aq")]
    w = st[mp.aero_id("aqueous aerosol", b, "H2O_aq")]
    ph = mp.rep.phases[0]
    vol = MIN_MASS / MIN_DENS + x / ph.density[0] + w / ph.density[1]
    n_conc = vol / (4.0 / 3.0 * PI * r ** 3)
    kn = mfp / r
    corr = 0.75 * alpha * (1.0 + kn) / (kn ** 2 + (1.0 + 0.283 * alpha) * kn + 0.75 * alpha)
    kc = 4.0 * PI * r * ETHANOL_D * corr             # m3 / particle / s
    moles = MIN_MASS / MIN_MW + x / ETHANOL_MW + w / WATER_MW
    cond = n_conc * kc * st[0]                       # ppm / s
    evap = n_conc * kc * psat_ppm * (x / ETHANOL_MW) / moles
    return cond, evap, kgm3_to_ppm


def simpol_f(mp, st, T=T_TEST, P=P_TEST):
    """d/dt of every state variable (gas ppm/s, aerosol kg m-3 s-1)."""
    f = np.zeros(mp.n_state)
    for b in range(mp.rep.n_bins):
        cond, evap, conv = _bin_terms(mp, st, b, T, P)
        f[0] += evap - cond
        f[mp.aero_id("aqueous aerosol", b, "ethanol_aq")] += (cond - evap) / conv
    return f


def simpol_equilibrium(mp, st0, T=T_TEST, P=P_TEST):
    """Raoult's law in every bin with the total ethanol mass conserved:
    solve for the gas mixing ratio y with
        y / kgm3_to_ppm + sum_b x_b(y) = total,
        x_b(y) = MW_e * (w_b / MW_w + m0) * (y / p) / (1 - y / p)
    (m0: the phase's minimum moles), by bisection to the last bit."""
    _, _, conv, psat = _simpol_constants(T, P)
    nb = mp.rep.n_bins
    xi = [mp.aero_id("aqueous aerosol", b, "ethanol_aq") for b in range(nb)]
    wi = [mp.aero_id("aqueous aerosol", b, "H2O_aq") for b in range(nb)]
    total = st0[0] / conv + sum(st0[i] for i in xi)

    def xs(y):
        q = y / psat
        return [ETHANOL_MW * (st0[w] / WATER_MW + MIN_MASS / MIN_MW) * q / (1.0 - q) for w in wi]

    lo, hi = 0.0, min(psat, total * conv)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid / conv + sum(xs(mid)) > total:
            hi = mid
        else:
            lo = mid
    y = 0.5 * (lo + hi)
    out = st0.copy()
    out[0] = y
    for i, x in zip(xi, xs(y)):
        out[i] = x
    return out


# --------------------------------------------------------------------------
# Aqueous equilibrium: the reference test's three reactions in two bins
# --------------------------------------------------------------------------
AQ_SPECIES = ["A", "B", "C", "D", "E", "F", "G", "H", "H2O_aq"]
AQ_MW = [48.0, 32.67, 114.3, 27.6, 202.4, 28.0, 35.67, 284.2, 18.01]
AQ_RXNS = [  # reactants (qty repeated), products, A, C, k_reverse
    (["A", "A"], ["B", "C"], 1.14e-2, 2300.0, 0.32),
    (["D", "E"], ["F", "F"], 12.3, 0.0, 3.25e-3),
    (["G"], ["H"], 2.35, 1245.7, 1.56e-4),
]


def aqeq_mechanism(n_bins=2):
    gas = Mechanism(["N2"], [2], [1.0e-8], [], 1.0e-10, "aqeq_gas")
    ph = AeroPhase("aqueous aerosol", AQ_SPECIES, AQ_MW, [1000.0] * len(AQ_SPECIES),
                   [1] * len(AQ_SPECIES))
    rep = BinnedRep(n_bins, 4.0e-7, 1.6e-6, [ph])
    rx = [AqEqRxn(reactants=r, products=p, phase="aqueous aerosol", A=a, C=c, k_reverse=kr)
          for r, p, a, c, kr in AQ_RXNS]
    mp = MultiPhaseMechanism(gas=gas, rep=rep, simpol=[], aqeq=rx, zsr=None, name="aqeq_box")
    mp.abs_tol = [1.0e-20] * mp.n_state
    return mp


def aqeq_state(mp, scale=(1.0, 0.6)):
    """The reference test's initial pattern (A, F, G present; B..E, H zero)
    scaled per bin, with water 1e-8 kg m-3 (molalities of order 1)."""
    st = np.zeros(mp.n_state)
    st[0] = 1.0
    for b in range(mp.rep.n_bins):
        s = scale[b]
        for name, v in (("A", 1.35e-9 * s), ("F", 0.8e-9 * s), ("G", 1.2e-9 * s),
                        ("B", 1.0e-12 * s), ("C", 2.0e-12 * s), ("D", 1.0e-12 * s),
                        ("E", 3.0e-12 * s), ("H", 1.0e-12 * s), ("H2O_aq", 1.0e-8)):
            st[mp.aero_id("aqueous aerosol", b, name)] = v
    return st


def _keq(a, c, T):
    return a if c == 0.0 else a * math.exp(c * (1.0 / T - 1.0 / 298.0))


def aqeq_f(mp, st, T=T_TEST):
    """d/dt (kg m-3 s-1) from the molality rates of every reaction."""
    f = np.zeros(mp.n_state)
    mw = dict(zip(AQ_SPECIES, AQ_MW))
    for b in range(mp.rep.n_bins):
        idx = {s: mp.aero_id("aqueous aerosol", b, s) for s in AQ_SPECIES}
        w = st[idx["H2O_aq"]]
        m = {s: st[idx[s]] / w * 1000.0 / mw[s] for s in AQ_SPECIES}
        for reac, prod, a, c, kr in AQ_RXNS:
            kf = _keq(a, c, T) * kr
            rate = kf * np.prod([m[s] for s in reac]) - kr * np.prod([m[s] for s in prod])
            for s in reac:
                f[idx[s]] -= rate * w * mw[s] / 1000.0
            for s in prod:
                f[idx[s]] += rate * w * mw[s] / 1000.0
    return f


def aqeq_equilibrium(mp, st0, T=T_TEST):
    """The reference test's equilibria (test_rxn_aqueous_equilibrium.F90:
    327-348) from each reaction's conserved total, per bin:
      2A <-> B + C:  [B] = [C] = T1 / (1/sqrt(K1) + 2) + ..., from the totals
      D + E <-> 2F, G <-> H likewise,
    solved exactly (with the small initial products included) by bisection
    on the extent of reaction."""
    out = st0.copy()
    mw = dict(zip(AQ_SPECIES, AQ_MW))
    for b in range(mp.rep.n_bins):
        idx = {s: mp.aero_id("aqueous aerosol", b, s) for s in AQ_SPECIES}
        w = st0[idx["H2O_aq"]]
        m = {s: st0[idx[s]] / w * 1000.0 / mw[s] for s in AQ_SPECIES}
        for reac, prod, a, c, kr in AQ_RXNS:
            K = _keq(a, c, T)
            nu = {}
            for s in reac:
                nu[s] = nu.get(s, 0) - 1
            for s in prod:
                nu[s] = nu.get(s, 0) + 1

            def q(xi):  # K * prod(reactants) - prod(products) at extent xi
                mm = {s: m[s] + nu[s] * xi for s in nu}
                return K * np.prod([mm[s] for s in reac]) - np.prod([mm[s] for s in prod])

            lo = max(-m[s] / nu[s] for s in nu if nu[s] > 0)   # products exhausted
            hi = min(-m[s] / nu[s] for s in nu if nu[s] < 0)   # reactants exhausted
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if q(mid) > 0.0:
                    lo = mid
                else:
                    hi = mid
            xi = 0.5 * (lo + hi)
            for s in nu:
                m[s] += nu[s] * xi
        for s in AQ_SPECIES[:-1]:
            out[idx[s]] = m[s] * w * mw[s] / 1000.0
    return out


def relax_GH(mp, st0, t, T=T_TEST):
    """G <-> H is linear: its exact transient (the reference's exponential,
    test_rxn_aqueous_equilibrium.F90:426-429) per bin."""
    K = _keq(2.35, 1245.7, T)
    kr = 1.56e-4
    kf = K * kr
    out = {}
    mw = dict(zip(AQ_SPECIES, AQ_MW))
    for b in range(mp.rep.n_bins):
        ig, ih = (mp.aero_id("aqueous aerosol", b, s) for s in ("G", "H"))
        w = st0[mp.aero_id("aqueous aerosol", b, "H2O_aq")]
        g0 = st0[ig] / w * 1000.0 / mw["G"]
        h0 = st0[ih] / w * 1000.0 / mw["H"]
        tot = g0 + h0
        geq = tot / (1.0 + K)
        g = geq + (g0 - geq) * math.exp(-(kf + kr) * t)
        out[ig] = g * w * mw["G"] / 1000.0
        out[ih] = (tot - g) * w * mw["H"] / 1000.0
    return out
