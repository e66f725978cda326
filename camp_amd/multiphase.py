"""Multi-phase CAMP mechanisms: aerosol phases, the modal/binned-mass aerosol
representation, SIMPOL phase transfer, aqueous equilibrium and the ZSR
aerosol-water sub-model, packed the way the reference Fortran initialisers
pack them:

* aerosol phase       src/aero_phase_data.F90:306-360   (layout aero_phase_solver.c:31-39)
* modal/binned mass   src/aero_reps/aero_rep_modal_binned_mass.F90:334-661
                                                      (layout aero_rep_modal_binned_mass.c:27-81)
* SIMPOL              src/rxns/rxn_SIMPOL_phase_transfer.F90:181-431
                                                      (layout rxn_SIMPOL_phase_transfer.c:31-74)
* aqueous equilibrium src/rxns/rxn_aqueous_equilibrium.F90:150-400
                                                      (layout rxn_aqueous_equilibrium.c:27-68)
* ZSR aerosol water   src/sub_models/sub_model_ZSR_aerosol_water.F90:266-700
                                                      (layout sub_model_ZSR_aerosol_water.c:30-68)

Species ids in packed ints are 1-based state indices (the C side subtracts
1); slots the C initialisation fills later (Jacobian ids, per-cell scratch)
are packed as zeros, as the Fortran code leaves them.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .mechanisms import (CHEM_SPEC_CONSTANT, CHEM_SPEC_VARIABLE, UNIV_GAS_CONST, Mechanism,
                         cb05_scale_mechanism, pack_reaction, synthetic_cells)
from .rng import CellStreams, Xoshiro

# type codes: rxn_solver.c:19-35, aero_rep_solver.c:19-20, sub_model_solver.c:16-18
RXN_AQUEOUS_EQUILIBRIUM = 7
RXN_SIMPOL_PHASE_TRANSFER = 10
AERO_REP_MODAL_BINNED_MASS = 2
SUB_MODEL_ZSR_AEROSOL_WATER = 2
BINNED, MODAL = 1, 2            # aero_rep_modal_binned_mass.c:18-19
ACT_TYPE_JACOBSON = 1           # sub_model_ZSR_aerosol_water.c:20-21
PI = 3.14159265358979323846     # constants.F90


@dataclass
class AeroPhase:
    name: str
    species: list                 # names, in phase order
    mw: list                      # kg mol-1
    density: list                 # kg m-3
    var_type: list


@dataclass
class BinnedRep:
    """One BINNED section of a modal/binned-mass representation."""
    n_bins: int
    min_dp: float
    max_dp: float
    phases: list                  # AeroPhase objects, each in every bin
    scale: str = "LOG"

    def bin_dp(self):
        # aero_rep_modal_binned_mass.F90:559-574
        n = self.n_bins
        if self.scale == "LINEAR":
            return [self.min_dp + i * (self.max_dp - self.min_dp) / (n - 1) for i in range(n)]
        d = (math.log10(self.max_dp) - math.log10(self.min_dp)) / (n - 1)
        return [10.0 ** (math.log10(self.min_dp) + i * d) for i in range(n)]


@dataclass
class SimpolRxn:
    gas: int                      # 0-based gas state id
    aero_name: str                # aerosol-phase species name
    phase: str
    B: tuple                      # B1..B4
    diff_coeff: float             # m2 s-1
    mw: float                     # kg mol-1 (aerosol species)
    gas_mw: float                 # kg mol-1 (gas species)
    n_star: float = None


@dataclass
class AqEqRxn:
    reactants: list               # aerosol species names (repeated for qty)
    products: list
    phase: str
    A: float
    C: float
    k_reverse: float
    water: str = "H2O_aq"


@dataclass
class IonPair:
    cation: str
    anion: str
    n_cation: int
    n_anion: int
    Y: list
    low_rh: float


@dataclass
class ZsrModel:
    phase: str
    gas_water: int                # 0-based gas state id
    aero_water: str
    pairs: list


@dataclass
class MultiPhaseMechanism:
    gas: Mechanism                # gas species first in the state, then the aerosol
    rep: BinnedRep
    simpol: list
    aqeq: list
    zsr: ZsrModel                 # None: no aerosol-water sub model
    species: list = field(default_factory=list)
    var_type: list = field(default_factory=list)
    abs_tol: list = field(default_factory=list)
    rel_tol: float = 1.0e-4
    name: str = "multiphase"

    def __post_init__(self):
        self.species = list(self.gas.species)
        self.var_type = list(self.gas.var_type)
        self.abs_tol = list(self.gas.abs_tol)
        self.n_gas = len(self.species)
        self.phase_start = {}     # (phase name, bin) -> first state id
        for b in range(self.rep.n_bins):
            for ph in self.rep.phases:
                self.phase_start[(ph.name, b)] = len(self.species)
                for s, vt in zip(ph.species, ph.var_type):
                    self.species.append("%s.%d.%s" % (ph.name, b + 1, s))
                    self.var_type.append(vt)
                    self.abs_tol.append(self.aero_abs_tol)
        self.rel_tol = self.gas.rel_tol

    aero_abs_tol = 1.0e-8

    @property
    def n_state(self):
        return len(self.species)

    @property
    def reactions(self):
        return self.gas.reactions

    def phase(self, name):
        return next(p for p in self.rep.phases if p.name == name)

    def aero_id(self, phase, b, spec):
        ph = self.phase(phase)
        return self.phase_start[(phase, b)] + ph.species.index(spec)


# -------------------------------------------------------------------------
# packers
# -------------------------------------------------------------------------

def pack_aero_phase(ph):
    """(int_params, float_params): [NUM_STATE_VAR, SPEC_TYPE...],
    [MW..., DENSITY...] (aero_phase_data.F90:318-360)."""
    n = len(ph.species)
    return [n] + list(ph.var_type), list(ph.mw) + list(ph.density)


def pack_binned_rep(mp, phase_model_ids):
    """(type, ints, floats, n_env) of one modal/binned-mass representation
    with a single BINNED section (aero_rep_modal_binned_mass.F90:354-661).
    phase_model_ids: 1-based index of each phase in the aerosol-phase list."""
    rep = mp.rep
    ns, nb, nph = 1, rep.n_bins, len(rep.phases)
    n_int = 4 + 2 * ns + 3 + 3 * nb * nph
    n_float = 3 * nb + 2 * nb * nph
    ints = [0] * n_int
    flts = [0.0] * n_float
    ints[0] = ns
    ints[1] = n_int
    ints[2] = n_float
    ints[3] = -1                            # AERO_REP_ID_ (F90:655)
    loc_i = 4 + 2 * ns + 1                  # 1-based
    loc_f = 1
    ints[4] = loc_i
    ints[5] = loc_f
    i0 = loc_i - 1
    ints[i0] = BINNED
    ints[i0 + 1] = nb
    ints[i0 + 2] = nph
    dps = rep.bin_dp()
    for b in range(nb):
        flts[loc_f - 1 + 3 * b] = dps[b]           # BIN_DP
        flts[loc_f - 1 + 3 * b + 1] = -9999.9      # NUMBER_CONC
        flts[loc_f - 1 + 3 * b + 2] = dps[b] / 2.0  # EFFECTIVE_RADIUS
    for j, ph in enumerate(rep.phases):
        for b in range(nb):
            ints[i0 + 3 + b * nph + j] = mp.phase_start[(ph.name, b)] + 1   # PHASE_STATE_ID
            ints[i0 + 3 + nb * nph + b * nph + j] = phase_model_ids[j]      # PHASE_MODEL_DATA_ID
    return AERO_REP_MODAL_BINNED_MASS, ints, flts, 2 * ns


def _phase_ids(mp, phase):
    """1-based aerosol phase instance ids of `phase` in the representation
    (bin-major, one phase per bin when the rep has a single phase)."""
    nph = len(mp.rep.phases)
    j = [p.name for p in mp.rep.phases].index(phase)
    return [b * nph + j + 1 for b in range(mp.rep.n_bins)]


def _num_jac_elem(mp, phase_id):
    # aero_rep_modal_binned_mass.F90:964-998 with aero_phase_data.F90:416-432
    nph = len(mp.rep.phases)
    b = (phase_id - 1) // nph
    return sum(sum(1 for vt in ph.var_type if vt in (1, 2, 3)) for ph in mp.rep.phases)


def pack_simpol(mp, r):
    """rxn_SIMPOL_phase_transfer.F90:181-431."""
    pids = _phase_ids(mp, r.phase)
    n = len(pids)
    njac = [_num_jac_elem(mp, p) for p in pids]
    n_int = 2 + 2 + 13 * n + 2 * sum(njac)
    n_float = 10 + 4 * sum(njac)
    ints = [0] * n_int
    flts = [0.0] * n_float
    ints[0] = n
    ints[1] = r.gas + 1
    # PHASE_INT_LOC_(1) = NUM_INT_PROP_ + 12*n + 3, PHASE_REAL_LOC_(1) = 11
    loc_i, loc_f = 2 + 12 * n + 3, 10 + 1
    for i, pid in enumerate(pids):
        b = pid - 1
        ints[2 + 2 + 10 * n + i] = loc_i            # PHASE_INT_LOC_
        ints[2 + 2 + 11 * n + i] = loc_f            # PHASE_REAL_LOC_
        ints[loc_i - 1] = njac[i]                   # NUM_AERO_PHASE_JAC_ELEM_
        ints[2 + i] = mp.aero_id(r.phase, b, r.aero_name) + 1      # AERO_SPEC_
        ints[2 + n + i] = -1                                       # AERO_ACT_ID_: none (F90:338)
        ints[2 + 2 * n + i] = pid                                  # AERO_PHASE_ID_
        ints[2 + 3 * n + i] = 1                                    # AERO_REP_ID_
        loc_i += 1 + 2 * njac[i]
        loc_f += 4 * njac[i]
    if r.n_star is not None:
        ns_ = r.n_star
        dh = (-10.0 * (ns_ - 1.0) + 7.53 * (ns_ ** (2.0 / 3.0) - 1.0) - 1.0) * 4184.0
        ds = (-32.0 * (ns_ - 1.0) + 9.21 * (ns_ ** (2.0 / 3.0) - 1.0) - 1.3) * 4.184
    else:
        dh = ds = 0.0
    flts[0] = dh
    flts[1] = ds
    flts[2] = r.diff_coeff
    flts[3] = math.sqrt(8.0 * UNIV_GAS_CONST / (PI * r.gas_mw))
    flts[4:8] = list(r.B)
    flts[8] = UNIV_GAS_CONST / r.mw * 1.0e6
    flts[9] = r.mw
    return RXN_SIMPOL_PHASE_TRANSFER, ints, flts, 4


def pack_aqeq(mp, r):
    """rxn_aqueous_equilibrium.F90:150-400 (no activity coefficients)."""
    pids = _phase_ids(mp, r.phase)
    nph = len(pids)
    nr, npd = len(r.reactants), len(r.products)
    n_int = 3 + (nr + npd + 2) * nph + (nr + npd) * nph + (nr + npd) * (nr + npd + 2) * nph
    n_float = 5 + 2 * (nr + npd) + 2 * nph
    ints = [0] * n_int
    flts = [0.0] * n_float
    ints[0], ints[1], ints[2] = nr, npd, nph
    for i, pid in enumerate(pids):
        b = pid - 1
        for j, s in enumerate(r.reactants):
            ints[3 + i * nr + j] = mp.aero_id(r.phase, b, s) + 1
        for j, s in enumerate(r.products):
            ints[3 + nr * nph + i * npd + j] = mp.aero_id(r.phase, b, s) + 1
        ints[3 + (nr + npd) * nph + i] = mp.aero_id(r.phase, b, r.water) + 1
        ints[3 + (nr + npd + 1) * nph + i] = 0   # ACTIVITY_COEFF_: none (read as -1 in C)
    flts[0] = r.A
    flts[1] = r.C
    flts[2] = r.k_reverse
    ph = mp.phase(r.phase)
    for j, s in enumerate(r.reactants):
        flts[5 + j] = 1.0e3 / ph.mw[ph.species.index(s)]
    for j, s in enumerate(r.products):
        flts[5 + nr + j] = 1.0e3 / ph.mw[ph.species.index(s)]
    return RXN_AQUEOUS_EQUILIBRIUM, ints, flts, 1


def pack_zsr(mp):
    """sub_model_ZSR_aerosol_water.F90:266-700 (JACOBSON ion pairs)."""
    z = mp.zsr
    pids = _phase_ids(mp, z.phase)
    nph, npair = len(pids), len(z.pairs)
    n_int = 5 + nph + 2 * npair + sum(6 + 3 * nph for _ in z.pairs)
    n_float = sum(3 + len(p.Y) for p in z.pairs)
    ints = [0] * n_int
    flts = [0.0] * n_float
    ints[0], ints[1], ints[2], ints[3], ints[4] = nph, z.gas_water + 1, npair, n_int, n_float
    water0 = []
    for i, pid in enumerate(pids):
        w = mp.aero_id(z.phase, pid - 1, z.aero_water)
        ints[5 + i] = w + 1
        water0.append(w)
    li, lf = 5 + nph + 2 * npair, 0
    for k, p in enumerate(z.pairs):
        ints[5 + nph + k] = li + 1
        ints[5 + nph + npair + k] = lf + 1
        ph = mp.phase(z.phase)
        ints[li] = ACT_TYPE_JACOBSON
        ints[li + 1] = p.n_cation
        ints[li + 2] = p.n_anion
        ints[li + 3] = mp.aero_id(z.phase, 0, p.cation) - water0[0]
        ints[li + 4] = mp.aero_id(z.phase, 0, p.anion) - water0[0]
        ints[li + 5] = len(p.Y)
        flts[lf] = p.low_rh
        flts[lf + 1] = ph.mw[ph.species.index(p.cation)]
        flts[lf + 2] = ph.mw[ph.species.index(p.anion)]
        flts[lf + 3:lf + 3 + len(p.Y)] = list(p.Y)
        li += 6 + 3 * nph
        lf += 3 + len(p.Y)
    return SUB_MODEL_ZSR_AEROSOL_WATER, ints, flts, 1


def pack_multiphase(mp):
    """Everything the C ABI needs, in the order camp_solver_data adds it:
    {'rxn': [(type, ints, floats, n_env)], 'phase': [(ints, floats)],
     'rep': [(type, ints, floats, n_env)], 'sub': [(type, ints, floats, n_env)]}"""
    rxn = [pack_reaction(r) for r in mp.gas.reactions]
    rxn += [pack_aqeq(mp, r) for r in mp.aqeq]
    rxn += [pack_simpol(mp, r) for r in mp.simpol]
    phases = [pack_aero_phase(ph) for ph in mp.rep.phases]
    reps = [pack_binned_rep(mp, list(range(1, len(mp.rep.phases) + 1)))]
    subs = [pack_zsr(mp)] if mp.zsr is not None else []
    return {"rxn": rxn, "phase": phases, "rep": reps, "sub": subs}


# -------------------------------------------------------------------------
# the seeded multi-phase recipe
# -------------------------------------------------------------------------

# Jacobson et al. (1996) Table 2 molality polynomials, as the reference's
# own test mechanism carries them (test/monarch/solver_test_mech.json)
_ION_PAIRS = [
    ("H_p", "HSO4_m", 1, 1, [5.611895, -1.387446e1, 1.750682e1, 7.138146e1, -3.109173e2,
                             4.662288e2, -3.128612e2, 7.767097e1], 0.0),
    ("H_p", "SO4_mm", 2, 1, [5.611895, -1.387446e1, 1.750682e1, 7.138146e1, -3.109173e2,
                             4.662288e2, -3.128612e2, 7.767097e1], 0.0),
    ("H_p", "NO3_m", 1, 1, [4.852977, -6.621314, 3.390133e1, -1.985191e2, 6.281150e2,
                            -1.038494e3, 8.498917e2, -2.729090e2], 0.0),
    ("NH4_p", "HSO4_m", 1, 1, [5.515580, 3.588744, -6.363443e1, 3.687630e2, -1.023972e3,
                               1.394436e3, -9.168213e2, 2.328726e2], 0.0),
    ("NH4_p", "SO4_mm", 2, 1, [4.363511e2, -4.947645e3, 2.399693e4, -6.364664e4, 9.952891e4,
                               -9.179112e4, 4.626748e4, -9.844195e3], 0.47),
]
# name, MW (kg mol-1), density (kg m-3)
_INORGANIC = [("H_p", 0.001008, 1000.0), ("NH4_p", 0.018038, 1770.0), ("Na_p", 0.0229898, 2160.0),
              ("SO4_mm", 0.09606, 1770.0), ("HSO4_m", 0.097071, 1770.0),
              ("NO3_m", 0.0620049, 1720.0), ("Cl_m", 0.035453, 2160.0),
              ("NH3_aq", 0.017031, 1000.0), ("HNO3_aq", 0.0630129, 1510.0)]


def multiphase_mechanism(seed=1, n_bins=8, n_simpol=20, n_aqeq=10):
    """The CB05-scale gas mechanism plus n_bins aerosol size bins of one
    aqueous phase with 30 species (water, n_simpol condensable organics, 9
    inorganic ions/solutes): SIMPOL partitioning of n_simpol gas species
    (vapour pressure at 298 K log-uniform 1e-10..1e-2 Pa), n_aqeq aqueous
    equilibria (K log-uniform 1e-8..1e2) and ZSR water uptake."""
    gas = cb05_scale_mechanism(seed)
    rng = Xoshiro(seed * 7919 + 17)
    # gas-phase water (CONSTANT), appended after the gas mechanism's species
    gas.species = list(gas.species) + ["H2O"]
    gas.var_type = list(gas.var_type) + [CHEM_SPEC_CONSTANT]
    gas.abs_tol = list(gas.abs_tol) + [1.0e-8]
    h2o = len(gas.species) - 1
    var_gas = [i for i in range(h2o) if gas.var_type[i] == CHEM_SPEC_VARIABLE]
    condensing = sorted(var_gas[i] for i in rng.sample(len(var_gas), n_simpol))
    soa = ["SOA%02d" % k for k in range(n_simpol)]
    soa_mw = [rng.uniform(0.1, 0.3) for _ in range(n_simpol)]
    names = ["H2O_aq"] + soa + [n for n, _, _ in _INORGANIC]
    mw = [0.01801] + soa_mw + [m for _, m, _ in _INORGANIC]
    dens = [1000.0] + [1400.0] * n_simpol + [d for _, _, d in _INORGANIC]
    vt = [CHEM_SPEC_CONSTANT] + [CHEM_SPEC_VARIABLE] * (len(names) - 1)
    phase = AeroPhase("aqueous", names, mw, dens, vt)
    rep = BinnedRep(n_bins, 1.0e-8, 1.0e-5, [phase])
    simpol = []
    for k, g in enumerate(condensing):
        p298 = rng.log_uniform(1.0e-10, 1.0e-2)     # Pa
        b1 = rng.uniform(-3000.0, -1000.0)          # K
        b2 = math.log10(p298 / 101325.0) - b1 / 298.0
        simpol.append(SimpolRxn(gas=g, aero_name=soa[k], phase="aqueous", B=(b1, b2, 0.0, 0.0),
                                diff_coeff=rng.uniform(5.0e-6, 2.0e-5), mw=soa_mw[k],
                                gas_mw=soa_mw[k], n_star=rng.uniform(1.5, 3.0)))
    solutes = names[1:]
    aqeq = []
    for _ in range(n_aqeq):
        picks = [solutes[i] for i in rng.sample(len(solutes), 4)]
        nr = 1 + rng.randint(0, 1)
        npd = 1 + rng.randint(0, 1)
        aqeq.append(AqEqRxn(reactants=picks[:nr], products=picks[nr:nr + npd], phase="aqueous",
                            A=rng.log_uniform(1.0e-8, 1.0e2), C=rng.uniform(0.0, 3000.0),
                            k_reverse=rng.log_uniform(1.0e-3, 1.0)))
    zsr = ZsrModel("aqueous", h2o, "H2O_aq",
                   [IonPair(c, a, nc, na, y, lo) for c, a, nc, na, y, lo in _ION_PAIRS])
    return MultiPhaseMechanism(gas=gas, rep=rep, simpol=simpol, aqeq=aqeq, zsr=zsr,
                               name="cb05scale_multiphase")


def ppm_to_rh(T, P):
    """sub_model_ZSR_aerosol_water.c:195-206 (1/ppm)."""
    a = 1.0 - 373.15 / T
    a = (((-0.1299 * a - 0.6445) * a - 1.976) * a + 13.3185) * a
    return P / (101325.0 * math.exp(a)) / 1.0e6


def multiphase_cells(mp, n_cells, seed=2, cell_offset=0):
    """Per-cell inputs: the gas cells of the CB05-scale recipe, RH uniform
    0.3..0.95 set through gas-phase water, aerosol species log-uniform
    (organics 1e-3..10, inorganics 1e-3..5 ug m-3), water from ZSR at the
    first f() evaluation."""
    gcells = synthetic_cells(mp.gas, n_cells, seed=seed, cell_offset=cell_offset)
    st = np.zeros((n_cells, mp.n_state))
    st[:, :mp.n_gas] = gcells.state[:, :mp.n_gas]
    cs = CellStreams(seed * 31 + 5, np.arange(cell_offset, cell_offset + n_cells))
    rh = cs.uniform(0.3, 0.95)
    for c in range(n_cells):
        T, P = gcells.env[c]
        st[c, mp.zsr.gas_water] = rh[c] / ppm_to_rh(T, P)
    ph = mp.rep.phases[0]
    for b in range(mp.rep.n_bins):
        for j, s in enumerate(ph.species):
            i = mp.aero_id(ph.name, b, s)
            if s == "H2O_aq":
                st[:, i] = 1.0
            elif s.startswith("SOA"):
                st[:, i] = cs.log_uniform(1.0e-3, 10.0)
            else:
                st[:, i] = cs.log_uniform(1.0e-3, 5.0)
    gcells.state = st
    return gcells
