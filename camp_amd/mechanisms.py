"""Mechanism descriptions, CAMP parameter packing, and seeded synthetic inputs.

This module plays the part of CAMP's Fortran model-description layer: it
turns a mechanism (from CAMP JSON ``camp-data`` objects or from the seeded
synthetic recipes) into the packed ``int``/``double`` parameter arrays that
the C ABI ``rxn_add_condensed_data`` consumes.  The packed layouts follow
the reference packers slot for slot:

* ARRHENIUS  src/rxns/rxn_arrhenius.F90:176-230 (layout rxn_arrhenius.c:20-36)
* TROE       src/rxns/rxn_troe.F90:196-245       (layout rxn_troe.c:20-39)
* PHOTOLYSIS src/rxns/rxn_photolysis.F90:229-309 (layout rxn_photolysis.c:20-33)
* EMISSION   src/rxns/rxn_emission.F90:193-217   (layout rxn_emission.c:20-28)
* FIRST_ORDER_LOSS src/rxns/rxn_first_order_loss.F90:196-220
                                                  (layout rxn_first_order_loss.c:20-29)

Species ids inside the packed ints are 1-based state indices, as in CAMP.
"""
import json
import math
from dataclasses import dataclass, field

import numpy as np

from .rng import Xoshiro, CellStreams

# Physical constants, src/constants.F90:37-41
BOLTZMANN = 1.3806505e-23
AVOGADRO = 6.02214179e23
UNIV_GAS_CONST = 8.314472
# CONV = N_A / R * 1e-12, rxn_arrhenius.F90:191 and rxn_troe.F90:194
CONV = AVOGADRO / UNIV_GAS_CONST * 1e-12

# Reaction type codes, src/rxn_solver.c:19-35
RXN_ARRHENIUS = 1
RXN_TROE = 2
RXN_CMAQ_H2O2 = 3
RXN_CMAQ_OH_HNO3 = 4
RXN_PHOTOLYSIS = 5
RXN_FIRST_ORDER_LOSS = 12
RXN_EMISSION = 13
TYPE_CODES = {"ARRHENIUS": RXN_ARRHENIUS, "TROE": RXN_TROE, "CMAQ_H2O2": RXN_CMAQ_H2O2,
              "CMAQ_OH_HNO3": RXN_CMAQ_OH_HNO3, "PHOTOLYSIS": RXN_PHOTOLYSIS,
              "FIRST_ORDER_LOSS": RXN_FIRST_ORDER_LOSS, "EMISSION": RXN_EMISSION}
UPDATABLE = ("PHOTOLYSIS", "EMISSION", "FIRST_ORDER_LOSS")

# State variable types, src/camp_common.h:34-38
CHEM_SPEC_VARIABLE = 1
CHEM_SPEC_CONSTANT = 2

DEFAULT_ABS_TOL = 1.0e-14     # chem_spec_data.F90:63
DEFAULT_REL_TOL = 1.0e-8      # camp_solver_data.F90:33
DEFAULT_MAX_STEPS = 10000     # camp_solver_data.F90:35
DEFAULT_MAX_CONV_FAILS = 1000  # camp_solver_data.F90:37


@dataclass
class Reaction:
    type: str
    reactants: list            # 0-based state ids, repeated for "qty" > 1
    products: list = field(default_factory=list)   # [(0-based id, yield)]
    params: dict = field(default_factory=dict)
    rxn_id: int = -1           # unique id of updatable reactions (> 0)


@dataclass
class Mechanism:
    species: list
    var_type: list
    abs_tol: list
    reactions: list
    rel_tol: float = DEFAULT_REL_TOL
    name: str = "mechanism"
    init_conc: list = None     # per species "init conc" of the input files (ppm)

    @property
    def n_state(self):
        return len(self.species)

    def updatable(self, rtype):
        return [r for r in self.reactions if r.type == rtype]


def pack_reaction(r):
    """Pack one reaction the way the reference Fortran initialize() does.

    Returns (type_code, int_params, float_params, n_env)."""
    p = r.params
    if r.type in ("ARRHENIUS", "TROE", "PHOTOLYSIS", "CMAQ_H2O2", "CMAQ_OH_HNO3"):
        nr, npd = len(r.reactants), len(r.products)
        head = [nr, npd] + ([r.rxn_id] if r.type == "PHOTOLYSIS" else [])
        ints = head + [s + 1 for s in r.reactants] + [s + 1 for s, _ in r.products]
        ints += [0] * (nr + npd) + [0] * (nr * (nr + npd))   # DERIV_ID, JAC_ID slots
        yields = [y for _, y in r.products]
        if r.type == "ARRHENIUS":
            A = p.get("A", 1.0)
            if p.get("time unit") == "MIN":
                A = A / 60.0
            if "Ea" in p:
                assert "C" not in p, "Received both Ea and C parameter for Arrhenius equation"
                C = -p["Ea"] / BOLTZMANN
            else:
                C = p.get("C", 0.0)
            D = p.get("D", 300.0)
            B = p.get("B", 0.0)
            E = p.get("E", 0.0)
            assert not (B != 0.0 and D == 0.0)
            floats = [A, B, C, D, E, CONV] + yields
            return RXN_ARRHENIUS, ints, floats, 1
        if r.type == "TROE":
            scaling = 1.0 / 60.0 if p.get("time unit") == "MIN" else 1.0
            floats = [p.get("k0_A", 1.0) * 1.0e6, p.get("k0_B", 0.0), p.get("k0_C", 0.0),
                      p.get("kinf_A", 1.0), p.get("kinf_B", 0.0), p.get("kinf_C", 0.0),
                      p.get("Fc", 0.6), p.get("N", 1.0), scaling, CONV] + yields
            return RXN_TROE, ints, floats, 1
        if r.type == "CMAQ_H2O2":
            # rxn_CMAQ_H2O2.F90 initialize: k = k1 + [M] k2, [M] (ppm -> #/cc
            # through conv) folded into k2_A with a factor 1e6
            k1A, k2A = p.get("k1_A", 1.0), p.get("k2_A", 1.0)
            if p.get("time unit") == "MIN":
                k1A, k2A = k1A / 60.0, k2A / 60.0
            floats = [k1A, p.get("k1_B", 0.0), p.get("k1_C", 0.0), k2A * 1.0e6,
                      p.get("k2_B", 0.0), p.get("k2_C", 0.0), CONV] + yields
            return RXN_CMAQ_H2O2, ints, floats, 1
        if r.type == "CMAQ_OH_HNO3":
            # rxn_CMAQ_OH_HNO3.F90 initialize: k = k0 + k3 [M] / (1 + k3 [M] /
            # k2), k3_A times 1e6, time unit MIN as SCALING = 1/60
            scaling = 1.0 / 60.0 if p.get("time unit") == "MIN" else 1.0
            floats = [p.get("k0_A", 1.0), p.get("k0_B", 0.0), p.get("k0_C", 0.0),
                      p.get("k2_A", 1.0), p.get("k2_B", 0.0), p.get("k2_C", 0.0),
                      p.get("k3_A", 1.0) * 1.0e6, p.get("k3_B", 0.0), p.get("k3_C", 0.0),
                      scaling, CONV] + yields
            return RXN_CMAQ_OH_HNO3, ints, floats, 1
        floats = [p.get("scaling factor", 1.0)] + yields
        return RXN_PHOTOLYSIS, ints, floats, 2
    if r.type == "EMISSION":
        return RXN_EMISSION, [r.rxn_id, r.products[0][0] + 1, 0], [p.get("scaling factor", 1.0)], 2
    if r.type == "FIRST_ORDER_LOSS":
        return RXN_FIRST_ORDER_LOSS, [r.rxn_id, r.reactants[0] + 1, 0, 0], \
            [p.get("scaling factor", 1.0)], 2
    raise ValueError("unsupported reaction type " + r.type)


def pack(mech):
    """Packed parameter arrays for every reaction, in mechanism order."""
    return [pack_reaction(r) for r in mech.reactions]


# --------------------------------------------------------------------------
# CAMP JSON (camp-data) loader for the supported object types
# --------------------------------------------------------------------------

def load_camp_json(objects, name="json"):
    """Build a Mechanism from a list of ``camp-data`` objects.

    Follows camp_core.F90:486-637 (object dispatch: CHEM_SPEC,
    RELATIVE_TOLERANCE, MECHANISM) and chem_spec_data.F90:212-320 (a
    species named again -- in another file -- is merged into the first
    object of that name, which fixes its place in the state; later
    properties update earlier ones; 'tracer type', 'absolute integration
    tolerance', 'init conc').  Mechanisms of the same name are merged too.
    Updatable reactions get ids 1, 2, ... in mechanism order, standing in
    for generate_int_id() (util.F90)."""
    if isinstance(objects, dict):
        objects = objects["camp-data"]
    spec = {}          # name -> merged properties, in order of first appearance
    rel_tol = DEFAULT_REL_TOL
    mechs = {}         # mechanism name -> reactions (first-appearance order)
    for obj in objects:
        t = obj.get("type")
        if t == "CHEM_SPEC":
            if obj.get("phase", "GAS") != "GAS":
                raise ValueError("only gas-phase species are supported")
            props = spec.setdefault(obj["name"], {})
            old = props.get("tracer type")
            new = obj.get("tracer type")
            if old is not None and new is not None and old != new:
                raise ValueError("Type mismatch for species " + obj["name"])
            props.update({k: v for k, v in obj.items() if k not in ("name", "type")})
        elif t == "RELATIVE_TOLERANCE":
            rel_tol = float(obj["value"])
        elif t == "MECHANISM":
            mechs.setdefault(obj.get("name", ""), []).extend(obj["reactions"])
    species = list(spec)
    var_type, abs_tol, init = [], [], []
    for s in species:
        props = spec[s]
        tt = props.get("tracer type", "VARIABLE")
        if tt not in ("VARIABLE", "CONSTANT"):
            raise ValueError("unsupported tracer type %s for %s" % (tt, s))
        var_type.append(CHEM_SPEC_CONSTANT if tt == "CONSTANT" else CHEM_SPEC_VARIABLE)
        abs_tol.append(float(props.get("absolute integration tolerance", DEFAULT_ABS_TOL)))
        init.append(float(props.get("init conc", 0.0)))
    idx = {s: i for i, s in enumerate(species)}
    reactions = []
    next_id = 1
    for rj in (r for rx in mechs.values() for r in rx):
        rtype = rj["type"]
        params = {k: v for k, v in rj.items() if k not in ("type", "reactants", "products")}
        reactants = []
        try:
            for s, props in rj.get("reactants", {}).items():
                reactants += [idx[s]] * int(props.get("qty", 1))
            products = [(idx[s], float(props.get("yield", 1.0)))
                        for s, props in rj.get("products", {}).items()]
            if rtype == "EMISSION":
                products = [(idx[rj["species"]], 1.0)]
            if rtype == "FIRST_ORDER_LOSS":
                reactants = [idx[rj["species"]]]
        except KeyError as e:
            raise ValueError("reaction %s names species %s, which no CHEM_SPEC object "
                             "declares" % (rj.get("rxn id", rtype), e.args[0])) from None
        r = Reaction(rtype, reactants, products, params)
        if rtype in UPDATABLE:
            r.rxn_id = next_id
            next_id += 1
        reactions.append(r)
    return Mechanism(species, var_type, abs_tol, reactions, rel_tol, name, init)


def load_camp_json_file(path):
    """One camp-data file, or a configuration file whose "camp-files" list
    names the files to load, as camp_core%load_files and load read them
    (camp_core.F90:310-360, 424-486).  Relative paths resolve against the
    configuration file's directory (the reference opens them relative to the
    working directory, the same place when run from there)."""
    import os
    with open(path) as f:
        top = json.load(f)
    if isinstance(top, dict) and "camp-files" in top:
        base = os.path.dirname(os.path.abspath(path))
        objects = []
        for rel in top["camp-files"]:
            with open(os.path.join(base, rel)) as f:
                objects.extend(json.load(f)["camp-data"])
        return load_camp_json(objects, name=path)
    return load_camp_json(top, name=path)


# --------------------------------------------------------------------------
# Seeded synthetic recipes
# --------------------------------------------------------------------------

def _random_gas_reaction(rng, rtype, n_species, n_react):
    n_prod = rng.randint(1, 3)
    ids = rng.sample(n_species, n_react + n_prod)
    products = [(s, rng.uniform(0.1, 1.0)) for s in ids[n_react:]]
    return Reaction(rtype, ids[:n_react], products, {})


def synthetic_mechanism(n_species, n_arrhenius, n_troe=0, n_photolysis=0, n_emission=0,
                        n_first_order_loss=0, seed=1, abs_tol=1.0e-8, rel_tol=1.0e-4,
                        name="synthetic"):
    """Seeded random gas-phase mechanism.

    Arrhenius: 1-2 reactants, 1-3 products (all species distinct), yields
    U(0.1, 1); A log-uniform 1e-4..1e2 s^-1 (unimolecular) or 1e-16..1e-10
    (bimolecular); C ~ U(0, 2000) K entering as k ~ exp(-C/T), which CAMP
    stores as a negative C (rxn_arrhenius.c:112); B ~ U(-2, 2); D = 300.
    Troe: 2 reactants; k0_A log-uniform 1e-33..1e-28, kinf_A log-uniform
    1e-12..1e-10, Fc = 0.6, N = 1.  Photolysis: 1 reactant, rate per cell.
    Emission and first-order loss: one species, rate per cell.  The
    reaction list is shuffled so the types interleave."""
    rng = Xoshiro(seed)
    rx = []
    for _ in range(n_arrhenius):
        nr = rng.randint(1, 2)
        r = _random_gas_reaction(rng, "ARRHENIUS", n_species, nr)
        lo, hi = (1e-4, 1e2) if nr == 1 else (1e-16, 1e-10)
        r.params = {"A": rng.log_uniform(lo, hi), "C": -rng.uniform(0.0, 2000.0),
                    "B": rng.uniform(-2.0, 2.0), "D": 300.0, "E": 0.0}
        rx.append(r)
    for _ in range(n_troe):
        r = _random_gas_reaction(rng, "TROE", n_species, 2)
        r.params = {"k0_A": rng.log_uniform(1e-33, 1e-28), "kinf_A": rng.log_uniform(1e-12, 1e-10),
                    "Fc": 0.6, "N": 1.0}
        rx.append(r)
    for _ in range(n_photolysis):
        rx.append(_random_gas_reaction(rng, "PHOTOLYSIS", n_species, 1))
    for _ in range(n_emission):
        s = rng.randint(0, n_species - 1)
        rx.append(Reaction("EMISSION", [], [(s, 1.0)], {}))
    for _ in range(n_first_order_loss):
        s = rng.randint(0, n_species - 1)
        rx.append(Reaction("FIRST_ORDER_LOSS", [s], [], {}))
    # Fisher-Yates shuffle so that reaction types interleave
    for i in range(len(rx) - 1, 0, -1):
        j = int(rng.uniform() * (i + 1)) % (i + 1)
        rx[i], rx[j] = rx[j], rx[i]
    next_id = 1
    for r in rx:
        if r.type in UPDATABLE:
            r.rxn_id = next_id
            next_id += 1
    species = ["S%03d" % i for i in range(n_species)]
    return Mechanism(species, [CHEM_SPEC_VARIABLE] * n_species, [abs_tol] * n_species, rx,
                     rel_tol, name)


def box_mechanism(seed=1):
    """10 species, 15 Arrhenius reactions (the single-cell box-model case)."""
    return synthetic_mechanism(10, 15, seed=seed, name="box10x15")


def cb05_scale_mechanism(seed=1):
    """72 species, 190 reactions: 133 Arrhenius, 28 Troe, 19 photolysis,
    5 emission, 5 first-order loss (CB05-scale synthetic gas mechanism)."""
    return synthetic_mechanism(72, 133, 28, 19, 5, 5, seed=seed, name="cb05scale72x190")


@dataclass
class CellInputs:
    state: np.ndarray   # (n_cells, n_state) ppm
    env: np.ndarray     # (n_cells, 2) [T K, P Pa]
    rates: dict         # rxn_id -> (n_cells,) base rates for update_data


def synthetic_cells(mech, n_cells, seed=2, cell_offset=0, T=None, P=None,
                    night_fraction=0.0, stressed_fraction=0.0,
                    photolysis_range=(1e-6, 1e-2), emission_range=(1e-7, 1e-4),
                    loss_range=(1e-6, 1e-3), y0_range=(1e-4, 1e2),
                    T_range=(220.0, 310.0), P_range=(2.0e4, 1.013e5)):
    """Per-cell initial state, environment and updatable rates.

    Each cell draws from its own stream (seed, cell index), in this order:
    T, P, y0[n_state], one base rate per updatable reaction (mechanism
    order), a night draw and a stress draw.  Night cells get photolysis
    J = 0; stressed cells get y0 x 1e3."""
    cells = np.arange(cell_offset, cell_offset + n_cells)
    st = CellStreams(seed, cells)
    tt = st.uniform(*T_range)
    pp = st.uniform(*P_range)
    if T is not None:
        tt = np.full(n_cells, float(T))
    if P is not None:
        pp = np.full(n_cells, float(P))
    ns = mech.n_state
    state = np.empty((n_cells, ns))
    for i in range(ns):
        state[:, i] = st.log_uniform(*y0_range)
    rates = {}
    for r in mech.reactions:
        if r.type == "PHOTOLYSIS":
            rates[r.rxn_id] = st.log_uniform(*photolysis_range)
        elif r.type == "EMISSION":
            rates[r.rxn_id] = st.log_uniform(*emission_range)
        elif r.type == "FIRST_ORDER_LOSS":
            rates[r.rxn_id] = st.log_uniform(*loss_range)
    night = st.uniform() < night_fraction
    stressed = st.uniform() < stressed_fraction
    for r in mech.reactions:
        if r.type == "PHOTOLYSIS":
            rates[r.rxn_id] = np.where(night, 0.0, rates[r.rxn_id])
    state[stressed] *= 1.0e3
    env = np.stack([tt, pp], axis=1)
    return CellInputs(np.ascontiguousarray(state), np.ascontiguousarray(env), rates)


def rate_update_payload(mech):
    """[(type_code, rxn_id)] of every updatable reaction, mechanism order."""
    return [(TYPE_CODES[r.type], r.rxn_id) for r in mech.reactions if r.type in UPDATABLE]


def arrhenius_k(A, B, C, D, E, T, P, n_react):
    """Closed form k of rxn_arrhenius.c:112-115 (for tests)."""
    return (A * math.exp(C / T) * (1.0 if B == 0.0 else (T / D) ** B)
            * (1.0 if E == 0.0 else 1.0 + E * P) * (CONV * P / T) ** (n_react - 1))
