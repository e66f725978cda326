"""Host-model interface: a grid model's tracer arrays through the GPU solver.

``MonarchInterface`` is the counterpart of the reference's
monarch_interface_t (test/monarch/camp_monarch_interface.F90): it loads a
CAMP configuration and an interface configuration (species map, gas-phase
water species, initial concentrations; :513-606), builds the species map
(create_map, :613-713), fills a mock model's arrays with the initial
concentrations (get_init_conc, :853-893) and integrates the model's cells
(integrate, :343-509) through libcamp_mi355x.so (include/camp_host_model.h):
gather, per-cell solve and scatter all run on the GPU.

Arrays are numpy float32 in Fortran order, shaped as the reference's:
temperature (ni, nj, nk), conc (ni, nj, nk, n_tracers), water_conc (ni, nj,
nk, n_water), air_density and pressure (ni, nk, nj).
"""
import ctypes as C
import json
import os

import numpy as np

from ._lib import lib
from .core import CampSolver
from .mechanisms import load_camp_json_file


def _f32(a):
    if a.dtype != np.float32 or not a.flags.f_contiguous:
        raise ValueError("host-model arrays are float32 in Fortran order")
    return C.c_void_p(a.ctypes.data)


def load_interface_config(path):
    """The interface file's SPECIES_MAP and INIT_CONC objects
    (camp_monarch_interface.F90:513-606): (gas-phase water species name,
    [(species, monarch id)], {species: init conc})."""
    with open(path) as f:
        data = json.load(f)["monarch-data"]
    water, smap, init = None, [], {}
    for obj in data:
        if obj["type"] == "SPECIES_MAP":
            water = obj["gas-phase water"]
            smap += [(k, int(v["monarch id"])) for k, v in obj["gas-phase species"].items()]
            if "aerosol-phase species" in obj:
                raise ValueError("aerosol-phase species maps need an aerosol representation")
        elif obj["type"] == "INIT_CONC":
            init.update({k: float(v["init conc"])
                         for k, v in obj.get("gas-phase species", {}).items()})
    if water is None:
        raise ValueError("Missing gas-phase water species for MONARCH interface.")
    return water, smap, init


class MonarchInterface:
    def __init__(self, camp_config, interface_config, tracer_start, tracer_end, n_cells,
                 grid, n_water, **solver_kw):
        """grid = (ni, nj, nk, n_tracers) of the host model's arrays."""
        self.mech = load_camp_json_file(camp_config)
        water, smap, init = load_interface_config(interface_config)
        idx = {s: i for i, s in enumerate(self.mech.species)}
        for s in [water] + [s for s, _ in smap] + list(init):
            if s not in idx:
                raise ValueError("Could not find species '%s' in PartMC-camp." % s)
        ni, nj, nk, nt = grid
        self.grid = grid
        self.map_monarch_id = np.array([mid + tracer_start - 1 for _, mid in smap], np.int32)
        if np.any(self.map_monarch_id > tracer_end):
            raise ValueError("Monarch id out of specified tracer array bounds.")
        self.map_camp_id = np.array([idx[s] + 1 for s, _ in smap], np.int32)
        self.gas_phase_water_id = idx[water] + 1
        self.init_conc = init
        self.solver = CampSolver(self.mech, n_cells, **solver_kw)
        self.h = lib().camp_hostmodel_new(
            C.c_void_p(self.solver.h), ni, nj, nk, nt, n_water, len(smap),
            self.map_monarch_id.ctypes.data_as(C.POINTER(C.c_int)),
            self.map_camp_id.ctypes.data_as(C.POINTER(C.c_int)), self.gas_phase_water_id)

    def get_init_conc(self, conc, water_conc, water_vapor_index, air_density):
        """The mock model's initial arrays (camp_monarch_interface.F90:853-893):
        every cell gets the initial concentration of each mapped species, the
        air density 1.225 and the water vapour of the water species' initial
        value."""
        state = np.zeros(self.mech.n_state)
        for s, v in self.init_conc.items():
            state[self.mech.species.index(s)] = v
        conc[...] = 0.0
        water_conc[..., water_vapor_index - 1] = 0.0
        air_density[...] = np.float32(1.225)
        for mid, cid in zip(self.map_monarch_id, self.map_camp_id):
            conc[..., mid - 1] = np.float32(state[cid - 1])
        water_conc[..., water_vapor_index - 1] = np.float32(
            state[self.gas_phase_water_id - 1] * 1.0e-9 / 1.225)

    def integrate(self, time_step, i_start, i_end, j_start, j_end, temperature, conc, water_conc,
                  water_vapor_index, air_density, pressure):
        return lib().camp_hostmodel_integrate(
            self.h, float(time_step), i_start, i_end, j_start, j_end, _f32(temperature),
            _f32(conc), _f32(water_conc), int(water_vapor_index), _f32(air_density),
            _f32(pressure))

    def gather(self, i_start, i_end, j_start, j_end, temperature, conc, water_conc,
               water_vapor_index, air_density, pressure):
        """Test hook: the gathered CAMP state and env of the range."""
        n = (i_end - i_start + 1) * (j_end - j_start + 1) * self.grid[2]
        st = np.zeros((n, self.mech.n_state))
        env = np.zeros((n, 2))
        dp = C.POINTER(C.c_double)
        lib().camp_hostmodel_gather(self.h, i_start, i_end, j_start, j_end, _f32(temperature),
                                    _f32(conc), _f32(water_conc), int(water_vapor_index),
                                    _f32(air_density), _f32(pressure), st.ctypes.data_as(dp),
                                    env.ctypes.data_as(dp))
        return st, env

    def set_state(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        lib().camp_hostmodel_set_state(self.h, state.ctypes.data_as(C.POINTER(C.c_double)))

    def close(self):
        if self.h:
            lib().camp_hostmodel_free(self.h)
            self.h = None
        self.solver.close()
