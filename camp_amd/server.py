"""Python front end of the request server (include/camp_server.h).

A CampServer admits chemistry solve requests for one or more initialised
per-cell solvers (mechanisms), runs them with continuous batching on the GPU
and reports per-request latency.  Inputs and outputs are staged in pinned
host memory so that copies overlap with running batches.
"""
import ctypes as C

import numpy as np

from ._lib import lib
from .mechanisms import RXN_EMISSION, RXN_FIRST_ORDER_LOSS, RXN_PHOTOLYSIS


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pinned(shape):
    import torch
    t = torch.empty(shape, dtype=torch.float64, pin_memory=True)
    return t, t.numpy()


def rate_layout(solver):
    """[(rxn_type, unique_id)] in the order a request's base rates follow."""
    L = lib()
    n = L.camp_solver_rate_layout(solver.h, None, None)
    t = np.zeros(max(n, 1), dtype=np.int32)
    u = np.zeros(max(n, 1), dtype=np.int32)
    L.camp_solver_rate_layout(solver.h, t.ctypes.data_as(C.POINTER(C.c_int)),
                              u.ctypes.data_as(C.POINTER(C.c_int)))
    return list(zip(t[:n].tolist(), u[:n].tolist()))


def base_rate_matrix(solver, rates, n_cells):
    """[n_cells][n_rates] from {(type, uid) or uid: per-cell array}, the form
    mechanisms.synthetic_cells returns (keyed by unique id)."""
    lay = rate_layout(solver)
    out = np.zeros((n_cells, max(len(lay), 1)))
    for j, (t, u) in enumerate(lay):
        v = rates.get((t, u), rates.get(u))
        if v is None:
            raise KeyError("no base rate for reaction type %d id %d" % (t, u))
        out[:, j] = np.asarray(v, dtype=np.float64)[:n_cells]
    return out[:, :len(lay)] if lay else out[:, :0]


class Request:
    def __init__(self, server, rid, keep):
        self.server, self.id, self._keep = server, rid, keep
        self.state = keep["out"]
        self.stats = keep["stats"]
        self.latency_s = None
        self.ok = None

    def wait(self):
        if self.ok is None:
            lat = np.zeros(1)
            self.ok = lib().camp_server_wait(self.server.h, self.id, _dp(lat)) == 0
            self.latency_s = float(lat[0])
        return self.ok


class CampServer:
    def __init__(self, n_streams=4, max_batch_cells=1 << 17):
        self.h = lib().camp_server_new(int(n_streams), int(max_batch_cells))
        self.mechs = []

    def add_mechanism(self, solver, max_cells=0):
        m = lib().camp_server_add_mechanism(self.h, solver.h, int(max_cells))
        self.mechs.append(solver)
        return m

    def capacity(self, mech):
        return lib().camp_server_capacity(self.h, int(mech))

    def submit(self, mech, state, env, base_rates, dt, n_steps=1, want_stats=False):
        solver = self.mechs[mech]
        n = state.shape[0]
        keep = {}
        for name, a in (("state", state), ("env", env), ("rates", base_rates)):
            t, v = _pinned(a.shape if a.size else (1,))
            if a.size:
                v[...] = a
            keep[name] = (t, v)
        out_t, out = _pinned((n, solver.ns))
        keep["out_t"], keep["out"] = out_t, out
        keep["stats"] = None
        stats_p = None
        if want_stats:
            st_t, st = _pinned((n, 10))
            keep["stats_t"], keep["stats"] = st_t, st
            stats_p = st.ctypes.data
        rid = lib().camp_server_submit(self.h, int(mech), int(n), keep["state"][1].ctypes.data,
                                       keep["env"][1].ctypes.data, keep["rates"][1].ctypes.data,
                                       float(dt), int(n_steps), out.ctypes.data, stats_p)
        if rid < 0:
            raise ValueError("request rejected (cells beyond the arena, or bad arguments)")
        return Request(self, rid, keep)

    def set_pool(self, mech, state, env, base_rates):
        state = np.ascontiguousarray(state, dtype=np.float64)
        env = np.ascontiguousarray(env, dtype=np.float64)
        br = np.ascontiguousarray(base_rates, dtype=np.float64)
        if lib().camp_server_set_pool(self.h, int(mech), state.shape[0], _dp(state), _dp(env),
                                      _dp(br) if br.size else None) != 0:
            raise ValueError("pool already set or bad arguments")

    def replay(self, t_arrive, mech, n_cells, offset, dt, n_steps):
        n = len(t_arrive)
        t_arrive = np.ascontiguousarray(t_arrive, dtype=np.float64)
        mech = np.ascontiguousarray(mech, dtype=np.int32)
        n_cells = np.ascontiguousarray(n_cells, dtype=np.int32)
        offset = np.ascontiguousarray(offset, dtype=np.int64)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        n_steps = np.ascontiguousarray(n_steps, dtype=np.int32)
        lat = np.zeros(n)
        span = np.zeros(1)
        ip = C.POINTER(C.c_int)
        failed = lib().camp_server_replay(
            self.h, n, _dp(t_arrive), mech.ctypes.data_as(ip), n_cells.ctypes.data_as(ip),
            offset.ctypes.data_as(C.POINTER(C.c_long)), _dp(dt), n_steps.ctypes.data_as(ip),
            _dp(lat), _dp(span))
        if failed < 0:
            raise ValueError("replay rejected a request (pool or mechanism)")
        return lat, float(span[0]), failed

    def counters(self):
        out = np.zeros(6)
        lib().camp_server_counters(self.h, _dp(out))
        keys = ("done", "failed", "cell_steps", "batches", "mean_batch_cells", "peak_resident")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            lib().camp_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["CampServer", "Request", "rate_layout", "base_rate_matrix",
           "RXN_PHOTOLYSIS", "RXN_EMISSION", "RXN_FIRST_ORDER_LOSS"]
