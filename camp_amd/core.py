"""Host-side mirror of CAMP's Fortran solver layer, above the C ABI.

``CampSolver`` replays the call sequence of camp_solver_data_t
(src/camp_solver_data.F90): ``initialize`` = solver_new -> one
rxn_add_condensed_data per reaction -> solver_initialize (:612-778);
``update_rxn_data`` = rxn_update_data with the rate payload helpers
(:812-827, rxn_photolysis.F90:135-155); ``solve`` = solver_run (:901-907);
``stats`` = solver_get_statistics (:943-961).  ``CampCore`` is the thin
camp_core_t equivalent: load a mechanism, keep the per-cell state and env,
and advance by dt from t = 0 (camp_core.F90:1354-1414).
"""
import ctypes as C

import numpy as np

from ._lib import lib
from .mechanisms import (DEFAULT_MAX_CONV_FAILS, DEFAULT_MAX_STEPS, TYPE_CODES, pack)

MODE_PER_CELL = 0
MODE_BLOCK = 1

_i, _d = C.c_int, C.c_double


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


_PAYLOAD = {
    "PHOTOLYSIS": ("rxn_photolysis_create_rate_update_data", "rxn_photolysis_set_rate_update_data"),
    "EMISSION": ("rxn_emission_create_rate_update_data", "rxn_emission_set_rate_update_data"),
    "FIRST_ORDER_LOSS": ("rxn_first_order_loss_create_rate_update_data",
                         "rxn_first_order_loss_set_rate_update_data"),
}


def packed_parts(mech):
    """{'rxn', 'phase', 'rep', 'sub'}: packed arrays of a gas-phase or
    multi-phase mechanism in the reference layouts."""
    if hasattr(mech, "rep"):
        from .multiphase import pack_multiphase
        return pack_multiphase(mech)
    return {"rxn": pack(mech), "phase": [], "rep": [], "sub": []}


class CampSolver:
    """camp_solver_data_t over libcamp_mi355x.so."""

    def __init__(self, mech, n_cells, rel_tol=None, abs_tol=None,
                 max_steps=DEFAULT_MAX_STEPS, max_conv_fails=DEFAULT_MAX_CONV_FAILS,
                 mode=MODE_PER_CELL):
        L = lib()
        self.mech = mech
        self.n_cells = int(n_cells)
        self.ns = mech.n_state
        parts = packed_parts(mech)
        packed = parts["rxn"]
        self._n_int = sum(len(p[1]) + 1 for p in packed)
        n_int = sum(len(p[1]) for p in packed)
        n_float = sum(len(p[2]) for p in packed)
        n_env = sum(p[3] for p in packed)
        vt = np.ascontiguousarray(mech.var_type, dtype=np.int32)
        ph, rp, sb = parts["phase"], parts["rep"], parts["sub"]
        sizes = [len(ph), sum(len(p[0]) for p in ph), sum(len(p[1]) for p in ph),
                 len(rp), sum(len(p[1]) for p in rp), sum(len(p[2]) for p in rp),
                 sum(p[3] for p in rp),
                 len(sb), sum(len(p[1]) for p in sb), sum(len(p[2]) for p in sb),
                 sum(p[3] for p in sb)]
        self.h = L.solver_new(self.ns, self.n_cells, _ip(vt), len(packed), n_int, n_float, n_env,
                              *sizes)
        # the order camp_solver_data adds them (camp_solver_data.F90:662-762)
        for t, ints, floats, ne in packed:
            ia = np.ascontiguousarray(ints, dtype=np.int32)
            fa = np.ascontiguousarray(floats if floats else [0.0], dtype=np.float64)
            L.rxn_add_condensed_data(t, len(ints), len(floats), ne, _ip(ia), _dp(fa), self.h)
        for ints, floats in ph:
            ia = np.ascontiguousarray(ints, dtype=np.int32)
            fa = np.ascontiguousarray(floats if floats else [0.0], dtype=np.float64)
            L.aero_phase_add_condensed_data(len(ints), len(floats), _ip(ia), _dp(fa), self.h)
        for t, ints, floats, ne in rp:
            ia = np.ascontiguousarray(ints, dtype=np.int32)
            fa = np.ascontiguousarray(floats if floats else [0.0], dtype=np.float64)
            L.aero_rep_add_condensed_data(t, len(ints), len(floats), ne, _ip(ia), _dp(fa), self.h)
        for t, ints, floats, ne in sb:
            ia = np.ascontiguousarray(ints, dtype=np.int32)
            fa = np.ascontiguousarray(floats if floats else [0.0], dtype=np.float64)
            L.sub_model_add_condensed_data(t, len(ints), len(floats), ne, _ip(ia), _dp(fa),
                                           self.h)
        at = np.ascontiguousarray(mech.abs_tol if abs_tol is None else abs_tol, dtype=np.float64)
        rt = mech.rel_tol if rel_tol is None else rel_tol
        L.solver_initialize(self.h, _dp(at), rt, max_steps, max_conv_fails)
        self.set_mode(mode)
        nv, nnz, lnz, nr = (C.c_int() for _ in range(4))
        L.camp_solver_dims(self.h, C.byref(nv), C.byref(nnz), C.byref(lnz), C.byref(nr))
        self.nv, self.nnz, self.lu_nnz, self.nnz_rxn = nv.value, nnz.value, lnz.value, nr.value
        self._types = {r.rxn_id: r.type for r in mech.reactions if r.rxn_id > 0}

    def close(self):
        if getattr(self, "h", None):
            lib().solver_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, mode):
        return lib().camp_solver_set_mode(self.h, mode)

    def set_graphs(self, on):
        """Block mode: Newton iterations replayed as HIP graphs (default)."""
        return lib().camp_solver_set_graphs(self.h, int(bool(on)))

    def set_cells_per_block(self, n):
        return lib().camp_solver_set_cells_per_block(self.h, n)

    def set_placement(self, place):
        """0: cell history in LDS; 1: in global memory (more cells per CU);
        2: history and LU factors in global memory (more still)."""
        return lib().camp_solver_set_placement(self.h, int(place))

    def autotune(self, state_ptr, env_ptr, t0, t1, n_sample=0):
        """Pick the per-cell kernel's cell placement and cells per workgroup
        by timing each candidate on a sample of the device-resident cells
        (copies).  Returns ((placement, cells per workgroup),
        {(placement, cells per workgroup): ms or None when it does not fit})."""
        ms = np.zeros(18)
        best = lib().camp_solver_autotune(self.h, C.c_void_p(state_ptr), C.c_void_p(env_ptr),
                                          float(t0), float(t1), int(n_sample), _dp(ms))
        cands = [(p, c) for p in (0, 1, 2) for c in (1, 2, 4, 8, 16, 32)]
        return (best // 100, best % 100), {k: (float(v) if v >= 0 else None)
                                           for k, v in zip(cands, ms)}

    def phase_cycles(self, state_ptr, env_ptr, n_sample=0, reps=4):
        """Cycles of f, J, LU and one solve in the per-cell kernel (dict), or
        None where not measurable (gas-phase programs at placement 2)."""
        out = np.zeros(4)
        rc = lib().camp_solver_phase_cycles(self.h, C.c_void_p(state_ptr), C.c_void_p(env_ptr),
                                            int(n_sample), int(reps), _dp(out))
        return None if rc else dict(zip(("f", "jac", "lu", "solve"), out.tolist()))

    def set_guess_helper(self, on):
        lib().camp_solver_set_guess_helper(self.h, int(bool(on)))

    def set_cell_order(self, order=None):
        """Queue order of the per-cell kernel (a permutation of the cells;
        None: the default, longest first by the previous call's step
        counts).  Results do not depend on it."""
        if order is None:
            rc = lib().camp_solver_set_cell_order(self.h, None, 0)
        else:
            o = np.ascontiguousarray(order, dtype=np.int32)
            rc = lib().camp_solver_set_cell_order(self.h, _ip(o), len(o))
        if rc:
            raise ValueError("cell order must be a permutation of the %d cells" % self.n_cells)

    def set_poison(self, on=True):
        """Fill each wavefront's working memory with NaNs before every cell
        (a check that no cell reads what an earlier one left)."""
        lib().camp_solver_set_poison(self.h, int(bool(on)))

    def update_rxn_data(self, cell, rxn_unique_id, rate):
        """One rxn_update_data call with a freshly packed payload."""
        L = lib()
        rtype = self._types[rxn_unique_id]
        create, setter = _PAYLOAD[rtype]
        payload = getattr(L, create)()
        getattr(L, setter)(payload, rxn_unique_id, float(rate))
        rid = C.c_int(0)
        L.rxn_update_data(int(cell), C.byref(rid), TYPE_CODES[rtype], payload, self.h)
        L.rxn_free_update_data(payload)

    def set_rates(self, rates):
        """rates: {unique id: scalar or array over cells} (bulk update)."""
        L = lib()
        for uid, vals in rates.items():
            v = np.ascontiguousarray(np.broadcast_to(np.asarray(vals, np.float64),
                                                     (self.n_cells,)))
            rc = L.camp_solver_set_base_rates(self.h, TYPE_CODES[self._types[uid]], int(uid),
                                              _dp(v))
            if rc:
                raise RuntimeError("rate update for reaction id %d did not match" % uid)

    def solve(self, state, env, t0, t1):
        """solver_run on host arrays; state (n_cells, ns) updated in place."""
        assert state.dtype == np.float64 and state.flags.c_contiguous
        assert state.shape == (self.n_cells, self.ns)
        env = np.ascontiguousarray(env, dtype=np.float64)
        return lib().solver_run(self.h, _dp(state), _dp(env), float(t0), float(t1))

    def solve_device(self, state_ptr, env_ptr, t0, t1):
        """solver run on device pointers (e.g. torch tensor data_ptr())."""
        return lib().camp_solver_run_device(self.h, C.c_void_p(state_ptr), C.c_void_p(env_ptr),
                                            float(t0), float(t1))

    def stats(self):
        ints = [C.c_int() for _ in range(9)]
        dbl = [C.c_double() for _ in range(2)]
        more = [C.c_int() for _ in range(3)]
        tail = [C.c_double() for _ in range(3)]
        lib().solver_get_statistics(self.h, *[C.byref(x) for x in ints],
                                    *[C.byref(x) for x in dbl], *[C.byref(x) for x in more],
                                    *[C.byref(x) for x in tail])
        keys = ("solver_flag", "num_steps", "RHS_evals", "LS_setups", "error_test_fails",
                "NLS_iters", "NLS_convergence_fails", "DLS_Jac_evals", "DLS_RHS_evals")
        d = {k: v.value for k, v in zip(keys, ints)}
        d["last_time_step__s"], d["next_time_step__s"] = dbl[0].value, dbl[1].value
        for k, v in zip(("Jac_eval_fails", "RHS_evals_total", "Jac_evals_total"), more):
            d[k] = v.value
        for k, v in zip(("RHS_time__s", "Jac_time__s", "max_loss_precision"), tail):
            d[k] = v.value
        return d

    full_stats = stats

    def cell_stats(self):
        out = np.zeros((self.n_cells, 10))
        lib().camp_solver_get_cell_stats(self.h, _dp(out))
        return out

    def rxn_env(self):
        n_env = sum(p[3] for p in packed_parts(self.mech)["rxn"])
        out = np.zeros((self.n_cells, n_env))
        lib().camp_solver_get_rxn_env(self.h, _dp(out))
        return out

    def last_kernel_ms(self):
        return lib().camp_solver_last_kernel_ms(self.h)

    def set_profiling(self, on=True):
        lib().camp_solver_set_profiling(self.h, int(bool(on)))

    def cell_profile(self):
        """Shader cycles per cell: [f, Jac, LU factor, LU solve, whole solve,
        f's state update + rates, f's gather + estimate blend]."""
        out = np.zeros((self.n_cells, 7))
        if lib().camp_solver_get_cell_profile(self.h, _dp(out)) != 0:
            raise RuntimeError("no profiled per-cell run yet")
        return out

    def work_model(self):
        """Per-cell algorithmic flops and HBM bytes of f, J, LU, solve."""
        out = np.zeros(8)
        lib().camp_solver_work_model(self.h, _dp(out))
        keys = ("f", "jac", "lu", "solve")
        return ({k: float(v) for k, v in zip(keys, out[:4])},
                {k: float(v) for k, v in zip(keys, out[4:])})

    def time_fj(self, state_ptr, env_ptr, reps=10):
        """Average ms of the plain-path f(y) and J(y) kernels (device data)."""
        a, b = C.c_double(), C.c_double()
        lib().camp_solver_time_fj(self.h, C.c_void_p(state_ptr), C.c_void_p(env_ptr), int(reps),
                                  C.byref(a), C.byref(b))
        return a.value, b.value

    def set_fj_lanes(self, f_lanes, j_lanes):
        """Lanes per cell (8, 16, 32) of the block path's f and J kernels."""
        return lib().camp_solver_set_fj_lanes(self.h, int(f_lanes), int(j_lanes))

    def tune_fj(self, state_ptr, env_ptr, reps=10):
        """Time the block path's f/J kernels at 8, 16 and 32 lanes per cell
        and keep the fastest.  Returns ((f lanes, J lanes),
        {"f": {lanes: ms}, "jac": {lanes: ms}})."""
        ms = np.zeros(6)
        best = lib().camp_solver_tune_fj(self.h, C.c_void_p(state_ptr), C.c_void_p(env_ptr),
                                         int(reps), _dp(ms))
        w = (8, 16, 32)
        return (best % 100, best // 100), {"f": dict(zip(w, ms[:3].tolist())),
                                           "jac": dict(zip(w, ms[3:].tolist()))}

    def pattern(self):
        L = lib()
        rcp = np.zeros(self.ns + 1, np.int32)
        rrv = np.zeros(max(self.nnz_rxn, 1), np.int32)
        cp = np.zeros(self.nv + 1, np.int32)
        rv = np.zeros(max(self.nnz, 1), np.int32)
        di = np.zeros(self.ns, np.int32)
        ri = np.zeros(max(self._n_int, 1), np.int32)
        L.camp_solver_get_pattern(self.h, _ip(rcp), _ip(rrv), _ip(cp), _ip(rv), _ip(di), _ip(ri))
        return dict(rxn_colptr=rcp, rxn_rowval=rrv[:self.nnz_rxn], colptr=cp,
                    rowval=rv[:self.nnz], deriv_ids=di, rxn_int_data=ri[:self._n_int])

    def eval_deriv(self, state, env, time_step=1.0, use_deriv_est=False):
        state = np.ascontiguousarray(state, dtype=np.float64)
        env = np.ascontiguousarray(env, dtype=np.float64)
        out = np.zeros((self.n_cells, self.nv))
        rc = lib().camp_eval_deriv(self.h, _dp(state), _dp(env), float(time_step),
                                   int(use_deriv_est), _dp(out))
        return rc, out

    def eval_jac(self, state, env, time_step=1.0):
        state = np.ascontiguousarray(state, dtype=np.float64)
        env = np.ascontiguousarray(env, dtype=np.float64)
        out = np.zeros((self.n_cells, self.nnz))
        rc = lib().camp_eval_jac(self.h, _dp(state), _dp(env), float(time_step), _dp(out))
        return rc, out

    def check_jac(self, state, env, time_step=1.0, jac=None):
        """Finite-difference check of J on the GPU (camp_solver_check_jac):
        (total failing elements or -1, failing elements per cell)."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        env = np.ascontiguousarray(env, dtype=np.float64)
        fails = np.zeros(self.n_cells, np.int32)
        jp = None
        if jac is not None:
            jac = np.ascontiguousarray(jac, dtype=np.float64)
            jp = _dp(jac)
        rc = lib().camp_solver_check_jac(self.h, _dp(state), _dp(env), float(time_step), jp,
                                         _ip(fails))
        return rc, fails

    def set_eval_kernels(self, group):
        """camp_eval_deriv / camp_eval_jac on the lane-group kernels (1) or
        the per-wavefront code of the per-cell kernel (0)."""
        return lib().camp_solver_set_eval_kernels(self.h, int(group))

    def set_eval_jac(self, on=True):
        """solver_set_eval_jac: check every cell's last Jacobian after each
        solve; failures are counted as Jac_eval_fails."""
        lib().solver_set_eval_jac(self.h, bool(on))

    def eval_lu_solve(self, jac, gamma, b):
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        x = np.ascontiguousarray(b, dtype=np.float64).copy()
        rc = lib().camp_eval_lu_solve(self.h, _dp(jac), float(gamma), _dp(x))
        return rc, x


class CampCore:
    """camp_core_t equivalent: a mechanism, per-cell state/env and solve(dt)."""

    def __init__(self, mech, n_cells=1, **solver_kw):
        self.mech = mech
        self.n_cells = n_cells
        self.solver = CampSolver(mech, n_cells, **solver_kw)

    def new_state(self):
        st = np.zeros((self.n_cells, self.mech.n_state))
        env = np.tile([298.15, 101325.0], (self.n_cells, 1))
        return st, env

    def update_data(self, rxn_unique_id, rate, cell=None):
        cells = range(self.n_cells) if cell is None else [cell]
        for c in cells:
            self.solver.update_rxn_data(c, rxn_unique_id, rate)

    def solve(self, state, env, dt):
        """camp_core%solve: integrate from t = 0 to dt (camp_core.F90:1408)."""
        return self.solver.solve(state, env, 0.0, dt)
