"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the product package never does.  ``Oracle`` drives
oracle/liboracle.so (the C restatement in camp_oracle.c) with the same call
sequence CAMP's Fortran layer uses on the C ABI (camp_solver_data.F90:
612-778): new -> add reactions -> initialize -> update_data -> run.
``RefLib`` drives oracle/_ref/libcampref.so, the reference's own C sources.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libcampref.so")
REF_ROOT = "/root/reference"

_i = C.c_int
_d = C.c_double
_p = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


# CAMP_ORACLE_NATIVE=1 selects liboracle_native.so (the same source built
# -O3 -march=native for this host: the benchmark's CPU baseline)
if os.environ.get("CAMP_ORACLE_NATIVE") == "1":
    LIB_PATH = os.path.join(HERE, "liboracle_native.so")

NATIVE_FLAGS = "-std=c99 -O3 -march=native -fPIC -ffp-contract=off"


def build_native(quiet=True):
    """Compile liboracle_native.so for this host (always rebuilt, since a
    copy built on another machine may not run here)."""
    path = os.path.join(HERE, "liboracle_native.so")
    if os.path.exists(path):
        os.remove(path)
    subprocess.run(["make", "-s", "-C", HERE, "native"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)
    return path


def build(ref=True, quiet=True):
    """Compile liboracle.so, and oracle/_ref when the reference tree exists."""
    targets = ["oracle"]
    if ref and os.path.isdir(os.path.join(REF_ROOT, "src")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def _ip_(a):
    return a.ctypes.data_as(_ip)


def _dp_(a):
    return a.ctypes.data_as(_dp)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.oracle_new.restype = _p
        L.oracle_new.argtypes = [_i, _i, _ip, _i, _i, _i, _i]
        L.oracle_add_rxn.argtypes = [_p, _i, _i, _i, _i, _ip, _dp]
        L.oracle_initialize.argtypes = [_p, _dp, _d, _i, _i]
        L.oracle_free.argtypes = [_p]
        L.oracle_update_data.argtypes = [_p, _i, _ip, _i, _i, _d]
        L.oracle_update_data.restype = _i
        for name in ("oracle_n_dep", "oracle_nnz", "oracle_nnz_rxn", "oracle_n_mapped",
                     "oracle_rxn_int_size", "oracle_n_env", "oracle_lu_nnz"):
            getattr(L, name).argtypes = [_p]
            getattr(L, name).restype = _i
        L.oracle_get_pattern.argtypes = [_p, _ip, _ip, _ip, _ip, _ip]
        L.oracle_get_jac_map.argtypes = [_p, _ip, _ip, _ip]
        L.oracle_get_rxn_int_data.argtypes = [_p, _ip, _i]
        L.oracle_get_rxn_env.argtypes = [_p, _dp]
        L.oracle_update_env_state.argtypes = [_p, _dp]
        L.oracle_eval_f.argtypes = [_p, _i, _dp, _dp, _d, _i]
        L.oracle_eval_f.restype = _i
        L.oracle_eval_jac.argtypes = [_p, _i, _dp, _dp, _d]
        L.oracle_eval_jac.restype = _i
        L.oracle_eval_est.argtypes = [_p, _i, _dp, _dp]
        L.oracle_eval_est.restype = None
        L.oracle_reset_jac_history.argtypes = [_p]
        L.oracle_solver_run.argtypes = [_p, _dp, _dp, _d, _d, _i]
        L.oracle_solver_run.restype = _i
        L.oracle_solver_run_cells.argtypes = [_p, _dp, _dp, _d, _d, _i, _i]
        L.oracle_solver_run_cells.restype = _i
        L.oracle_get_stats.argtypes = [_p, C.POINTER(C.c_long), _dp]
        L.oracle_set_guess_helper.argtypes = [_p, _i]
        L.oracle_lu_solve_sparse.argtypes = [_p, _dp, _d, _dp, _dp]
        L.oracle_lu_solve_sparse.restype = _i
        L.oracle_lu_solve_dense.argtypes = [_p, _dp, _d, _dp, _dp]
        L.oracle_lu_solve_dense.restype = _i
        L.oracle_set_neg_mode.argtypes = [_p, _i]
        L.oracle_kat_pattern.argtypes = [_i, _i, _ip, _ip, _ip, _ip]
        L.oracle_kat_pattern.restype = _i
        L.oracle_kat_jac_accum.argtypes = [_i, _i, _ip, _ip, _dp, _dp]
        L.oracle_kat_time_derivative.argtypes = [_i, _i, _ip, _dp, _dp, _dp]
        L.oracle_set_aero_sizes.argtypes = [_p] + [_i] * 11
        L.oracle_add_aero_phase.argtypes = [_p, _i, _i, _ip, _dp]
        L.oracle_add_aero_rep.argtypes = [_p, _i, _i, _i, _i, _ip, _dp]
        L.oracle_add_sub_model.argtypes = [_p, _i, _i, _i, _i, _ip, _dp]
        L.oracle_param_pattern.argtypes = [_p, _ip, _ip]
        L.oracle_param_pattern.restype = _i
        L.oracle_get_arrays.argtypes = [_p, _i, _p, _i]
        L.oracle_cell_update.argtypes = [_p, _i, _dp]
        L.oracle_rep_get.argtypes = [_p, _i, _i, _dp, _i, _i, _dp, _dp]
        L.oracle_eval_jac_parts.argtypes = [_p, _i, _dp, _dp, _dp]
        L.oracle_eval_jac_parts.restype = _i
        _lib = L
    return _lib


def packed_parts(mech):
    """{'rxn', 'phase', 'rep', 'sub'} packed arrays of a gas or multi-phase
    mechanism, in the order the Fortran layer adds them."""
    if hasattr(mech, "rep"):
        from camp_amd.multiphase import pack_multiphase
        return pack_multiphase(mech)
    from camp_amd.mechanisms import pack
    return {"rxn": pack(mech), "phase": [], "rep": [], "sub": []}


def aero_sizes(parts):
    """solver_new's aerosol/sub-model size arguments for packed parts."""
    ph, rp, sb = parts["phase"], parts["rep"], parts["sub"]
    return [len(ph), sum(len(p[0]) for p in ph), sum(len(p[1]) for p in ph),
            len(rp), sum(len(p[1]) for p in rp), sum(len(p[2]) for p in rp),
            sum(p[3] for p in rp),
            len(sb), sum(len(p[1]) for p in sb), sum(len(p[2]) for p in sb),
            sum(p[3] for p in sb)]


def _arr(v, dt):
    return np.asarray(v if len(v) else [0], dtype=dt)


class Oracle:
    """One oracle model: mechanism packed into the reference layout."""

    def __init__(self, mech, n_cells, abs_tol=None, rel_tol=None,
                 max_steps=10000, max_conv_fails=1000):
        L = lib()
        self.mech = mech
        self.n_cells = n_cells
        self.ns = mech.n_state
        parts = packed_parts(mech)
        packed = parts["rxn"]
        n_int = sum(len(p[1]) for p in packed)
        n_float = sum(len(p[2]) for p in packed)
        n_env = sum(p[3] for p in packed)
        vt = np.asarray(mech.var_type, dtype=np.int32)
        self.h = L.oracle_new(self.ns, n_cells, _ip_(vt), len(packed), n_int, n_float, n_env)
        self.sizes = aero_sizes(parts)
        L.oracle_set_aero_sizes(self.h, *self.sizes)
        for ints, floats in parts["phase"]:
            L.oracle_add_aero_phase(self.h, len(ints), len(floats), _ip_(_arr(ints, np.int32)),
                                    _dp_(_arr(floats, np.float64)))
        for t, ints, floats, ne in parts["rep"]:
            L.oracle_add_aero_rep(self.h, t, len(ints), len(floats), ne,
                                  _ip_(_arr(ints, np.int32)), _dp_(_arr(floats, np.float64)))
        for t, ints, floats, ne in parts["sub"]:
            L.oracle_add_sub_model(self.h, t, len(ints), len(floats), ne,
                                   _ip_(_arr(ints, np.int32)), _dp_(_arr(floats, np.float64)))
        for t, ints, floats, ne in packed:
            ia = np.asarray(ints, dtype=np.int32)
            fa = np.asarray(floats, dtype=np.float64)
            L.oracle_add_rxn(self.h, t, len(ints), len(floats), ne, _ip_(ia), _dp_(fa))
        self.n_float = n_float
        at = np.asarray(mech.abs_tol if abs_tol is None else abs_tol, dtype=np.float64)
        rt = mech.rel_tol if rel_tol is None else rel_tol
        L.oracle_initialize(self.h, _dp_(at), rt, max_steps, max_conv_fails)
        self.nv = L.oracle_n_dep(self.h)
        self.nnz = L.oracle_nnz(self.h)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.oracle_free(self.h)
            self.h = None

    def set_rates(self, rates, cells=None):
        """rates: {rxn_id: array over cells}; uses rxn_update_data per cell."""
        L = lib()
        payload = self.mech
        types = {r.rxn_id: r.type for r in payload.reactions if r.rxn_id > 0}
        from camp_amd.mechanisms import TYPE_CODES
        cells = range(self.n_cells) if cells is None else cells
        rid = C.c_int(0)
        for uid, vals in rates.items():
            code = TYPE_CODES[types[uid]]
            vals = np.broadcast_to(np.asarray(vals, dtype=np.float64), (self.n_cells,))
            for c in cells:
                rid.value = 0
                L.oracle_update_data(self.h, c, C.byref(rid), code, uid, float(vals[c]))

    def pattern(self):
        L = lib()
        nnz_r = L.oracle_nnz_rxn(self.h)
        rcp = np.zeros(self.ns + 1, np.int32)
        rrv = np.zeros(max(nnz_r, 1), np.int32)
        cp = np.zeros(self.nv + 1, np.int32)
        rv = np.zeros(max(self.nnz, 1), np.int32)
        di = np.zeros(self.ns, np.int32)
        L.oracle_get_pattern(self.h, _ip_(rcp), _ip_(rrv), _ip_(cp), _ip_(rv), _ip_(di))
        return dict(rxn_colptr=rcp, rxn_rowval=rrv[:nnz_r], colptr=cp, rowval=rv[:self.nnz],
                    deriv_ids=di)

    def jac_map(self):
        L = lib()
        n = L.oracle_n_mapped(self.h)
        a, b, c = (np.zeros(max(n, 1), np.int32) for _ in range(3))
        L.oracle_get_jac_map(self.h, _ip_(a), _ip_(b), _ip_(c))
        return a[:n], b[:n], c[:n]

    def int_data(self):
        L = lib()
        n = L.oracle_rxn_int_size(self.h)
        a = np.zeros(n, np.int32)
        L.oracle_get_rxn_int_data(self.h, _ip_(a), n)
        return a

    def update_env(self, env):
        env = np.ascontiguousarray(env, dtype=np.float64)
        lib().oracle_update_env_state(self.h, _dp_(env))

    def rxn_env(self):
        L = lib()
        a = np.zeros(self.n_cells * L.oracle_n_env(self.h))
        L.oracle_get_rxn_env(self.h, _dp_(a))
        return a.reshape(self.n_cells, -1)

    def eval_f(self, cell, state_cell, time_step=1.0, use_deriv_est=False):
        s = np.ascontiguousarray(state_cell, dtype=np.float64)
        out = np.zeros(self.nv)
        rc = lib().oracle_eval_f(self.h, cell, _dp_(s), _dp_(out), time_step, int(use_deriv_est))
        return rc, out

    def eval_jac(self, cell, state_cell, time_step=1.0):
        s = np.ascontiguousarray(state_cell, dtype=np.float64)
        out = np.zeros(self.nnz)
        rc = lib().oracle_eval_jac(self.h, cell, _dp_(s), _dp_(out), time_step)
        return rc, out

    def eval_est(self, cell, state_cell):
        """J_deriv + J_solver (y - J_state) from the cell's J history."""
        s = np.ascontiguousarray(state_cell, dtype=np.float64)
        out = np.zeros(self.nv)
        lib().oracle_eval_est(self.h, cell, _dp_(s), _dp_(out))
        return out

    def reset_jac_history(self):
        lib().oracle_reset_jac_history(self.h)

    def run(self, state, env, t0, t1, mode=1):
        """solver_run; state (n_cells, ns) is updated in place."""
        assert state.flags.c_contiguous and state.dtype == np.float64
        env = np.ascontiguousarray(env, dtype=np.float64)
        return lib().oracle_solver_run(self.h, _dp_(state), _dp_(env), t0, t1, mode)

    def run_cells(self, state, env, t0, t1, c0, c1):
        env = np.ascontiguousarray(env, dtype=np.float64)
        return lib().oracle_solver_run_cells(self.h, _dp_(state), _dp_(env), t0, t1, c0, c1)

    def stats(self):
        c = (C.c_long * 8)()
        h = np.zeros(2)
        lib().oracle_get_stats(self.h, c, _dp_(h))
        keys = ("num_steps", "RHS_evals", "LS_setups", "error_test_fails", "NLS_iters",
                "NLS_convergence_fails", "DLS_Jac_evals", "solver_flag")
        d = dict(zip(keys, list(c)))
        d["last_time_step__s"], d["next_time_step__s"] = float(h[0]), float(h[1])
        return d

    def set_guess_helper(self, on):
        lib().oracle_set_guess_helper(self.h, int(on))

    def lu_solve(self, jac, gamma, b, dense=False):
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.nv)
        fn = lib().oracle_lu_solve_dense if dense else lib().oracle_lu_solve_sparse
        rc = fn(self.h, _dp_(jac), gamma, _dp_(b), _dp_(x))
        return rc, x

    def lu_nnz(self):
        return lib().oracle_lu_nnz(self.h)

    # multi-phase introspection
    def param_pattern(self):
        L = lib()
        n = L.oracle_param_pattern(self.h, None, None)
        cp = np.zeros(self.ns + 1, np.int32)
        rv = np.zeros(max(n, 1), np.int32)
        L.oracle_param_pattern(self.h, _ip_(cp), _ip_(rv))
        return cp, rv[:n]

    def arrays(self, which, n):
        """0 rxn float, 1 rep int, 2 rep float, 3 sub int, 4 sub env data."""
        a = np.zeros(n, np.int32 if which in (1, 3) else np.float64)
        lib().oracle_get_arrays(self.h, which, a.ctypes.data, n)
        return a

    def cell_update(self, cell, state_cell):
        s = np.array(state_cell, dtype=np.float64)
        lib().oracle_cell_update(self.h, cell, _dp_(s))
        return s

    def jac_parts(self, cell, state_cell):
        """(J_rxn over the reaction pattern, J_params over the parameter pattern)"""
        s = np.ascontiguousarray(state_cell, dtype=np.float64)
        jr = np.zeros(lib().oracle_nnz_rxn(self.h))
        jp = np.zeros(lib().oracle_param_pattern(self.h, None, None))
        rc = lib().oracle_eval_jac_parts(self.h, cell, _dp_(s), _dp_(jr), _dp_(jp))
        return rc, jr, jp

    def rep_get(self, rep, cell, state_cell, phase, which, n_partial):
        s = np.ascontiguousarray(state_cell, dtype=np.float64)
        v = np.zeros(1)
        part = np.zeros(max(n_partial, 1))
        lib().oracle_rep_get(self.h, rep, cell, _dp_(s), phase, which, _dp_(v), _dp_(part))
        return float(v[0]), part[:n_partial]


_ref = None


def ref_available():
    return os.path.exists(REF_LIB_PATH)


def reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB_PATH)
        L.ref_new.restype = _p
        L.ref_new.argtypes = [_i, _i, _ip, _i, _i, _i, _i]
        L.ref_add_rxn.argtypes = [_p, _i, _i, _i, _i, _ip, _dp]
        L.ref_build_pattern.argtypes = [_p, _ip, _ip, _i]
        L.ref_build_pattern.restype = _i
        L.ref_update_ids.argtypes = [_p, _ip]
        L.ref_get_int_data.argtypes = [_p, _ip, _i]
        L.ref_update_env.argtypes = [_p, _i, _dp]
        L.ref_update_data.argtypes = [_p, _i, _ip, _i, _i, _d]
        L.ref_get_env_data.argtypes = [_p, _dp, _i]
        L.ref_time_derivative.argtypes = [_i, _i, _ip, _dp, _dp, _dp]
        L.ref_set_aero.argtypes = [_p] + [_i] * 11 + [_dp]
        L.ref_add_phase.argtypes = [_p, _i, _i, _ip, _dp]
        L.ref_add_rep.argtypes = [_p, _i, _i, _i, _i, _ip, _dp]
        L.ref_add_sub.argtypes = [_p, _i, _i, _i, _i, _ip, _dp]
        L.ref_param_pattern.argtypes = [_p, _ip, _ip, _ip, _i]
        L.ref_param_pattern.restype = _i
        L.ref_update_env_all.argtypes = [_p, _i, _dp]
        L.ref_cell_update.argtypes = [_p, _i, _dp]
        L.ref_rep_get.argtypes = [_p, _i, _dp, _i, _i, _i, _dp, _dp]
        L.ref_get_arrays.argtypes = [_p, _i, _p, _i]
        _ref = L
    return _ref


class RefModel:
    """The reference's own rxn_* / Jacobian functions on a packed mechanism."""

    def __init__(self, mech, n_cells):
        L = reflib()
        self.mech = mech
        self.n_cells = n_cells
        parts = packed_parts(mech)
        packed = parts["rxn"]
        self.n_int = sum(len(p[1]) for p in packed) + len(packed)
        n_float = sum(len(p[2]) for p in packed)
        self.n_float = n_float
        self.n_env = sum(p[3] for p in packed)
        vt = np.asarray(mech.var_type, dtype=np.int32)
        self.h = L.ref_new(mech.n_state, n_cells, _ip_(vt), len(packed),
                           self.n_int - len(packed), n_float, self.n_env)
        self.sizes = aero_sizes(parts)
        at = np.asarray(mech.abs_tol, dtype=np.float64)
        L.ref_set_aero(self.h, *self.sizes, _dp_(at))
        # the reference's add functions copy from (and may advance) these
        for ints, floats in parts["phase"]:
            L.ref_add_phase(self.h, len(ints), len(floats), _ip_(_arr(ints, np.int32)),
                            _dp_(_arr(floats, np.float64)))
        for t, ints, floats, ne in parts["rep"]:
            L.ref_add_rep(self.h, t, len(ints), len(floats), ne, _ip_(_arr(ints, np.int32)),
                          _dp_(_arr(floats, np.float64)))
        for t, ints, floats, ne in parts["sub"]:
            L.ref_add_sub(self.h, t, len(ints), len(floats), ne, _ip_(_arr(ints, np.int32)),
                          _dp_(_arr(floats, np.float64)))
        for t, ints, floats, ne in packed:
            ia = np.asarray(ints, dtype=np.int32)
            fa = np.asarray(floats if floats else [0.0], dtype=np.float64)
            L.ref_add_rxn(self.h, t, len(ints), len(floats), ne, _ip_(ia), _dp_(fa))

    def pattern(self):
        ns = self.mech.n_state
        cp = np.zeros(ns + 1, np.int32)
        rv = np.zeros(ns * ns, np.int32)
        nnz = reflib().ref_build_pattern(self.h, _ip_(cp), _ip_(rv), ns * ns)
        return cp, rv[:nnz]

    def update_ids(self, deriv_ids):
        di = np.ascontiguousarray(deriv_ids, dtype=np.int32)
        reflib().ref_update_ids(self.h, _ip_(di))
        out = np.zeros(self.n_int, np.int32)
        reflib().ref_get_int_data(self.h, _ip_(out), self.n_int)
        return out

    def update_env(self, cell, T, P):
        e = np.array([T, P], dtype=np.float64)
        reflib().ref_update_env(self.h, cell, _dp_(e))

    def update_data(self, cell, rtype, uid, rate):
        rid = C.c_int(0)
        reflib().ref_update_data(self.h, cell, C.byref(rid), rtype, uid, rate)

    def env_data(self):
        out = np.zeros(self.n_cells * self.n_env)
        reflib().ref_get_env_data(self.h, _dp_(out), out.size)
        return out.reshape(self.n_cells, self.n_env)

    # multi-phase
    def param_pattern(self, deriv_ids):
        ns = self.mech.n_state
        di = np.ascontiguousarray(deriv_ids, dtype=np.int32)
        cp = np.zeros(ns + 1, np.int32)
        rv = np.zeros(ns * ns, np.int32)
        n = reflib().ref_param_pattern(self.h, _ip_(di), _ip_(cp), _ip_(rv), ns * ns)
        return cp, rv[:n]

    def update_env_all(self, cell, T, P):
        e = np.array([T, P], dtype=np.float64)
        reflib().ref_update_env_all(self.h, cell, _dp_(e))

    def cell_update(self, cell, state_cell):
        s = np.array(state_cell, dtype=np.float64)
        reflib().ref_cell_update(self.h, cell, _dp_(s))
        return s

    def rep_get(self, rep, cell, state_cell, phase, which, n_partial):
        s = np.ascontiguousarray(state_cell, dtype=np.float64)
        v = np.zeros(1)
        part = np.zeros(max(n_partial, 1))
        reflib().ref_rep_get(self.h, cell, _dp_(s), rep, phase, which, _dp_(v), _dp_(part))
        return float(v[0]), part[:n_partial]

    def arrays(self, which, n):
        a = np.zeros(n, np.int32 if which in (1, 3) else np.float64)
        reflib().ref_get_arrays(self.h, which, a.ctypes.data, n)
        return a


def ref_time_derivative(n_spec, spec, contrib, est=None):
    spec = np.ascontiguousarray(spec, dtype=np.int32)
    contrib = np.ascontiguousarray(contrib, dtype=np.float64)
    out = np.zeros(n_spec)
    e = None if est is None else np.ascontiguousarray(est, dtype=np.float64)
    reflib().ref_time_derivative(n_spec, len(spec), _ip_(spec), _dp_(contrib),
                                 None if e is None else _dp_(e), _dp_(out))
    return out


def kat_pattern(n, pairs):
    dep = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
    ind = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
    cp = np.zeros(n + 1, np.int32)
    rv = np.zeros(max(len(pairs), 1), np.int32)
    nnz = lib().oracle_kat_pattern(n, len(pairs), _ip_(dep), _ip_(ind), _ip_(cp), _ip_(rv))
    return cp, rv[:nnz]


def kat_jac_accum(nnz, adds):
    elem = np.ascontiguousarray([a[0] for a in adds], dtype=np.int32)
    flag = np.ascontiguousarray([a[1] for a in adds], dtype=np.int32)
    val = np.ascontiguousarray([a[2] for a in adds], dtype=np.float64)
    out = np.zeros(nnz)
    lib().oracle_kat_jac_accum(nnz, len(adds), _ip_(elem), _ip_(flag), _dp_(val), _dp_(out))
    return out


def kat_time_derivative(n_spec, spec, contrib, est=None):
    spec = np.ascontiguousarray(spec, dtype=np.int32)
    contrib = np.ascontiguousarray(contrib, dtype=np.float64)
    out = np.zeros(n_spec)
    e = None if est is None else _dp_(np.ascontiguousarray(est, dtype=np.float64))
    lib().oracle_kat_time_derivative(n_spec, len(spec), _ip_(spec), _dp_(contrib), e, _dp_(out))
    return out
