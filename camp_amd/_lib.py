"""ctypes binding of libcamp_mi355x.so (the HIP solver's C ABI).

The library is built in-tree by camp_amd/build.py (hipcc, gfx950).  There is
no fallback: if the library is missing or fails to load, every entry point
raises, so nothing can silently run on the CPU instead of the GPU.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcamp_mi355x.so")

_i, _d, _p = C.c_int, C.c_double, C.c_void_p
_ip, _dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
_l, _lp = C.c_long, C.POINTER(C.c_long)

# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    "solver_new": (_p, [_i, _i, _ip, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i]),
    "solver_initialize": (None, [_p, _dp, _d, _i, _i]),
    "solver_run": (_i, [_p, _dp, _dp, _d, _d]),
    "solver_get_statistics": (None, [_p, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp,
                                     _ip, _ip, _ip, _dp, _dp, _dp]),
    "solver_reset_timers": (None, [_p]),
    "solver_free": (None, [_p]),
    "rxn_add_condensed_data": (None, [_i, _i, _i, _i, _ip, _dp, _p]),
    "rxn_update_data": (None, [_i, _ip, _i, _p, _p]),
    "rxn_print_data": (None, [_p]),
    "rxn_free_update_data": (None, [_p]),
    "rxn_photolysis_create_rate_update_data": (_p, []),
    "rxn_photolysis_set_rate_update_data": (None, [_p, _i, _d]),
    "rxn_emission_create_rate_update_data": (_p, []),
    "rxn_emission_set_rate_update_data": (None, [_p, _i, _d]),
    "rxn_first_order_loss_create_rate_update_data": (_p, []),
    "rxn_first_order_loss_set_rate_update_data": (None, [_p, _i, _d]),
    "aero_phase_add_condensed_data": (None, [_i, _i, _ip, _dp, _p]),
    "aero_rep_add_condensed_data": (None, [_i, _i, _i, _i, _ip, _dp, _p]),
    "sub_model_add_condensed_data": (None, [_i, _i, _i, _i, _ip, _dp, _p]),
    "sub_model_update_data": (None, [_i, _ip, _i, _p, _p]),
    "sub_model_print_data": (None, [_p]),
    "aero_phase_print_data": (None, [_p]),
    "aero_rep_update_data": (None, [_i, _ip, _i, _p, _p]),
    "aero_rep_print_data": (None, [_p]),
    "aero_rep_modal_binned_mass_create_gmd_update_data": (_p, []),
    "aero_rep_modal_binned_mass_set_gmd_update_data": (None, [_p, _i, _i, _d]),
    "aero_rep_modal_binned_mass_create_gsd_update_data": (_p, []),
    "aero_rep_modal_binned_mass_set_gsd_update_data": (None, [_p, _i, _i, _d]),
    "aero_rep_free_update_data": (None, [_p]),
    "solver_set_debug_out": (_i, [_p, C.c_bool]),
    "solver_set_eval_jac": (_i, [_p, C.c_bool]),
    "camp_solver_set_mode": (_i, [_p, _i]),
    "camp_solver_set_cells_per_block": (_i, [_p, _i]),
    "camp_solver_set_guess_helper": (None, [_p, _i]),
    "camp_solver_set_cell_order": (_i, [_p, _ip, _i]),
    "camp_solver_set_poison": (None, [_p, _i]),
    "camp_solver_run_device": (_i, [_p, _p, _p, _d, _d]),
    "camp_solver_set_base_rates": (_i, [_p, _i, _i, _dp]),
    "camp_solver_dims": (_i, [_p, _ip, _ip, _ip, _ip]),
    "camp_solver_get_pattern": (None, [_p, _ip, _ip, _ip, _ip, _ip, _ip]),
    "camp_eval_deriv": (_i, [_p, _dp, _dp, _d, _i, _dp]),
    "camp_eval_jac": (_i, [_p, _dp, _dp, _d, _dp]),
    "camp_eval_lu_solve": (_i, [_p, _dp, _d, _dp]),
    "camp_solver_check_jac": (_i, [_p, _dp, _dp, _d, _dp, _ip]),
    "camp_solver_set_placement": (_i, [_p, _i]),
    "camp_solver_set_eval_kernels": (_i, [_p, _i]),
    "camp_solver_tune_fj": (_i, [_p, _p, _p, _i, _dp]),
    "camp_solver_set_fj_lanes": (_i, [_p, _i, _i]),
    "camp_solver_set_graphs": (_i, [_p, _i]),
    "camp_solver_cell_lds_bytes": (_l, [_p, _i]),
    "camp_solver_run_device_noalloc": (_i, [_p, _p, _p, _d, _d]),
    "rxn_wet_deposition_create_rate_update_data": (_p, []),
    "rxn_wet_deposition_set_rate_update_data": (None, [_p, _i, _d]),
    "rxn_condensed_phase_photolysis_create_rate_update_data": (_p, []),
    "rxn_condensed_phase_photolysis_set_rate_update_data": (None, [_p, _i, _d]),
    "aero_rep_single_particle_create_number_update_data": (_p, []),
    "aero_rep_single_particle_set_number_update_data__n_m3": (None, [_p, _i, _i, _d]),
    "camp_solver_get_cell_stats": (None, [_p, _dp]),
    "camp_solver_last_kernel_ms": (_d, [_p]),
    "camp_solver_set_profiling": (None, [_p, _i]),
    "camp_solver_get_cell_profile": (_i, [_p, _dp]),
    "camp_solver_get_rxn_env": (None, [_p, _dp]),
    "camp_cr_math": (None, [_i, _i, _dp, _dp, _dp]),
    "camp_solver_autotune": (_i, [_p, _p, _p, _d, _d, _i, _dp]),
    "camp_solver_phase_cycles": (_i, [_p, _p, _p, _i, _i, _dp]),
    "camp_solver_time_fj": (_i, [_p, _p, _p, _i, _dp, _dp]),
    "camp_solver_work_model": (None, [_p, _dp]),
    # include/camp_server.h
    "camp_server_new": (_p, [_i, _l]),
    "camp_hostmodel_new": (_p, [_p, _i, _i, _i, _i, _i, _i, _ip, _ip, _i]),
    "camp_hostmodel_integrate": (_i, [_p, _d, _i, _i, _i, _i, _p, _p, _p, _i, _p, _p]),
    "camp_hostmodel_set_state": (None, [_p, _dp]),
    "camp_hostmodel_gather": (None, [_p, _i, _i, _i, _i, _p, _p, _p, _i, _p, _p, _dp, _dp]),
    "camp_hostmodel_free": (None, [_p]),
    "camp_server_add_mechanism": (_i, [_p, _p, _l]),
    "camp_server_capacity": (_l, [_p, _i]),
    "camp_solver_rate_layout": (_i, [_p, _ip, _ip]),
    "camp_server_submit": (_l, [_p, _i, _i, _p, _p, _p, _d, _i, _p, _p]),
    "camp_server_wait": (_i, [_p, _l, _dp]),
    "camp_server_set_pool": (_i, [_p, _i, _l, _dp, _dp, _dp]),
    "camp_server_replay": (_i, [_p, _i, _dp, _ip, _ip, _lp, _dp, _ip, _dp, _dp]),
    "camp_server_counters": (None, [_p, _dp]),
    "camp_server_free": (None, [_p]),
}

_lib = None


def lib():
    """Load libcamp_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcamp_mi355x.so is not built: run __graft_entry__.build() "
                               "or python camp_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
