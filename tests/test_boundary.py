"""The drop-in boundary on the host (no GPU): every C function the
reference's Fortran layer binds with bind(c) for the solver path is exported
by libcamp_mi355x.so, the update-data helpers pack the reference's payload
layouts, and the plugin-data calls that need no device follow the
reference's semantics (sub_model_solver.c:466-541, aero_rep_solver.c:
492-589, aero_rep_modal_binned_mass.c:594-738, rxn_solver.c:708-851)."""
import ctypes as C
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIXTURE = os.path.join(HERE, "golden", "fortran_bindc_symbols.json")


def _symbols():
    return json.load(open(FIXTURE))


def _exported():
    from camp_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_every_fortran_bound_function_is_exported():
    """A Fortran build of the reference links against this library: every
    bind(c) name of camp_solver_data.F90, of the supported types' modules
    (rxn_photolysis/emission/first_order_loss.F90,
    aero_rep_modal_binned_mass.F90) and of the modules of types it does not
    integrate (rxn_wet_deposition.F90, rxn_condensed_phase_photolysis.F90,
    aero_rep_single_particle.F90) is a defined dynamic symbol."""
    exported = _exported()
    sym = _symbols()
    want = sorted({n for part in ("supported", "unsupported_types")
                   for names in sym[part].values() for n in names})
    assert len(want) >= 33
    missing = [n for n in want if n not in exported]
    assert not missing, missing


def test_unsupported_types_payloads_follow_the_reference_layouts():
    """rxn_wet_deposition.c:218-239, rxn_condensed_phase_photolysis.c:
    464-485: {int id; double rate}; aero_rep_single_particle.c:405-433:
    {int aero_rep_id; int update_type = 0 (number); int particle_id; double
    number_conc}."""
    from camp_amd._lib import lib
    L = lib()
    for kind in ("wet_deposition", "condensed_phase_photolysis"):
        p = getattr(L, "rxn_%s_create_rate_update_data" % kind)()
        getattr(L, "rxn_%s_set_rate_update_data" % kind)(p, 5, 1.25)
        assert struct.unpack("<id", C.string_at(p, 12)) == (5, 1.25)
        L.rxn_free_update_data(p)
    p = L.aero_rep_single_particle_create_number_update_data()
    L.aero_rep_single_particle_set_number_update_data__n_m3(p, 3, 7, 2.0e9)
    assert struct.unpack("<iiid", C.string_at(p, 20)) == (3, 0, 7, 2.0e9)
    L.aero_rep_free_update_data(p)


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="reference tree absent")
def test_symbol_fixture_matches_the_reference_sources():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_bindc_symbols
    assert make_bindc_symbols.build() == _symbols()


def test_rate_payloads_pack_id_then_rate():
    """rxn_photolysis.c:262-283 (and emission, first-order loss): malloc'd
    {int id; double rate} with the double right after the int."""
    from camp_amd._lib import lib
    L = lib()
    for kind in ("photolysis", "emission", "first_order_loss"):
        p = getattr(L, "rxn_%s_create_rate_update_data" % kind)()
        getattr(L, "rxn_%s_set_rate_update_data" % kind)(p, 17, 2.5e-3)
        raw = C.string_at(p, 12)
        assert struct.unpack("<id", raw) == (17, 2.5e-3)
        L.rxn_free_update_data(p)


def test_section_payloads_pack_three_ints_then_value():
    """aero_rep_modal_binned_mass.c:677-738: {aero_rep_id, update_type (0 GMD,
    1 GSD), section_id, double value}."""
    from camp_amd._lib import lib
    L = lib()
    for kind, code in (("gmd", 0), ("gsd", 1)):
        p = getattr(L, "aero_rep_modal_binned_mass_create_%s_update_data" % kind)()
        getattr(L, "aero_rep_modal_binned_mass_set_%s_update_data" % kind)(p, 42, 3, 1.7)
        assert struct.unpack("<iiid", C.string_at(p, 20)) == (42, code, 3, 1.7)
        L.aero_rep_free_update_data(p)


def _mp_solver(n_cells=2):
    from camp_amd.core import CampSolver
    from camp_amd.multiphase import multiphase_mechanism
    mp = multiphase_mechanism(1, n_bins=2, n_simpol=2, n_aqeq=2)
    return mp, CampSolver(mp, n_cells)


def test_sub_model_update_data_scans_to_the_end():
    """No supported sub-model type takes update data: the scan advances the
    id past the last sub model (sub_model_solver.c:486-503)."""
    from camp_amd._lib import lib
    mp, s = _mp_solver()
    sid = C.c_int(0)
    lib().sub_model_update_data(1, C.byref(sid), 2, None, s.h)
    assert sid.value == 1  # one ZSR sub model
    s.close()


def test_aero_rep_update_data_skips_other_ids():
    """An update whose id matches no representation leaves every cell alone
    and advances the id to the number of representations."""
    from camp_amd._lib import lib
    L = lib()
    mp, s = _mp_solver()
    p = L.aero_rep_modal_binned_mass_create_gmd_update_data()
    L.aero_rep_modal_binned_mass_set_gmd_update_data(p, 12345, 0, 1e-7)
    rid = C.c_int(0)
    L.aero_rep_update_data(0, C.byref(rid), 2, p, s.h)
    assert rid.value == 1
    rid = C.c_int(0)
    L.aero_rep_update_data(0, C.byref(rid), 9, p, s.h)  # another type: no match
    assert rid.value == 1
    L.aero_rep_free_update_data(p)
    s.close()


def test_gmd_update_of_a_binned_section_exits_like_the_reference():
    """aero_rep_modal_binned_mass.c:609-616: a GMD update for a section that
    is not MODAL prints an error and exits with status 1."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import ctypes as C\n"
            "from camp_amd._lib import lib\n"
            "from camp_amd.core import CampSolver\n"
            "from camp_amd import multiphase as M\n"
            "mp = M.multiphase_mechanism(1, n_bins=2, n_simpol=2, n_aqeq=2)\n"
            "orig = M.pack_binned_rep\n"
            "def rep(*a):\n"
            "    t, ints, fl, ne = orig(*a)\n"
            "    ints = list(ints); ints[3] = 77\n"
            "    return t, ints, fl, ne\n"
            "M.pack_binned_rep = rep\n"
            "s = CampSolver(mp, 1)\n"
            "L = lib()\n"
            "p = L.aero_rep_modal_binned_mass_create_gmd_update_data()\n"
            "L.aero_rep_modal_binned_mass_set_gmd_update_data(p, 77, 0, 1e-7)\n"
            "rid = C.c_int(0)\n"
            "L.aero_rep_update_data(0, C.byref(rid), 2, p, s.h)\n"
            "print('returned')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "non-modal" in r.stdout and "returned" not in r.stdout


def test_print_functions_write_the_reference_headers(capfd):
    from camp_amd._lib import lib
    L = lib()
    mp, s = _mp_solver()
    L.aero_phase_print_data(s.h)
    L.aero_rep_print_data(s.h)
    L.sub_model_print_data(s.h)
    L.rxn_print_data(s.h)
    out = capfd.readouterr().out
    assert "Aerosol Phase 0" in out
    assert "number of aerosol representations: 1" in out
    assert "Modal/binned mass-only aerosol representation" in out
    assert "Sub model 0" in out and "Reaction data" in out
    s.close()


def test_debug_setters_return_success():
    """camp_solver.c:502-533: both CAMP_DEBUG setters return
    CAMP_SOLVER_SUCCESS."""
    from camp_amd._lib import lib
    mp, s = _mp_solver()
    assert lib().solver_set_debug_out(s.h, True) == 0
    assert lib().solver_set_eval_jac(s.h, False) == 0
    s.close()


def test_rxn_update_data_scans_and_resumes():
    """rxn_solver.c:708-763: the scan starts at *rxn_id, stops after the
    first reaction whose unique id matches (leaving *rxn_id on it), and a
    payload with an unknown id scans to the end."""
    from camp_amd.core import CampSolver
    from camp_amd.mechanisms import RXN_PHOTOLYSIS, cb05_scale_mechanism
    from camp_amd._lib import lib
    L = lib()
    mech = cb05_scale_mechanism(1)
    s = CampSolver(mech, 3)
    photo = [r for r in mech.reactions if r.type == "PHOTOLYSIS"]
    p = L.rxn_photolysis_create_rate_update_data()
    for cell, r in ((0, photo[0]), (2, photo[3]), (2, photo[0])):
        L.rxn_photolysis_set_rate_update_data(p, r.rxn_id, 1e-3 * (cell + 1) + r.rxn_id * 1e-6)
        rid = C.c_int(0)
        L.rxn_update_data(cell, C.byref(rid), RXN_PHOTOLYSIS, p, s.h)
        k = mech.reactions.index(r)
        assert rid.value == k
    L.rxn_photolysis_set_rate_update_data(p, 10 ** 6, 1.0)
    rid = C.c_int(0)
    L.rxn_update_data(1, C.byref(rid), RXN_PHOTOLYSIS, p, s.h)
    assert rid.value == len(mech.reactions)
    L.rxn_free_update_data(p)
    s.close()
