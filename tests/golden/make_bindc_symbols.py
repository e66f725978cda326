"""Writes tests/golden/fortran_bindc_symbols.json: the C functions the
reference's Fortran layer declares with bind(c) for the solver path --
src/camp_solver_data.F90 (the solver and plugin-data interface, including
the CAMP_DEBUG-only setters) and the modules of the supported reaction and
aerosol-representation types whose update-data helpers are C functions --
scraped from the Fortran interface blocks.  The modules of types this
solver does not support are listed apart.

Run (where the reference tree is present): python tests/golden/make_bindc_symbols.py
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

SUPPORTED = ["src/camp_solver_data.F90", "src/rxns/rxn_photolysis.F90",
             "src/rxns/rxn_emission.F90", "src/rxns/rxn_first_order_loss.F90",
             "src/aero_reps/aero_rep_modal_binned_mass.F90"]
UNSUPPORTED = ["src/rxns/rxn_wet_deposition.F90", "src/rxns/rxn_condensed_phase_photolysis.F90",
               "src/aero_reps/aero_rep_single_particle.F90"]


def scrape(path):
    """Names of the procedures whose declaration carries bind(c)."""
    text = open(path).read()
    # join Fortran continuation lines, drop comments
    text = re.sub(r"!.*", "", text)
    text = re.sub(r"&\s*\n\s*&?", " ", text)
    names = []
    for line in text.splitlines():
        m = re.search(r"\b(?:function|subroutine)\s+([a-z_0-9]+)\s*\(.*\bbind\s*\(\s*c\s*\)",
                      line, flags=re.I)
        if m and not re.match(r"\s*end\b", line, flags=re.I):
            names.append(m.group(1).lower())
    return sorted(set(names))


def build(ref=REF):
    return {"supported": {f: scrape(os.path.join(ref, f)) for f in SUPPORTED},
            "unsupported_types": {f: scrape(os.path.join(ref, f)) for f in UNSUPPORTED}}


def main():
    if not os.path.isdir(os.path.join(REF, "src")):
        sys.exit("reference tree not found at " + REF)
    with open(os.path.join(HERE, "fortran_bindc_symbols.json"), "w") as f:
        json.dump(build(), f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
