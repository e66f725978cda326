"""Shared helpers for the parity tests."""
import numpy as np

from camp_amd.mechanisms import box_mechanism, cb05_scale_mechanism, synthetic_cells


def cb05_inputs(n_cells, seed_cells=2, night_fraction=0.0, stressed_fraction=0.0):
    mech = cb05_scale_mechanism(1)
    cells = synthetic_cells(mech, n_cells, seed=seed_cells, night_fraction=night_fraction,
                            stressed_fraction=stressed_fraction)
    return mech, cells


def box_inputs(n_cells=1):
    mech = box_mechanism(1)
    cells = synthetic_cells(mech, n_cells, seed=2, T=298.0, P=101325.0)
    return mech, cells


def rel_err(a, b, atol):
    """max |a-b| / (|b| + atol)"""
    return float(np.max(np.abs(a - b) / (np.abs(b) + atol)))


def close_mask(a, b, rtol, atol):
    return np.abs(a - b) <= rtol * np.abs(b) + atol
