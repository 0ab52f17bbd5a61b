"""Shared fixtures: grids, balanced initial states and the oracle driver.

Markers: `gpu` tests need an MI355X (they call the HIP library through the
C ABI); everything else runs on the CPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


_grids = {}


def get_grid(res, orography="flat", moist=0, n_lloyd=0, n_layers=26, n_oro=23):
    from game_archived_amd.grid import Grid
    key = (res, orography, moist, n_lloyd, n_layers, n_oro)
    if key not in _grids:
        _grids[key] = Grid(res, n_layers=n_layers, n_oro_layers=n_oro, moisture_on=moist, n_lloyd=n_lloyd,
                           orography=orography, cache=n_lloyd > 0)
    return _grids[key]


def oracle_state(grid, moist=False, seed=20261202):
    import pyoracle
    from game_archived_amd.state import initial_state
    return initial_state(grid, pyoracle.balance_terms(grid), moist=moist, seed=seed)


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


CONSTITUENT_NAMES = ("snow", "rain", "cloud_ice", "cloud_water", "gas", "vapour")


def state_errors(dims, a, b, fields=("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind")):
    """max |a - b| / max |b| per prognostic field, with the densities compared
    one constituent at a time (each against its own maximum: a condensate
    error is not hidden under the scale of the gas density)"""
    out = {}
    for f in fields:
        x, y = getattr(a, f), getattr(b, f)
        if f == "rho":
            names = CONSTITUENT_NAMES if dims.n_const == 6 else ("gas",)
            for k, name in enumerate(names):
                sl = slice(k * dims.s, (k + 1) * dims.s)
                yk = y[sl]
                # an all-zero constituent (a dry-started condensate) must stay exactly zero
                out[f"rho_{name}"] = rel_err(x[sl], yk) if np.abs(yk).max() > 0 else float(np.abs(x[sl]).max())
        else:
            out[f] = rel_err(x, y)
    return out
