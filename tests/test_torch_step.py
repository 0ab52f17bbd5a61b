"""The PyTorch restatement of the dry step (tests/torch_step.py) against the
C oracle: its forward must be the oracle's up to rounding order, so that its
autograd products are the derivatives of the oracle's (and the device's) map."""
import numpy as np
import pytest
import torch

import pyoracle
import torch_step as TS
from conftest import get_grid, oracle_state


@pytest.mark.parametrize("orography", ["flat", "gaussian"])
def test_torch_step_matches_oracle(orography):
    g = get_grid(4, orography)
    d = g.dims
    s0 = oracle_state(g)
    dt, n = 300.0, 3
    om = pyoracle.OracleModel(g, s0, dt)
    om.run(n)
    G = TS.TorchGrid(g)
    u, w = G.split_v(s0.wind)
    s = {"rho": G.split_s(s0.rho), "rt": G.split_s(s0.rhotheta_v), "ep": G.split_s(s0.exner_pert), "u": u, "w": w}
    with torch.no_grad():
        out = TS.run(G, s, dt, n)
    ref_u, ref_w = G.split_v(om.new.wind)
    pairs = (("rho", out["rho"], G.split_s(om.new.rho)), ("rhotheta_v", out["rt"], G.split_s(om.new.rhotheta_v)),
             ("exner_pert", out["ep"], G.split_s(om.new.exner_pert)), ("u", out["u"], ref_u),
             ("w", out["w"], ref_w))
    for name, a, b in pairs:
        err = float((a - b).abs().max() / b.abs().max())
        print(name, err)
        assert err <= 1e-11, (name, err)
