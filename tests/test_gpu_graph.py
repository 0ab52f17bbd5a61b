"""Steady steps replayed from HIP graphs against the same steps launched
eagerly (GAME_GPU_NO_GRAPH): bit-identical states after every replay, with
the program's own work (state downloads, integrals, a changed dt that
rebuilds the graphs) and synchronisations between the replays.  The same for
the step's independent branches on two streams against one stream
(GAME_GPU_ONE_STREAM)."""
import os

import numpy as np
import pytest

from conftest import get_grid, oracle_state

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind")
# (steps, dt) of successive run() calls: the first call's first step is eager
# in both modes, later calls replay the captured steps; 250 s rebuilds them
CALLS = [(3, 300.0), (1, 300.0), (2, 300.0), (2, 250.0), (3, 300.0)]


def trajectory(grid, s0, precision, graphs, one_stream=False):
    from game_archived_amd.rkhevi import DeviceModel
    env = {"GAME_GPU_NO_GRAPH": None if graphs else "1", "GAME_GPU_ONE_STREAM": "1" if one_stream else None}
    for k, v in env.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    try:
        dm = DeviceModel(grid, precision=precision)
    finally:
        for k in env:
            os.environ.pop(k, None)
    out = []
    try:
        dm.start(s0)
        for n, dt in CALLS:
            dm.run(n, dt)
            dm.synchronize()
            st = dm.get_state()
            out.append(({f: np.array(getattr(st, f)) for f in FIELDS}, dm.integrals()))
    finally:
        dm.close()
    return out


@pytest.mark.parametrize("precision_name", ["fp64", "mixed"])
def test_graph_replays_match_eager_steps(precision_name):
    from game_archived_amd.rkhevi import FP64, MIXED
    prec = {"fp64": FP64, "mixed": MIXED}[precision_name]
    print(f"case: graph replay vs eager, res 4 moist gaussian, {precision_name}, calls {CALLS}", flush=True)
    g = get_grid(4, "gaussian", moist=1)
    s0 = oracle_state(g, moist=True)
    eager = trajectory(g, s0, prec, graphs=False)
    graph = trajectory(g, s0, prec, graphs=True)
    for k, ((fe, ie), (fg, ig)) in enumerate(zip(eager, graph)):
        for f in FIELDS:
            assert np.array_equal(fe[f], fg[f]), (k, f)
            assert np.isfinite(fg[f]).all(), (k, f)
        assert np.array_equal(ie, ig), k


@pytest.mark.parametrize("precision_name", ["fp64", "mixed"])
@pytest.mark.parametrize("res,layers,n_oro", [(4, 26, 23), (3, 55, 40)])
def test_two_streams_match_one_stream(precision_name, res, layers, n_oro):
    """the branches on the second stream (kinetic energy, pressure gradient,
    the first substep's tracer solve) change no bit against one stream, the
    two-stream steps replayed from graphs and the one-stream steps eager, at
    26 layers and with the 55-layer column kernels"""
    from game_archived_amd.rkhevi import FP64, MIXED
    prec = {"fp64": FP64, "mixed": MIXED}[precision_name]
    print(f"case: two streams (graphs) vs one stream (eager), res {res} x {layers} moist gaussian, {precision_name}",
          flush=True)
    g = get_grid(res, "gaussian", moist=1, n_layers=layers, n_oro=n_oro)
    s0 = oracle_state(g, moist=True)
    one = trajectory(g, s0, prec, graphs=False, one_stream=True)
    two = trajectory(g, s0, prec, graphs=True)
    for k, ((fo, io), (ft, it)) in enumerate(zip(one, two)):
        for f in FIELDS:
            assert np.array_equal(fo[f], ft[f]), (k, f)
        assert np.array_equal(io, it), k
