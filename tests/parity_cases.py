"""The cases on which the oracle is pinned to the reference's own compiled
code, run either way: through the oracle (this module) or through the
reference (refharness.reference_*).  Both return {name: float64 array} with
the same keys, so the comparison is bit for bit, and `digest` turns one into
the committed golden vectors (tests/golden/).

Case: RES_ID 5 x 26 layers, moist (the reference's compiled-in layout,
src/game_types.h:14-18), Gaussian mountain, 2000 Lloyd iterations, the
balanced analytic state with seeded condensates and super-saturation.
"""
import hashlib

import numpy as np

DT = 300.0
N_STEPS = 3
OPERATOR_SEED = 1


def case_grid():
    from conftest import get_grid
    return get_grid(5, "gaussian", moist=1, n_lloyd=2000)


def case_state(grid):
    """balanced state plus condensates up to 2e-3 of the gas density in half
    of the boxes and a third of the boxes super-saturated, so every branch of
    the moist physics runs"""
    from conftest import oracle_state
    state = oracle_state(grid, moist=True)
    rng = np.random.default_rng(7)
    d = grid.dims
    gas = state.rho[4 * d.s:5 * d.s]
    for k in range(4):
        state.rho[k * d.s:(k + 1) * d.s] = gas * rng.uniform(0, 2e-3, d.s) * (rng.random(d.s) < 0.5)
    vap = state.rho[5 * d.s:6 * d.s]
    vap *= np.where(rng.random(d.s) < 0.33, 1.6, 1.0)
    return state


def operator_inputs(grid):
    d = grid.dims
    rng = np.random.default_rng(OPERATOR_SEED)
    scal = rng.standard_normal(d.s)
    vec = rng.standard_normal(d.v)
    dens = np.abs(rng.standard_normal(d.s))
    return scal, vec, dens


SCALAR_TIMES_VECTOR = ("scalar_times_vector", "scalar_times_vector_h", "scalar_times_vector_h_upstream",
                       "scalar_times_vector_v")


def oracle_operators(grid, state):
    """grad (gradient_operators.c:66-74), divv_h (divergences.c:14-64),
    divv_h_tracer (divergences.c:66-148), the scalar-times-vector family
    (multiplications.c) and inner_product (inner_product.c:13-43)"""
    import pyoracle
    from game_archived_amd import _native as N
    d = grid.dims
    ctx = pyoracle.ctx(grid)
    O = pyoracle.lib()
    scal, vec, dens = operator_inputs(grid)
    out = {}
    a = np.zeros(d.v)
    O.or_grad(ctx, N.ptr(scal), N.ptr(a))
    out["grad"] = a
    a = np.zeros(d.s)
    O.or_divv_h(ctx, N.ptr(vec), N.ptr(a))
    out["divv_h"] = a
    a = np.zeros(d.s)
    O.or_divv_h_tracer(ctx, N.ptr(vec), N.ptr(dens), N.ptr(state.wind), N.ptr(a))
    out["divv_h_tracer"] = a
    for name in SCALAR_TIMES_VECTOR:
        a = np.zeros(d.v)
        getattr(O, "or_" + name)(ctx, N.ptr(scal), N.ptr(vec), N.ptr(a))
        out[name] = a
    a = np.zeros(d.s)
    O.or_inner_product(ctx, N.ptr(vec), N.ptr(state.wind), N.ptr(a))
    out["inner_product"] = a
    return out


COMPOSITE_FIELDS = ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind",
                    "pgrad_acc_old", "pressure_gradient_acc_neg_nl", "pressure_gradient_acc_neg_l",
                    "pressure_grad_condensates_v", "v_squared_grad", "pot_vort_tend",
                    "tendency_wind", "tendency_rho", "phase_trans_rates", "condensates_sediment_heat",
                    "temperature")


def oracle_composite(grid, state):
    """N_STEPS RKHEVI steps of the oracle with the driver's state copies"""
    import pyoracle
    om = pyoracle.OracleModel(grid, state, DT)
    om.run(N_STEPS)
    out = {f: np.array(getattr(om.new, f)) for f in ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind")}
    for f in ("pgrad_acc_old", "pressure_gradient_acc_neg_nl", "pressure_gradient_acc_neg_l",
              "pressure_grad_condensates_v", "v_squared_grad", "pot_vort_tend"):
        out[f] = np.array(getattr(om.forc, f))
    out["tendency_wind"] = np.array(om.tend.wind)
    out["tendency_rho"] = np.array(om.tend.rho)
    out["phase_trans_rates"] = np.array(om.irrev.phase_trans_rates)
    out["condensates_sediment_heat"] = np.array(om.irrev.condensates_sediment_heat)
    out["temperature"] = np.array(om.diag.temperature)
    return out


def digest(a, n_sample=16, seed=0):
    """sha256 of the little-endian float64 bytes, plus a few exact samples
    (hex floats) that locate a mismatch"""
    a = np.ascontiguousarray(a, dtype="<f8")
    idx = np.sort(np.random.default_rng(seed).choice(a.size, size=min(n_sample, a.size), replace=False))
    return {"n": int(a.size), "sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "sample": [[int(i), float(a[i]).hex()] for i in idx]}


def input_digests(grid, state):
    """digests of what both sides start from: the grid's arrays and the state"""
    out = {}
    for k, v in grid.g.items():
        if isinstance(v, np.ndarray):
            out["grid." + k] = digest(v.astype(np.float64))
    for k, v in grid.dg.items():
        if isinstance(v, np.ndarray):
            out["dualgrid." + k] = digest(v.astype(np.float64))
    for f in ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind"):
        out["state0." + f] = digest(getattr(state, f))
    return out
