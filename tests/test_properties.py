"""Property tests (hypothesis) of the host code and the oracle, over
generated inputs rather than fixed cases:

* the netCDF classic codec: any set of int32 / float64 variables (NaN,
  infinities, signed zeros and subnormals included) comes back bit for bit,
  and scipy reads the same bits;
* the three-band solve of the oracle (column_solvers.c:641-667) against
  LAPACK's banded solver for diagonally dominant systems of any size the
  model uses;
* the equation of state of the restart path (read_init_data,
  set_initial_state.c:244-336) inverts temperature_diagnostics for
  perturbed states;
* the grid generator's self-checks (index.c:53-107,
  coriolis.c:393-427, rhombus_averaging.c:242-266) on grids of every
  resolution the tests use, Lloyd iterations included.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st
from scipy.io import netcdf_file

import pyoracle
from conftest import get_grid, oracle_state
from game_archived_amd import _native as N
from test_files import c_read, c_write

SETTINGS = settings(max_examples=40, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])

names = st.text(alphabet="abcdefghijklmnopqrstuvwxyz_0123456789", min_size=1, max_size=24).filter(
    lambda s: s[0].isalpha())
doubles = st.floats(allow_nan=True, allow_infinity=True, allow_subnormal=True, width=64)
ints = st.integers(min_value=-2**31, max_value=2**31 - 1)


@st.composite
def variable_sets(draw):
    n = draw(st.integers(1, 6))
    vs_names = draw(st.lists(names, min_size=n, max_size=n, unique=True))
    out = []
    for k, name in enumerate(vs_names):
        length = draw(st.integers(1, 50))
        if draw(st.booleans()):
            a = np.array(draw(st.lists(doubles, min_size=length, max_size=length)), dtype=np.float64)
        else:
            a = np.array(draw(st.lists(ints, min_size=length, max_size=length)), dtype=np.int32)
        units = draw(st.one_of(st.none(), st.text(alphabet="mKs/^*()2", min_size=1, max_size=10)))
        out.append((name, f"dim_{k}", a, units))
    return out


@SETTINGS
@given(variable_sets())
def test_codec_round_trip(tmp_path, variables):
    p = str(tmp_path / "h.nc")
    assert c_write(p, variables) == 0
    f = netcdf_file(p, "r", mmap=False)
    try:
        for name, dim, a, units in variables:
            b = c_read(p, name, a.dtype, a.size)
            assert b.tobytes() == a.tobytes(), name                  # bit for bit, NaN payloads included
            v = f.variables[name]
            assert np.asarray(v.data).astype(a.dtype).tobytes() == a.tobytes(), name
            assert getattr(v, "units", None) == (units.encode() if units else None)
    finally:
        f.close()


@SETTINGS
@given(st.integers(2, 64), st.integers(0, 2**32 - 1))
def test_thomas_solves_diagonally_dominant_systems(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(n - 1) * rng.uniform(0.01, 10.0)
    e = rng.standard_normal(n - 1) * rng.uniform(0.01, 10.0)
    off = np.zeros(n)
    off[1:] += np.abs(c)
    off[:-1] += np.abs(e)
    d = (off + rng.uniform(0.1, 5.0, n)) * rng.choice([-1.0, 1.0])
    r = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3)
    x = np.zeros(n)
    pyoracle.lib().or_thomas(N.ptr(np.ascontiguousarray(c)), N.ptr(np.ascontiguousarray(d)),
                             N.ptr(np.ascontiguousarray(e)), N.ptr(np.ascontiguousarray(r)), N.ptr(x), n)
    ab = np.zeros((3, n))
    ab[0, 1:] = e
    ab[1] = d
    ab[2, :-1] = c
    xr = scipy.linalg.solve_banded((1, 1), ab, r)
    assert np.allclose(x, xr, rtol=1e-12, atol=1e-12 * np.abs(xr).max())


def _host():
    return N.host_lib()


@settings(max_examples=12, deadline=None)
@given(st.integers(0, 2**32 - 1), st.floats(0.0, 0.02), st.booleans())
def test_state_from_temperature_inverts_the_diagnostics(seed, amplitude, moist):
    """perturb a balanced state's densities and rhotheta_v, diagnose the
    temperature, and rebuild the state from densities and temperature: the
    rebuilt rhotheta_v and exner are the perturbed state's"""
    from game_archived_amd.state import State
    g = get_grid(3, "gaussian", moist=1 if moist else 0)
    d = g.dims
    s0 = oracle_state(g, moist=moist)
    rng = np.random.default_rng(seed)
    s = s0.copy()
    s.rho *= 1.0 + amplitude * rng.uniform(-1, 1, s.rho.size)
    if moist:
        # keep the vapour below saturation, so the limiter leaves it alone
        s.rho[5 * d.s:6 * d.s] *= 0.5
    gas = s.rho[d.n_cond * d.s:(d.n_cond + 1) * d.s]
    theta = (s0.rhotheta_v / s0.rho[d.n_cond * d.s:(d.n_cond + 1) * d.s]) * (1 + amplitude * rng.uniform(-1, 1, d.s))
    s.rhotheta_v = gas * theta
    s.theta_v_pert[:] = theta - g.g["theta_v_bg"]
    R_D, C_D_V, P_0 = 287.057811, 717.942189, 100000.0
    s.exner_pert[:] = (R_D * s.rhotheta_v / P_0) ** (R_D / C_D_V) - g.g["exner_bg"]
    T = np.zeros(d.s)
    pyoracle.lib().or_temperature_diagnostics(pyoracle.ctx(g), s.view(), N.ptr(T))
    out = State(d)
    rc = _host().game_state_from_temperature(d, g.grid_view(), N.ptr(s.rho.copy()), N.ptr(T), out.view())
    assert rc == 0
    assert np.allclose(out.rhotheta_v, s.rhotheta_v, rtol=1e-13, atol=0)
    assert np.allclose(out.exner_pert + g.g["exner_bg"], s.exner_pert + g.g["exner_bg"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("res,oro,lloyd", [(3, "flat", 0), (3, "gaussian", 0), (4, "gaussian", 0), (5, "flat", 2000)])
def test_grid_self_checks(res, oro, lloyd):
    """the reference's abort-on-failure invariants as checks: every edge in
    two cells with opposite signs, five or six edges per cell, TRSK
    antisymmetry (Thuburn et al. 2009, Eq. 39), rhombus weights in (0, 1)
    summing to 1"""
    g = get_grid(res, oro, n_lloyd=lloyd)
    d = g.dims
    adj = g.g["adjacent_vector_indices_h"].reshape(-1, 6)
    sg = g.g["adjacent_signs_h"].reshape(-1, 6)
    cnt = np.zeros(d.nvh, dtype=int)
    ssum = np.zeros(d.nvh, dtype=int)
    for j in range(6):
        valid = np.ones(d.nsh, bool)
        valid[:12] = j < 5
        np.add.at(cnt, adj[valid, j], 1)
        np.add.at(ssum, adj[valid, j], sg[valid, j])
    assert (cnt == 2).all() and (ssum == 0).all()
    assert (sg[:12, 5] == 0).all() and (np.abs(sg[12:]) == 1).all()
    w = g.g["trsk_weights"].reshape(-1, 10)
    ti = g.g["trsk_indices"].reshape(-1, 10)
    fr, to = g.g["from_index"], g.g["to_index"]

    def used(e, j):
        """the TRSK slots the generator fills: beside a pentagon a side has
        four terms, and its fifth slot is padding (index 0, weight 0)"""
        return (j != 4 or fr[e] >= 12) and (j != 9 or to[e] >= 12)

    # coriolis.c:393-427: w_ef d_e / d'_f + w_fe d_f / d'_e = 0 within 1e-10,
    # d the primal normal distance of the top layer's edges, d' the dual one
    # of the top level (the generator's height rescaling is ~1 and divides
    # both terms alike); every term's partner is found exactly once
    nd = g.g["normal_distance"][d.nsh:d.nsh + d.nvh]
    dnd = g.dg["normal_distance"][:d.nvh]
    worst = 0.0
    for e in range(d.nvh):
        for j in range(10):
            if not used(e, j):
                continue
            f = ti[e, j]
            ks = [k for k in range(10) if ti[f, k] == e and used(f, k)]
            assert len(ks) == 1, (e, j, f, ks)
            worst = max(worst, abs(w[e, j] * nd[e] / dnd[f] + w[f, ks[0]] * nd[f] / dnd[e]))
    assert worst < 1e-9
    drw = g.g["density_to_rhombi_weights"].reshape(-1, 4)
    assert ((drw > 0) & (drw < 1)).all()
    assert np.abs(drw.sum(axis=1) - 1).max() < 1e-9
