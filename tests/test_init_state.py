"""The analytic balanced initial state (set_ideal_init restated,
src/io/set_initial_state.c:23-242).

The balance is checked in its hydrostatic form, not in the quadratic the
builder solves: for every column and pair of adjacent layers,
  c_p * (theta_k + theta_k+1)/2 * (Pi_k - Pi_k+1)
    + (Phi_k - Phi_k+1) + (v2_k - v2_k+1)/2 - (z_k - z_k+1) * pvt_w = 0
with theta and Pi the full fields (background plus perturbation), which is
the discrete vertical momentum balance the reference's quadratic
(set_initial_state.c:183-189) encodes.  A wrong sign, a swapped layer or a
missing term fails it by many orders of magnitude.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle
from conftest import get_grid
from game_archived_amd import _native as N
from game_archived_amd.state import initial_state

C_D_P = 1005.0
R_D = 287.057811
P_0 = 100000.0


@pytest.fixture(scope="module", params=[False, True], ids=["dry", "moist"])
def case(request):
    moist = request.param
    g = get_grid(4, "gaussian", moist=1 if moist else 0)
    terms = {}

    def balance(wind, density):
        vsq, pvt = pyoracle.balance_terms(g)(wind, density)
        terms["vsq"], terms["pvt"] = vsq, pvt
        return vsq, pvt

    st = initial_state(g, balance, moist=moist)
    return g, st, terms, moist


def test_discrete_hydrostatic_balance(case):
    g, st, terms, moist = case
    d = g.dims
    nsh, L = d.nsh, d.n_layers
    exner = (st.exner_pert + g.g["exner_bg"]).reshape(L, nsh)
    theta = (st.theta_v_pert + g.g["theta_v_bg"]).reshape(L, nsh)
    phi = g.g["gravity_potential"].reshape(L, nsh)
    z = g.g["z_scalar"].reshape(L, nsh)
    vsq = terms["vsq"].reshape(L, nsh)
    pvt = terms["pvt"]
    resid, scale = [], []
    for k in range(L - 1):
        pvt_w = pvt[(k + 1) * d.vpl + np.arange(nsh)]
        r = (C_D_P * 0.5 * (theta[k] + theta[k + 1]) * (exner[k] - exner[k + 1])
             + (phi[k] - phi[k + 1]) + 0.5 * (vsq[k] - vsq[k + 1]) - (z[k] - z[k + 1]) * pvt_w)
        resid.append(np.abs(r))
        scale.append(np.abs(phi[k] - phi[k + 1]))
    rel = float(np.max(np.array(resid) / np.array(scale)))
    print("max relative residual", rel)
    assert rel <= 1e-11


def test_equation_of_state_and_definitions(case):
    """rho theta_v = rho * theta_v; p = rho R_d T_v with p = P_0 Pi^(c_p/R_d)"""
    g, st, terms, moist = case
    d = g.dims
    theta = st.theta_v_pert + g.g["theta_v_bg"]
    exner = st.exner_pert + g.g["exner_bg"]
    gas = st.rho[d.n_cond * d.s:(d.n_cond + 1) * d.s]
    assert np.allclose(st.rhotheta_v, gas * theta, rtol=1e-14, atol=0)
    tv = theta * exner
    p = P_0 * exner ** (C_D_P / R_D)
    assert np.allclose(p, gas * R_D * tv, rtol=1e-12, atol=0)


def test_profile_and_humidity(case):
    """T = 288 K - 6.5 K/km below 11 km, isothermal above (plus the seeded
    noise of std 0.1 K); moist: vapour only below 8 km, q <= 0.8 (1 + 5 %) q_sat"""
    g, st, terms, moist = case
    d = g.dims
    z = g.g["z_scalar"]
    T = st.init_temperature
    prof = 288.0 - 0.0065 * np.minimum(z, 11000.0)
    dev = T - prof
    assert abs(dev.mean()) < 0.01 and 0.08 < dev.std() < 0.12
    if moist:
        vap = st.rho[(d.n_cond + 1) * d.s:]
        assert np.all(vap[z >= 8000.0] == 0.0)
        assert np.all(vap[z < 8000.0] > 0.0)
        for k in range(d.n_cond):
            assert np.all(st.rho[k * d.s:(k + 1) * d.s] == 0.0)


def test_noise_is_seeded_and_standard():
    lib = N.host_lib()
    x = np.array([lib.game_gauss(C.c_uint64(20261202), h, l) for h in range(2000) for l in range(10)])
    assert abs(x.mean()) < 0.03 and abs(x.std() - 1.0) < 0.03
    y = np.array([lib.game_gauss(C.c_uint64(20261202), h, l) for h in range(2000) for l in range(10)])
    assert np.array_equal(x, y)
    z = np.array([lib.game_gauss(C.c_uint64(20261203), h, l) for h in range(2000) for l in range(10)])
    assert abs(np.corrcoef(x, z)[0, 1]) < 0.05
