"""The mixed-precision form of the device step against the FP64 form.

Mixed form: the work arrays of the horizontal operators, the momentum
tendency, temperature and the phase-transition rates are stored in FP32;
the prognostic state, the density and rhotheta_v tendencies and both
vertical implicit solvers stay FP64.  Bars: relative L2 difference of rho
and rhotheta_v against the FP64 form after the run (<= 1e-4 after one
simulated day is the target; a few steps must sit far below it), and the
dry-air mass conserved as tightly as in the FP64 form.
"""
import numpy as np
import pytest

from conftest import get_grid, oracle_state

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(grid, state0, n_steps, dt, precision):
    from game_archived_amd.rkhevi import DeviceModel
    dm = DeviceModel(grid, precision=precision)
    try:
        dm.start(state0)
        i0 = dm.integrals()
        dm.run(n_steps, dt)
        i1 = dm.integrals()
        return dm.get_state(), i0, i1
    finally:
        dm.close()


@pytest.mark.parametrize("orography", ["flat", "gaussian"])
def test_mixed_dry_close_to_fp64_and_conserves_mass(orography):
    from game_archived_amd.rkhevi import FP64, MIXED
    print(f"case: res 4 dry {orography}, 20 steps dt 300, mixed vs FP64", flush=True)
    g = get_grid(4, orography)
    s0 = oracle_state(g)
    ref, _, _ = run(g, s0, 20, 300.0, FP64)
    mix, i0, i1 = run(g, s0, 20, 300.0, MIXED)
    d = g.dims
    for f in ("rho", "rhotheta_v", "wind"):
        e = rel_l2(getattr(mix, f), getattr(ref, f))
        print(f, e)
        assert e <= 1e-5, (f, e)
    drift = abs(i1[0] - i0[0]) / i0[0]
    print("mass drift", drift)
    assert drift <= 1e-13
    assert np.isfinite(mix.rho).all() and np.isfinite(mix.wind).all()


def test_mixed_moist_close_to_fp64():
    from game_archived_amd.rkhevi import FP64, MIXED
    print("case: res 4 moist flat, 10 steps dt 300, mixed vs FP64", flush=True)
    g = get_grid(4, "flat", moist=1)
    s0 = oracle_state(g, moist=True)
    ref, r0, r1 = run(g, s0, 10, 300.0, FP64)
    mix, m0, m1 = run(g, s0, 10, 300.0, MIXED)
    d = g.dims
    gas = slice(4 * d.s, 5 * d.s)
    for name, a, b in (("rho_gas", mix.rho[gas], ref.rho[gas]), ("rhotheta_v", mix.rhotheta_v, ref.rhotheta_v),
                       ("rho_vapour", mix.rho[5 * d.s:], ref.rho[5 * d.s:])):
        e = rel_l2(a, b)
        print(name, e)
        assert e <= 1e-5, (name, e)
    # total mass of all constituents: moved between them by the phase
    # transitions, conserved in sum up to the sedimentation out of the bottom
    tot_m = sum(m1[:d.n_const]) - sum(m0[:d.n_const])
    tot_r = sum(r1[:d.n_const]) - sum(r0[:d.n_const])
    total = sum(r0[:d.n_const])
    print("total mass change mixed / fp64", tot_m, tot_r)
    # both forms move the same mass out of the bottom: the mixed form's change
    # is the FP64 form's up to its own relative error, plus the conservation floor
    assert abs(tot_m - tot_r) <= 1e-13 * total + 1e-4 * abs(tot_r), (tot_m, tot_r)
    # the dry-air mass (gas minus vapour) is conserved in both forms
    for i0, i1 in ((m0, m1), (r0, r1)):
        dry0, dry1 = i0[d.n_cond] - i0[d.n_cond + 1], i1[d.n_cond] - i1[d.n_cond + 1]
        assert abs(dry1 - dry0) <= 1e-13 * dry0, (dry0, dry1)
