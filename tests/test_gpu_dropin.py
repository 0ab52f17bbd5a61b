"""The drop-in boundary end to end on the device: the reference's own driver
protocol (src/coordinator.c:292-324) calls the drop-in library's
`linear_combine_two_states` and `manage_rkhevi` on the reference's structs
(their layout read from the reference's compiled code), and the result is
compared with the same protocol run on the reference's own compiled
functions (refharness.reference_composite).

Case: RES_ID 5 x 26 moist (the reference's compiled-in layout), Gaussian
mountain, 2000 Lloyd iterations, seeded condensates and super-saturation,
3 steps at dt = 300 s.  Moist steps go through `exp`/`pow`, where the device
math library and glibc may differ by an ulp, so the bar is 1e-10 relative.
"""
import ctypes as C
import os

import numpy as np
import pytest

import parity_cases as P
import refharness as R
from conftest import rel_err

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference sources)")]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "game_archived_amd", "lib")
VP = C.c_void_p


@pytest.mark.parametrize("aux", ["per_step", "deferred"])
def test_dropin_driver_protocol_matches_reference(aux):
    """aux per_step: the carried fields reach the host structs after every
    step, as the reference writes them in place; deferred
    (GAME_DROPIN_DEFER_AUX): they stay on the device until
    game_dropin_sync_aux"""
    from game_archived_amd import _native as N
    print(f"case: drop-in r5l26m1, res 5 moist gaussian, 3 steps dt 300, carried fields {aux}", flush=True)
    C.CDLL(os.path.join(LIB, "libgame_rkhevi.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(os.path.join(LIB, "libgame_dropin_r5l26m1.so"))
    lib.linear_combine_two_states.argtypes = [VP, VP, VP, C.c_double, C.c_double, VP]
    lib.manage_rkhevi.argtypes = [VP] * 9 + [C.c_double, C.c_double]
    lib.game_dropin_sync_aux.argtypes = [VP] * 4
    lib.game_dropin_init.argtypes = [VP] * 3
    grid = P.case_grid()
    st = P.case_state(grid)
    d = grid.dims
    ref = R.reference_composite(grid, st)

    g, dg = R.grid_structs(grid)
    so, sn, td = R.state_struct(st), R.state_struct(st), R.RefStruct("State")
    diag, forc, irrev = R.RefStruct("Diagnostics"), R.RefStruct("Forcings"), R.RefStruct("Irreversible_quantities")
    cfg = R.config_struct(N.default_config())
    first = cfg.member("totally_first_step_bool", np.int32, 1)
    os.environ["GAME_GPU_NOABORT"] = "1"
    if aux == "deferred":
        os.environ["GAME_DROPIN_DEFER_AUX"] = "1"
    try:
        # the maintainer's added call after the grid is read (before coordinator.c:250)
        assert lib.game_dropin_init(g.ptr, dg.ptr, cfg.ptr) == 0
        assert lib.linear_combine_two_states(so.ptr, so.ptr, sn.ptr, 1.0, 0.0, g.ptr) == 0   # coordinator.c:292
        first[0] = 1
        for step in range(P.N_STEPS):
            assert lib.linear_combine_two_states(sn.ptr, so.ptr, so.ptr, 1.0, 0.0, g.ptr) == 0   # :305
            assert lib.manage_rkhevi(so.ptr, sn.ptr, g.ptr, dg.ptr, td.ptr, diag.ptr, forc.ptr, irrev.ptr, cfg.ptr,
                                     P.DT, step * P.DT) == 0
            first[0] = 0
        if aux == "deferred":
            assert not np.array(forc.member("pot_vort_tend", np.float64, d.v)).any()   # not yet on the host
            assert lib.game_dropin_sync_aux(td.ptr, diag.ptr, forc.ptr, irrev.ptr) == 0
    finally:
        lib.game_dropin_finalize()
        os.environ.pop("GAME_GPU_NOABORT", None)
        os.environ.pop("GAME_DROPIN_DEFER_AUX", None)

    got = {k: np.array(v) for k, v in R.state_fields(sn, d).items()}
    for f in ("pgrad_acc_old", "pressure_gradient_acc_neg_nl", "pressure_gradient_acc_neg_l", "v_squared_grad",
              "pot_vort_tend"):
        got[f] = np.array(forc.member(f, np.float64, d.v))
    got["tendency_rho"] = np.array(td.member("rho", np.float64, d.n_const * d.s))
    got["phase_trans_rates"] = np.array(irrev.member("phase_trans_rates", np.float64, 5 * d.s))
    got["temperature"] = np.array(diag.member("temperature", np.float64, d.s))
    for k, v in got.items():
        assert np.isfinite(v).all(), k
        e = rel_err(v, ref[k])
        print(k, e)
        assert e <= 1e-10, (k, e)
