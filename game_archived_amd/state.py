"""Host containers for the step's fields (numpy), mirroring the reference's
struct state, diagnostics, forcings and irreversible_quantities
(src/game_types.h:128-238), restricted to the members the step touches,
plus the analytic initial state of the measured workloads."""
import numpy as np

from . import _native as N


class Fields:
    """A bag of named float64 arrays plus its ctypes view."""
    VIEW = None
    SPEC = ()

    def __init__(self, dims):
        self.dims = dims
        for name, size in self.SPEC:
            setattr(self, name, np.zeros(int(size(dims)), dtype=np.float64))

    def arrays(self):
        return {name: getattr(self, name) for name, _ in self.SPEC}

    def view(self):
        return N.fill_view(self.VIEW(), self.arrays())

    def copy(self):
        o = type(self)(self.dims)
        for name, _ in self.SPEC:
            getattr(o, name)[:] = getattr(self, name)
        return o


class State(Fields):
    VIEW = N.StateView
    SPEC = (("rho", lambda d: d.n_const * d.s), ("rhotheta_v", lambda d: d.s), ("theta_v_pert", lambda d: d.s),
            ("exner_pert", lambda d: d.s), ("wind", lambda d: d.v))


class Diagnostics(Fields):
    VIEW = N.DiagView
    SPEC = (("flux_density", lambda d: d.v), ("flux_density_divv", lambda d: d.s),
            ("rel_vort_on_triangles", lambda d: d.dual_v_vectors), ("rel_vort", lambda d: d.curl),
            ("pot_vort", lambda d: d.curl), ("temperature", lambda d: d.s),
            ("scalar_field_placeholder", lambda d: d.s), ("v_squared", lambda d: d.s))


class Forcings(Fields):
    VIEW = N.ForcingsView
    SPEC = (("pgrad_acc_old", lambda d: d.v), ("pressure_gradient_acc_neg_nl", lambda d: d.v),
            ("pressure_gradient_acc_neg_l", lambda d: d.v), ("pressure_grad_condensates_v", lambda d: d.v),
            ("v_squared_grad", lambda d: d.v), ("pot_vort_tend", lambda d: d.v), ("radiation_tendency", lambda d: d.s))


class Irrev(Fields):
    VIEW = N.IrrevView
    SPEC = (("temperature_diffusion_heating", lambda d: d.s), ("friction_acc", lambda d: d.v),
            ("heating_diss", lambda d: d.s), ("pressure_gradient_decel_factor", lambda d: d.s),
            ("condensates_sediment_heat", lambda d: d.s), ("mass_diff_tendency", lambda d: d.n_const * d.s),
            ("phase_trans_rates", lambda d: (d.n_cond + 1) * d.s), ("phase_trans_heating_rate", lambda d: d.s))


def init_params(moist=False, seed=20261202, sigma_t=0.1):
    """The analytic workload: T = 288 K - 6.5 K/km up to 11 km, isothermal above;
    u = 35 m/s cos^2(lat) sin(pi min(z, 20 km)/20 km); seeded N(0, sigma_t) noise
    on T per cell and layer; moist: q = 0.8 q_sat (1 + U(-5 %, 5 %)) below 8 km."""
    p = N.InitParams()
    N.host_lib().game_init_default_params(p)
    p.seed = seed
    p.sigma_t = sigma_t
    p.moist = 1 if moist else 0
    return p


def initial_state(grid, balance_terms, moist=False, seed=20261202, sigma_t=0.1):
    """Builds a balanced state (set_ideal_init restated).  `balance_terms(wind,
    density)` returns (v_squared, pot_vort_tend) of the initial wind, evaluated
    with the step's operators (set_initial_state.c:155-166)."""
    d = grid.dims
    lib = N.host_lib()
    p = init_params(moist, seed, sigma_t)
    st = State(d)
    work = {n: np.zeros(d.s) for n in ("temperature", "temperature_v", "pressure", "specific_humidity", "density")}
    wv = N.fill_view(N.InitWork(), work)
    sv = st.view()
    rc = lib.game_init_phase1(p, d, grid.grid_view(), N.ptr(grid.x["latitude_vector"]), sv, wv)
    if rc:
        raise RuntimeError(f"initial state phase 1 failed ({rc})")
    v_squared, pot_vort_tend = balance_terms(st.wind, work["density"])
    rc = lib.game_init_phase2(p, d, grid.grid_view(), N.ptr(np.ascontiguousarray(v_squared)),
                              N.ptr(np.ascontiguousarray(pot_vort_tend)), sv, wv)
    if rc:
        raise RuntimeError(f"initial state phase 2 failed ({rc})")
    st.init_temperature = work["temperature"]
    return st


def write_state_file(path, grid, state, temperature, tke=None, n_soil_layers=5):
    """the model's state file (src/io/write_output.c:1428-1503): densities,
    temperature and wind of `state`; rh, rel_vort, divv_h_all_layers and
    t_soil are written as zeros (output diagnostics), tke if given"""
    import os
    d = grid.dims
    lib = N.host_lib()
    tk = N.ptr(np.ascontiguousarray(tke, dtype=np.float64)) if tke is not None else None
    rc = lib.game_state_file_write(os.fsencode(path), d, N.ptr(state.rho), N.ptr(np.ascontiguousarray(temperature)),
                                   N.ptr(state.wind), None, None, None, tk, None, n_soil_layers)
    if rc != 0:
        raise RuntimeError(f"state file write failed ({rc}): {lib.game_nc_last_error().decode()}")


def read_state_file(path, grid):
    """read_init_data (src/io/set_initial_state.c:244-336): densities,
    temperature and wind from the file, vapour limited to saturation, then
    rhotheta_v, theta_v_pert and exner_pert from the equation of state
    (game_state_from_temperature).  Returns (state, temperature, tke or None)."""
    import ctypes as C
    import os
    d = grid.dims
    lib = N.host_lib()
    st = State(d)
    temperature = np.zeros(d.s)
    tke = np.zeros(d.s)
    found = C.c_int(0)
    rc = lib.game_state_file_read(os.fsencode(path), d, N.ptr(st.rho), N.ptr(temperature), N.ptr(st.wind),
                                  N.ptr(tke), C.byref(found))
    if rc != 0:
        raise RuntimeError(f"state file read failed ({rc}): {lib.game_nc_last_error().decode()}")
    rho = st.rho.copy()
    rc = lib.game_state_from_temperature(d, grid.grid_view(), N.ptr(rho), N.ptr(temperature), st.view())
    if rc != 0:
        raise RuntimeError(f"state from temperature failed ({rc})")
    return st, temperature, (tke if found.value else None)
