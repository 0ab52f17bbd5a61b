"""ctypes bindings of the C ABI in include/game_rkhevi.h.

The structs below mirror the header field for field; the libraries are the
in-tree builds made by game_archived_amd.build (lib/ next to this file).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GAME_LIB_DIR: another directory of the same libraries (an instrumented build)
LIBDIR = os.environ.get("GAME_LIB_DIR") or os.path.join(HERE, "lib")

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class Dims(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "res_id", "n_layers", "moisture_on", "n_oro_layers", "n_const", "n_cond",
        "n_levels", "nsh", "nvh", "nt", "vpl", "dvpl")] + [
        (n, C.c_longlong) for n in ("s", "v", "curl", "dual_v", "dual_v_vectors")]


def make_dims(res_id, n_layers, moisture_on, n_oro_layers):
    """game_dims_derive() of include/game_rkhevi.h (src/game_types.h:12-60)."""
    d = Dims()
    d.res_id, d.n_layers, d.moisture_on, d.n_oro_layers = res_id, n_layers, moisture_on, n_oro_layers
    four_r = 1 << (2 * res_id)
    hexagons = 10 * (four_r - 1)
    d.n_cond = 4 * moisture_on
    d.n_const = 1 + moisture_on + d.n_cond
    d.n_levels = n_layers + 1
    d.nsh = 12 + hexagons
    d.nvh = 30 + 3 * hexagons
    d.nt = 20 * four_r
    d.vpl = d.nsh + d.nvh
    d.dvpl = d.nvh + d.nt
    d.s = d.nsh * n_layers
    d.v = n_layers * d.nvh + d.n_levels * d.nsh
    d.curl = n_layers * 2 * d.nvh + d.nvh
    d.dual_v = d.n_levels * d.nvh + n_layers * d.nt
    d.dual_v_vectors = n_layers * d.nt
    return d


GRID_FIELDS = [
    ("no_of_oro_layers", C.c_int),
    ("normal_distance", DP), ("volume", DP), ("area", DP), ("z_scalar", DP), ("z_vector", DP),
    ("gravity_potential", DP), ("gravity_m", DP), ("slope", DP), ("theta_v_bg", DP), ("exner_bg", DP),
    ("exner_bg_grad", DP), ("layer_thickness", DP),
    ("trsk_indices", IP), ("trsk_modified_curl_indices", IP), ("from_index", IP), ("to_index", IP),
    ("adjacent_vector_indices_h", IP), ("adjacent_signs_h", IP), ("density_to_rhombi_indices", IP),
    ("latitude_scalar", DP), ("longitude_scalar", DP), ("inner_product_weights", DP), ("direction", DP),
    ("density_to_rhombi_weights", DP), ("trsk_weights", DP), ("is_land", IP),
    ("toa", C.c_double), ("radius", C.c_double), ("stretching_parameter", C.c_double), ("oro_id", C.c_int),
]


class GridView(C.Structure):
    _fields_ = GRID_FIELDS


class DualgridView(C.Structure):
    _fields_ = [("area", DP), ("z_vector", DP), ("normal_distance", DP), ("from_index", IP), ("to_index", IP),
                ("vorticity_indices_triangles", IP), ("vorticity_signs_triangles", IP), ("f_vec", DP)]


class StateView(C.Structure):
    _fields_ = [(n, DP) for n in ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind")]


class DiagView(C.Structure):
    _fields_ = [(n, DP) for n in ("flux_density", "flux_density_divv", "rel_vort_on_triangles", "rel_vort",
                                  "pot_vort", "temperature", "scalar_field_placeholder", "v_squared")]


class ForcingsView(C.Structure):
    _fields_ = [(n, DP) for n in ("pgrad_acc_old", "pressure_gradient_acc_neg_nl", "pressure_gradient_acc_neg_l",
                                  "pressure_grad_condensates_v", "v_squared_grad", "pot_vort_tend",
                                  "radiation_tendency")]


class IrrevView(C.Structure):
    _fields_ = [(n, DP) for n in ("temperature_diffusion_heating", "friction_acc", "heating_diss",
                                  "pressure_gradient_decel_factor", "condensates_sediment_heat",
                                  "mass_diff_tendency", "phase_trans_rates", "phase_trans_heating_rate")]


class Config(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "totally_first_step_bool", "temperature_diff_h", "temperature_diff_v", "momentum_diff_h",
        "momentum_diff_v", "mass_diff_h", "mass_diff_v", "rad_on", "prog_soil_temp", "sfc_phase_trans",
        "sfc_sensible_heat_flux", "rad_update", "time_to_next_analysis", "pbl_scheme", "total_run_span")] + [
        (n, C.c_double) for n in ("damping_start_height_over_toa", "damping_coeff_max", "impl_thermo_weight",
                                  "cloud_droplets_velocity", "rain_velocity", "snow_velocity", "radiation_delta_t")]


def default_config():
    """The switch vector of every measured workload: all diffusion, radiation,
    PBL and surface fluxes off; the constants of src/coordinator.c:54-57, 221-223."""
    c = Config()
    c.totally_first_step_bool = 1
    c.impl_thermo_weight = 0.75
    c.damping_start_height_over_toa = 0.53
    c.damping_coeff_max = 0.25
    c.cloud_droplets_velocity = 0.01
    c.rain_velocity = 10.0
    c.snow_velocity = 5.0
    return c


class GridParams(C.Structure):
    _fields_ = [("res_id", C.c_int), ("n_layers", C.c_int), ("n_oro_layers", C.c_int),
                ("n_lloyd_iterations", C.c_int), ("stretching_parameter", C.c_double), ("toa", C.c_double),
                ("radius_rescale", C.c_double), ("oro_id", C.c_int), ("oro_h0", C.c_double),
                ("oro_half_width", C.c_double), ("oro_lat", C.c_double), ("oro_lon", C.c_double),
                ("generator_latlon", DP), ("rhombus_tolerance", C.c_double),
                ("trsk_tolerance", C.c_double)]


class GridExtras(C.Structure):
    _fields_ = [("latitude_vector", DP), ("longitude_vector", DP), ("latitude_scalar_dual", DP),
                ("longitude_scalar_dual", DP), ("direction_dual", DP), ("pent_hex_face_unity_sphere", DP),
                ("triangle_face_unit_sphere", DP), ("oro", DP), ("vorticity_indices_rhombi", IP)]


class GridFileExtras(C.Structure):
    """game_grid_file_extras_t (include/game_files.h)"""
    _fields_ = [("latitude_vector", DP), ("longitude_vector", DP), ("latitude_scalar_dual", DP),
                ("longitude_scalar_dual", DP), ("interpol_indices", IP), ("interpol_weights", DP),
                ("sfc_albedo", DP), ("sfc_rho_c", DP), ("t_conductivity", DP), ("roughness_length", DP),
                ("n_lloyd_iterations", C.c_int)]


class InitParams(C.Structure):
    _fields_ = [("t_sfc", C.c_double), ("lapse_rate", C.c_double), ("z_tropopause", C.c_double),
                ("p_sfc", C.c_double), ("u_max", C.c_double), ("z_jet_top", C.c_double), ("sigma_t", C.c_double),
                ("seed", C.c_uint64), ("moist", C.c_int), ("rh_fraction", C.c_double), ("q_noise", C.c_double),
                ("z_moist_top", C.c_double)]


class InitWork(C.Structure):
    _fields_ = [(n, DP) for n in ("temperature", "temperature_v", "pressure", "specific_humidity", "density")]


def ptr(a):
    """Raw pointer of a contiguous numpy array (double or int32)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    if a.dtype == np.float64:
        return a.ctypes.data_as(DP)
    if a.dtype == np.int32:
        return a.ctypes.data_as(IP)
    raise TypeError(f"unsupported dtype {a.dtype}")


def fill_view(view, arrays):
    """Point every pointer field of a ctypes view at the numpy array of the same name."""
    for name, typ in view._fields_:
        if typ in (DP, IP) and name in arrays and arrays[name] is not None:
            setattr(view, name, ptr(arrays[name]))
    return view


_libs = {}


def load(name):
    """Loads lib/<name>.so built in-tree; raises if it is missing."""
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIBDIR, name + ".so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() first")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def host_lib():
    lib = load("libgame_host")
    lib.game_grid_generate.argtypes = [C.POINTER(GridParams), C.POINTER(Dims), C.POINTER(GridView),
                                       C.POINTER(DualgridView), C.POINTER(GridExtras)]
    lib.game_grid_generate.restype = C.c_int
    lib.game_grid_generators.argtypes = [C.c_int, C.c_int, DP]
    lib.game_grid_generators.restype = C.c_int
    lib.game_grid_lloyd_state.argtypes = [C.c_int, C.c_int, DP, DP, IP, IP, IP]
    lib.game_grid_lloyd_state.restype = C.c_int
    lib.game_grid_last_error.restype = C.c_char_p
    lib.game_init_default_params.argtypes = [C.POINTER(InitParams)]
    for fn in (lib.game_init_phase1,):
        fn.argtypes = [C.POINTER(InitParams), C.POINTER(Dims), C.POINTER(GridView), DP, C.POINTER(StateView),
                       C.POINTER(InitWork)]
        fn.restype = C.c_int
    lib.game_init_phase2.argtypes = [C.POINTER(InitParams), C.POINTER(Dims), C.POINTER(GridView), DP, DP,
                                     C.POINTER(StateView), C.POINTER(InitWork)]
    lib.game_init_phase2.restype = C.c_int
    lib.game_gauss.argtypes = [C.c_uint64, C.c_longlong, C.c_int]
    lib.game_gauss.restype = C.c_double
    P = C.POINTER
    lib.game_grid_file_write.argtypes = [C.c_char_p, P(Dims), P(GridView), P(DualgridView), P(GridFileExtras)]
    lib.game_grid_file_read.argtypes = [C.c_char_p, P(Dims), P(GridView), P(DualgridView), P(GridFileExtras)]
    lib.game_latlon_interpolation.argtypes = [C.c_int, DP, DP, IP, IP, IP, IP, DP]
    lib.game_sfc_properties.argtypes = [C.c_int, IP, DP, DP, DP, DP]
    lib.game_grid_statistics_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, DP, DP, DP]
    lib.game_state_file_write.argtypes = [C.c_char_p, P(Dims), DP, DP, DP, DP, DP, DP, DP, DP, C.c_int]
    lib.game_state_file_read.argtypes = [C.c_char_p, P(Dims), DP, DP, DP, DP, IP]
    lib.game_nc_last_error.restype = C.c_char_p
    lib.game_state_from_temperature.argtypes = [P(Dims), P(GridView), DP, DP, P(StateView)]
    lib.game_state_from_temperature.restype = C.c_int
    return lib


def gpu_lib():
    lib = load("libgame_rkhevi")
    P = C.POINTER
    lib.game_gpu_init.argtypes = [P(Dims), P(GridView), P(DualgridView), P(Config), C.c_int]
    lib.game_gpu_set_state.argtypes = [C.c_int, P(StateView)]
    lib.game_gpu_get_state.argtypes = [C.c_int, P(StateView)]
    lib.game_gpu_get_forcings.argtypes = [P(ForcingsView)]
    lib.game_gpu_get_diagnostics.argtypes = [P(DiagView)]
    lib.game_gpu_get_irrev.argtypes = [P(IrrevView)]
    lib.game_gpu_set_config.argtypes = [P(Config)]
    lib.game_gpu_manage_rkhevi.argtypes = [C.c_double, C.c_int]
    lib.game_gpu_linear_combine_two_states.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    lib.game_gpu_run_steps.argtypes = [C.c_int, C.c_double, C.c_int]
    lib.game_gpu_temperature_diagnostics.argtypes = [C.c_int]
    lib.game_gpu_integrals.argtypes = [C.c_int, DP]
    lib.game_gpu_write_integrals.argtypes = [C.c_int, C.c_double, C.c_char_p]
    lib.game_gpu_run_steps_write_integrals.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_char_p]
    lib.game_gpu_enable_timing.argtypes = [C.c_int]
    lib.game_gpu_kernel_stats.argtypes = [C.c_int, P(C.c_char_p), P(C.c_longlong), P(C.c_double)]
    lib.game_gpu_last_error.restype = C.c_char_p
    lib.game_gpu_bytes_allocated.restype = C.c_longlong
    lib.game_gpu_init_balance_terms.argtypes = [DP, DP, DP, DP]
    lib.game_gpu_kernel_count.restype = C.c_int
    VP = C.c_void_p   # device pointers of the operator-level interface
    for name, n in (("grad", 2), ("divv_h", 2), ("curl_triangles", 2), ("rhombus_density", 2),
                    ("vorticity_flux", 3), ("inner_product", 3)):
        getattr(lib, f"game_gpu_op_{name}").argtypes = [VP] * n
        getattr(lib, f"game_gpu_op_{name}_adjoint").argtypes = [VP] * n
    lib.game_gpu_op_thomas.argtypes = [C.c_int, C.c_int] + [VP] * 6
    lib.game_gpu_op_thomas_noalloc.argtypes = [C.c_int, C.c_int] + [VP] * 5
    lib.game_gpu_window_forward.argtypes = [C.c_int, C.c_double] + [VP] * 7
    lib.game_gpu_window_vjp.argtypes = [C.c_int, C.c_double, C.c_int] + [VP] * 10
    lib.game_gpu_window_gradient.argtypes = [C.c_int, C.c_double, C.c_int] + [VP] * 4 + [DP] + [VP] * 3
    lib.game_gpu_window_last_timing.argtypes = [DP]
    lib.game_gpu_op_thomas_adjoint.argtypes = [C.c_int, C.c_int] + [VP] * 10
    return lib
