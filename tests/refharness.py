"""Reference structs in numpy buffers, for calling the reference's own
compiled functions (oracle/_ref/libgame_ref.so) from the tests.

The reference fixes its sizes at compile time (src/game_types.h:14-18:
RES_ID 5, 26 layers, moist), so everything here is at that layout.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libgame_ref.so")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def available():
    return os.path.exists(REF_SO) and os.path.exists(ORACLE_SO)


def lib():
    global _lib
    if _lib is None:
        C.CDLL(ORACLE_SO, mode=C.RTLD_GLOBAL)
        _lib = C.CDLL(REF_SO, mode=C.RTLD_GLOBAL)
        _lib.ref_sizeof.restype = C.c_long
        _lib.ref_offsetof.restype = C.c_long
    return _lib


def sizes():
    a = (C.c_long * 8)()
    lib().ref_sizes(a)
    keys = ("res_id", "n_layers", "moisture_on", "nsh", "nvh", "s", "v", "n_const")
    return dict(zip(keys, list(a)))


class RefStruct:
    """A reference struct as a zeroed byte buffer with typed member views."""

    def __init__(self, name):
        self.name = name
        n = lib().ref_sizeof(name.encode())
        assert n > 0, name
        self.buf = np.zeros(n, dtype=np.uint8)

    @property
    def ptr(self):
        return C.c_void_p(self.buf.ctypes.data)

    def member(self, m, dtype, count):
        off = lib().ref_offsetof(self.name.encode(), m.encode())
        assert off >= 0, (self.name, m)
        return self.buf[off:off + np.dtype(dtype).itemsize * count].view(dtype)

    def set(self, m, arr):
        arr = np.ascontiguousarray(arr)
        self.member(m, arr.dtype, arr.size)[:] = arr

    def set_scalar(self, m, value, dtype):
        self.member(m, dtype, 1)[0] = value


GRID_MEMBERS = ("normal_distance", "volume", "area", "z_scalar", "z_vector", "gravity_potential", "gravity_m",
                "slope", "theta_v_bg", "exner_bg", "exner_bg_grad", "layer_thickness", "trsk_indices",
                "trsk_modified_curl_indices", "from_index", "to_index", "adjacent_vector_indices_h",
                "adjacent_signs_h", "density_to_rhombi_indices", "latitude_scalar", "longitude_scalar",
                "inner_product_weights", "direction", "density_to_rhombi_weights", "trsk_weights", "is_land")
DUAL_MEMBERS = ("area", "z_vector", "normal_distance", "from_index", "to_index", "vorticity_indices_triangles",
                "vorticity_signs_triangles", "f_vec")


def grid_structs(grid):
    g, dg = RefStruct("Grid"), RefStruct("Dualgrid")
    for m in GRID_MEMBERS:
        g.set(m, grid.g[m])
    g.set_scalar("no_of_oro_layers", grid.dims.n_oro_layers, np.int32)
    g.set_scalar("toa", grid.gv.toa, np.float64)
    g.set_scalar("radius", grid.gv.radius, np.float64)
    for m in DUAL_MEMBERS:
        dg.set(m, grid.dg[m])
    return g, dg


def state_struct(state):
    s = RefStruct("State")
    for m in ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind"):
        s.set(m, getattr(state, m))
    return s


def state_fields(s, dims):
    return {"rho": s.member("rho", np.float64, dims.n_const * dims.s),
            "rhotheta_v": s.member("rhotheta_v", np.float64, dims.s),
            "theta_v_pert": s.member("theta_v_pert", np.float64, dims.s),
            "exner_pert": s.member("exner_pert", np.float64, dims.s),
            "wind": s.member("wind", np.float64, dims.v)}


def config_struct(cfg):
    c = RefStruct("Config")
    c.buf[:] = np.frombuffer(bytes(cfg), dtype=np.uint8)
    return c
