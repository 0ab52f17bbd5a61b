"""Grid construction (host) for the RKHEVI step.

`Grid` owns numpy arrays for every member of the reference's struct grid and
struct dualgrid (src/game_types.h:70-126) and fills them with the host C
generator (csrc/gridgen.c), which restates grid_generator/src/*.c and the
derived metrics of src/io/set_grid_properties.c:196-221.

The Lloyd iterations that place the generators dominate the cost of building
a grid (O(N) per iteration, 2000 iterations in the reference's run script);
their result depends only on (res_id, iterations) and is cached as a .npy
file of lat/lon pairs, the analogue of the reference's
use_scalar_h_coords_file option (grid_generator/run_script.sh:14-15).
"""
import math
import os
import time

import numpy as np

from . import _native as N

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def generators(res_id, n_lloyd, cache=True):
    """Generator coordinates after n_lloyd Lloyd iterations, (NSH, 2) lat/lon."""
    path = os.path.join(DATA_DIR, f"generators_res{res_id}_lloyd{n_lloyd}.npy")
    if cache and os.path.exists(path):
        return np.ascontiguousarray(np.load(path, allow_pickle=False))
    nsh = 12 + 10 * ((1 << (2 * res_id)) - 1)
    out = np.zeros((nsh, 2))
    rc = N.host_lib().game_grid_generators(res_id, n_lloyd, N.ptr(out))
    if rc != 0:
        raise RuntimeError(f"generator construction failed: {N.host_lib().game_grid_last_error().decode()}")
    if cache:
        os.makedirs(DATA_DIR, exist_ok=True)
        np.save(path + ".tmp.npy", out)
        os.replace(path + ".tmp.npy", path)
    return out


class Grid:
    """One grid: sizes (`dims`), struct grid (`g`), struct dualgrid (`dg`) and
    the extra grid-file arrays (`x`), as dicts of numpy arrays."""

    def __init__(self, res_id, n_layers=26, n_oro_layers=23, moisture_on=0, n_lloyd=2000,
                 toa=41152.0, stretching=1.3, orography="flat", oro_h0=2000.0, oro_half_width=1500e3,
                 oro_lat_deg=30.0, oro_lon_deg=90.0, radius_rescale=1.0, cache=True, rhombus_tolerance=0.0,
                 trsk_tolerance=0.0):
        t0 = time.time()
        self._alloc(res_id, n_layers, n_oro_layers, moisture_on)
        d = self.dims
        gen = generators(res_id, n_lloyd, cache=cache)
        p = N.GridParams()
        p.res_id, p.n_layers, p.n_oro_layers, p.n_lloyd_iterations = res_id, n_layers, n_oro_layers, n_lloyd
        p.stretching_parameter, p.toa, p.radius_rescale = stretching, toa, radius_rescale
        p.oro_id = {"flat": 0, "gaussian": 2}[orography]
        p.oro_h0, p.oro_half_width = oro_h0, oro_half_width
        p.oro_lat, p.oro_lon = math.radians(oro_lat_deg), math.radians(oro_lon_deg)
        p.generator_latlon = N.ptr(gen)
        # 0: the reference's bound on the rhombus weight sums (1e-10)
        p.rhombus_tolerance = rhombus_tolerance
        # 0: the reference's bound on the TRSK kite-area and antisymmetry checks (1e-10)
        p.trsk_tolerance = trsk_tolerance
        self._gen = gen
        self.orography = orography
        self.n_lloyd = n_lloyd
        self.gv.no_of_oro_layers = n_oro_layers
        rc = N.host_lib().game_grid_generate(p, d, self.gv, self.dgv, self.extras_view())
        if rc != 0:
            raise RuntimeError(f"grid generation failed ({rc}): {N.host_lib().game_grid_last_error().decode()}")
        self.build_seconds = time.time() - t0

    def _alloc(self, res_id, n_layers, n_oro_layers, moisture_on):
        self.dims = N.make_dims(res_id, n_layers, moisture_on, n_oro_layers)
        d = self.dims
        f8 = lambda n: np.zeros(int(n), dtype=np.float64)
        i4 = lambda n: np.zeros(int(n), dtype=np.int32)
        self.g = {
            "normal_distance": f8(d.v), "volume": f8(d.s), "area": f8(d.v), "z_scalar": f8(d.s),
            "z_vector": f8(d.v), "gravity_potential": f8(d.s), "gravity_m": f8(d.v), "slope": f8(d.v),
            "theta_v_bg": f8(d.s), "exner_bg": f8(d.s), "exner_bg_grad": f8(d.v), "layer_thickness": f8(d.s),
            "trsk_indices": i4(10 * d.nvh), "trsk_modified_curl_indices": i4(10 * d.nvh),
            "from_index": i4(d.nvh), "to_index": i4(d.nvh), "adjacent_vector_indices_h": i4(6 * d.nsh),
            "adjacent_signs_h": i4(6 * d.nsh), "density_to_rhombi_indices": i4(4 * d.nvh),
            "latitude_scalar": f8(d.nsh), "longitude_scalar": f8(d.nsh), "inner_product_weights": f8(8 * d.s),
            "direction": f8(d.nvh), "density_to_rhombi_weights": f8(4 * d.nvh), "trsk_weights": f8(10 * d.nvh),
            "is_land": i4(d.nsh),
        }
        self.dg = {
            "area": f8(d.curl), "z_vector": f8(d.dual_v), "normal_distance": f8(d.dual_v),
            "from_index": i4(d.nvh), "to_index": i4(d.nvh), "vorticity_indices_triangles": i4(3 * d.nt),
            "vorticity_signs_triangles": i4(3 * d.nt), "f_vec": f8(2 * d.nvh),
        }
        self.x = {
            "latitude_vector": f8(d.nvh), "longitude_vector": f8(d.nvh), "latitude_scalar_dual": f8(d.nt),
            "longitude_scalar_dual": f8(d.nt), "direction_dual": f8(d.nvh), "pent_hex_face_unity_sphere": f8(d.nsh),
            "triangle_face_unit_sphere": f8(d.nt), "oro": f8(d.nsh), "vorticity_indices_rhombi": i4(4 * d.nvh),
        }
        self.gv = N.fill_view(N.GridView(), self.g)
        self.gv.no_of_oro_layers = n_oro_layers
        self.dgv = N.fill_view(N.DualgridView(), self.dg)

    def grid_view(self):
        return self.gv

    def dualgrid_view(self):
        return self.dgv

    def extras_view(self):
        return N.fill_view(N.GridExtras(), self.x)

    def file_arrays(self):
        """the grid file's arrays that the generator does not keep: the
        lat-lon output interpolation (grid_generator/src/interpolate_ll.c)
        and the surface properties of cells that are not land"""
        d = self.dims
        lib = N.host_lib()
        nll = 8 << (2 * d.res_id)
        a = {"interpol_indices": np.zeros(5 * nll, dtype=np.int32), "interpol_weights": np.zeros(5 * nll)}
        for k in ("sfc_albedo", "sfc_rho_c", "t_conductivity", "roughness_length"):
            a[k] = np.zeros(d.nsh)
        g = self.g
        if lib.game_latlon_interpolation(d.res_id, N.ptr(g["latitude_scalar"]), N.ptr(g["longitude_scalar"]),
                                         N.ptr(g["adjacent_vector_indices_h"]), N.ptr(g["from_index"]),
                                         N.ptr(g["to_index"]), N.ptr(a["interpol_indices"]),
                                         N.ptr(a["interpol_weights"])) != 0:
            raise RuntimeError(lib.game_grid_last_error().decode())
        if lib.game_sfc_properties(d.nsh, N.ptr(g["is_land"]), N.ptr(a["sfc_albedo"]), N.ptr(a["sfc_rho_c"]),
                                   N.ptr(a["t_conductivity"]), N.ptr(a["roughness_length"])) != 0:
            raise RuntimeError(lib.game_grid_last_error().decode())
        return a

    def write_file(self, path):
        """the reference's grid file (grid_generator/src/coordinator.c:354-611)"""
        a = self.file_arrays()
        x = N.fill_view(N.GridFileExtras(), {**self.x, **a})
        x.n_lloyd_iterations = self.n_lloyd
        rc = N.host_lib().game_grid_file_write(os.fsencode(path), self.dims, self.gv, self.dgv, x)
        if rc != 0:
            raise RuntimeError(f"grid file write failed ({rc}): {N.host_lib().game_nc_last_error().decode()}")

    def write_statistics(self, path, oro_id=None):
        """the generator's statistics file (grid_generator/src/miscellaneous.c:92-135),
        named RES<r>_L<l>_ORO<o> inside as the generator names it"""
        d = self.dims
        if oro_id is None:
            oro_id = {"flat": 0, "gaussian": 2}.get(self.orography, 0)
        name = f"RES{d.res_id}_L{d.n_layers}_ORO{oro_id}".encode()
        rc = N.host_lib().game_grid_statistics_write(os.fsencode(path), name, int(self.n_lloyd), d.nsh, d.nvh,
                                                     N.ptr(self.x["pent_hex_face_unity_sphere"]),
                                                     N.ptr(self.g["normal_distance"]), N.ptr(self.dg["normal_distance"]))
        if rc != 0:
            raise RuntimeError(f"statistics file write failed ({rc})")

    @classmethod
    def from_file(cls, path, res_id, n_layers=26, n_oro_layers=23, moisture_on=0):
        """a grid read as src/io/set_grid_properties.c:21-221 reads it; the
        file's extra arrays are in .file (interpolation, surface properties)"""
        self = cls.__new__(cls)
        t0 = time.time()
        self._alloc(res_id, n_layers, n_oro_layers, moisture_on)
        d = self.dims
        nll = 8 << (2 * d.res_id)
        self.file = {"interpol_indices": np.zeros(5 * nll, dtype=np.int32), "interpol_weights": np.zeros(5 * nll)}
        for k in ("sfc_albedo", "sfc_rho_c", "t_conductivity", "roughness_length"):
            self.file[k] = np.zeros(d.nsh)
        x = N.fill_view(N.GridFileExtras(), {**self.x, **self.file})
        rc = N.host_lib().game_grid_file_read(os.fsencode(path), d, self.gv, self.dgv, x)
        if rc != 0:
            raise RuntimeError(f"grid file read failed ({rc}): {N.host_lib().game_nc_last_error().decode()}")
        self.n_lloyd = x.n_lloyd_iterations
        self.orography = "file"
        self._gen = None
        self.build_seconds = time.time() - t0
        return self
