"""The C ABI: struct layout identical to the reference's, and every entry
point declared in include/*.h exported by the libraries (no device needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import refharness as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "game_archived_amd", "lib")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:unsigned\s+)?(?:int|long long|long|double|void|char)\s*\**\s*"
                       r"(\w+)\s*\(", text, flags=re.M)
    return [n for n in names if not n.startswith("game_dims_derive")]


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_dropin_layout_equals_reference(tmp_path):
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-O0", "-I", os.path.join(ROOT, "include"), "-DGAME_RES_ID=5", "-DGAME_NO_OF_LAYERS=26",
                    "-DGAME_MOISTURE_ON=1", os.path.join(ROOT, "tests", "abi_layout.c"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout
    n = 0
    for line in out.splitlines():
        s, m, v = line.split()
        if m == "sizeof":
            assert R.lib().ref_sizeof(s.encode()) == int(v), (s, v)
        else:
            assert R.lib().ref_offsetof(s.encode(), m.encode()) == int(v), (s, m, v)
        n += 1
    assert n > 60


def test_rkhevi_library_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(LIB, "libgame_rkhevi.so"))
    names = declared("game_rkhevi.h")
    assert "game_gpu_manage_rkhevi" in names and len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("layout", ["r5l26m1", "r4l26m0", "r6l26m0"])
def test_dropin_library_exports_reference_entry_points(layout):
    C.CDLL(os.path.join(LIB, "libgame_rkhevi.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(os.path.join(LIB, f"libgame_dropin_{layout}.so"))
    names = declared("game_abi.h")
    for n in ("manage_rkhevi", "linear_combine_two_states", "temperature_diagnostics", "inner_product",
              "game_abi_fingerprint"):
        assert n in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.game_abi_fingerprint.restype = C.c_ulonglong
    assert lib.game_abi_fingerprint() != 0


def test_failed_init_leaves_no_model():
    """A game_gpu_init that fails (a physics switch outside the path, or no HIP
    device on this host) leaves no half-built model behind: every later entry
    point answers GAME_ERR_NOT_INIT instead of launching on missing buffers."""
    import numpy as np
    from conftest import get_grid
    from game_archived_amd import _native as N
    lib = N.gpu_lib()
    g = get_grid(3, "flat")
    bad = N.default_config()
    bad.rad_on = 1
    assert lib.game_gpu_init(g.dims, g.grid_view(), g.dualgrid_view(), bad, 0) == 5   # GAME_ERR_UNSUPPORTED
    assert b"physics switch" in lib.game_gpu_last_error()
    ok = N.default_config()
    rc = lib.game_gpu_init(g.dims, g.grid_view(), g.dualgrid_view(), ok, 7)
    assert rc == 2                                                                    # GAME_ERR_ARG: no such form
    out = np.zeros(g.dims.n_const + 4)
    assert lib.game_gpu_run_steps(1, 60.0, 1) == 3                                    # GAME_ERR_NOT_INIT
    assert lib.game_gpu_manage_rkhevi(60.0, 1) == 3
    assert lib.game_gpu_integrals(1, N.ptr(out)) == 3
    # every other entry point that needs a model, with null buffers: none may
    # reach them without one
    dt = C.c_double(60.0)
    calls = {
        "game_gpu_set_state": (1, None), "game_gpu_get_state": (1, None), "game_gpu_get_forcings": (None,),
        "game_gpu_get_diagnostics": (None,), "game_gpu_get_irrev": (None,), "game_gpu_set_config": (None,),
        "game_gpu_linear_combine_two_states": (0, 1, 1, C.c_double(0.5), C.c_double(0.5)),
        "game_gpu_inner_product": (None, None, None), "game_gpu_temperature_diagnostics": (1,),
        "game_gpu_write_integrals": (1, C.c_double(0.0), b"."),
        "game_gpu_run_steps_write_integrals": (1, dt, 1, C.c_double(0.0), b"."),
        "game_gpu_init_balance_terms": (None, None, None, None), "game_gpu_enable_timing": (1,),
        "game_gpu_reset_timing": (), "game_gpu_synchronize": (), "game_gpu_op_grad": (None, None),
        "game_gpu_op_grad_adjoint": (None, None), "game_gpu_op_thomas": (2, 2, None, None, None, None, None, None),
        "game_gpu_window_forward": (1, dt, None, None, None, None, None, None, None),
        "game_gpu_window_vjp": (1, dt, 1, None, None, None, None, None, None, None, None, None, None),
        "game_gpu_window_gradient": (1, dt, 1, None, None, None, None, None, None, None, None),
        "game_gpu_window_last_timing": (None,),
    }
    for name, args in calls.items():
        assert getattr(lib, name)(*args) == 3, name                                  # GAME_ERR_NOT_INIT
    assert lib.game_gpu_kernel_count() == 0
    assert lib.game_gpu_bytes_allocated() == 0
    assert lib.game_gpu_finalize() == 0


def test_host_library_exports_every_declared_file_function():
    lib = C.CDLL(os.path.join(LIB, "libgame_host.so"))
    names = declared("game_files.h") + ["game_nc_open"]
    assert "game_grid_file_read" in names and "game_state_file_write" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_dropin_set_grid_properties_reads_a_grid_file(tmp_path):
    """the drop-in's set_grid_properties (src/io/set_grid_properties.c:21-253)
    fills the reference's own struct grid and dualgrid (res 5 x 26, the
    reference's compiled layout) from a grid file: every member equals the
    generated grid's, the lat-lon interpolation and surface arrays included"""
    import numpy as np
    from conftest import get_grid
    g = get_grid(5, n_lloyd=2000)
    d = g.dims
    p = str(tmp_path / "grid.nc")
    g.write_file(p)
    extra = g.file_arrays()
    C.CDLL(os.path.join(LIB, "libgame_host.so"), mode=C.RTLD_GLOBAL)
    C.CDLL(os.path.join(LIB, "libgame_rkhevi.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(os.path.join(LIB, "libgame_dropin_r5l26m1.so"))
    lib.set_grid_properties.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p]
    G, DG = R.RefStruct("Grid"), R.RefStruct("Dualgrid")
    assert lib.set_grid_properties(G.ptr, DG.ptr, p.encode()) == 0
    for m in R.GRID_MEMBERS:
        a = g.g[m]
        assert np.array_equal(G.member(m, a.dtype, a.size), a), m
    for m in R.DUAL_MEMBERS:
        a = g.dg[m][:d.dual_v] if m == "area" else g.dg[m]
        assert np.array_equal(DG.member(m, a.dtype, a.size), a), m
    assert G.member("no_of_oro_layers", np.int32, 1)[0] == 23
    assert G.member("toa", np.float64, 1)[0] == 41152.0
    assert np.array_equal(G.member("latlon_interpol_indices", np.int32, extra["interpol_indices"].size),
                          extra["interpol_indices"])
    assert np.array_equal(G.member("latlon_interpol_weights", np.float64, extra["interpol_weights"].size),
                          extra["interpol_weights"])
    assert np.array_equal(G.member("t_conduc_soil", np.float64, d.nsh), extra["t_conductivity"])
    z = G.member("z_soil_interface", np.float64, 6)
    assert z[0] == 0.0 and abs(z[5] + 10.0) < 1e-12 and (np.diff(z) < 0).all()
    assert G.member("z_t_const", np.float64, 1)[0] == -10.0
