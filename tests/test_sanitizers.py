"""The host C code (grid generator, initial state, file formats) and the
oracle under AddressSanitizer and UndefinedBehaviorSanitizer: both are
rebuilt with -fsanitize=address,undefined into a scratch directory and a
child process runs a moist and a dry workload through them (grid generation
with orography, the lat-lon interpolation, grid and state files written and
read, the balanced initial state, the restart equation of state, oracle
steps).  Any sanitizer report makes the child fail.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "game_archived_amd", "csrc")
FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fopenmp", "-ffp-contract=off", "-fPIC", "-std=gnu99", "-shared"]

CHILD = r"""
import os, sys, tempfile
import numpy as np
sys.path.insert(0, os.path.join(sys.argv[1], "tests")); sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
import pyoracle
from game_archived_amd.grid import Grid
from game_archived_amd.state import initial_state, write_state_file, read_state_file
for moist, oro in ((1, "gaussian"), (0, "flat")):
    g = Grid(3, n_layers=26, n_oro_layers=23, moisture_on=moist, n_lloyd=3, orography=oro, cache=False)
    d = g.dims
    with tempfile.TemporaryDirectory() as t:
        g.write_file(os.path.join(t, "grid.nc"))
        h = Grid.from_file(os.path.join(t, "grid.nc"), 3, 26, 23, moist)
        assert all(np.array_equal(g.g[k], h.g[k]) for k in g.g)
        s0 = initial_state(g, pyoracle.balance_terms(g), moist=bool(moist))
        T = np.zeros(d.s)
        pyoracle.lib().or_temperature_diagnostics(pyoracle.ctx(g), s0.view(), pyoracle.N.ptr(T))
        write_state_file(os.path.join(t, "state.nc"), g, s0, T)
        st, T2, _ = read_state_file(os.path.join(t, "state.nc"), g)
        assert np.array_equal(T, T2)
        m = pyoracle.OracleModel(g, s0, 300.0)
        m.run(2)
        assert np.isfinite(m.new.rho).all() and np.isfinite(m.new.wind).all()
maps = open("/proc/self/maps").read()
assert os.environ["GAME_LIB_DIR"] in maps and os.environ["GAME_ORACLE_LIB"] in maps and "libasan" in maps
print("sanitized run finished")
"""


@pytest.mark.skipif(bool(os.environ.get("LD_PRELOAD")), reason="another library is preloaded in this process tree")
@pytest.mark.timeout(900)
def test_host_code_and_oracle_under_asan_and_ubsan(tmp_path):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True).stdout.strip()
    if not os.path.isabs(asan):
        pytest.skip("gcc has no libasan")
    lib = tmp_path / "lib"
    lib.mkdir()
    subprocess.run(["gcc", *FLAGS, "-o", str(lib / "libgame_host.so"),
                    *(os.path.join(CSRC, f) for f in ("gridgen.c", "init_state.c", "ncio.c", "gridfile.c")), "-lm"],
                   check=True)
    subprocess.run(["gcc", *FLAGS, "-Wall", "-o", str(tmp_path / "liboracle.so"),
                    os.path.join(ROOT, "oracle", "oracle_rkhevi.c"), "-lm"], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, GAME_LIB_DIR=str(lib), GAME_ORACLE_LIB=str(tmp_path / "liboracle.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""), OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=800)
    report = r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, report
    assert "sanitized run finished" in r.stdout
