"""The allocation-free three-band solve (game_gpu_op_thomas_noalloc) against
the entry point it varies (game_gpu_op_thomas, which takes a caller-allocated
work buffer): bit for bit on two calls in a row, the second smaller than the
first (the scratch keeps the first call's contents), and against LAPACK's
banded solver; sizes beyond the scratch are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg

from conftest import get_grid

pytestmark = pytest.mark.gpu


def test_noalloc_thomas_matches_the_workspace_entry_point():
    import torch
    from game_archived_amd.rkhevi import DeviceModel
    print("case: game_gpu_op_thomas_noalloc, (26 x 160000) then (7 x 3001)", flush=True)
    g = get_grid(3, "flat")
    dm = DeviceModel(g)
    lib = dm.lib
    p = lambda t: C.c_void_p(t.data_ptr())
    rng = np.random.default_rng(31)
    try:
        for n, ncol in ((26, 160000), (7, 3001)):
            c = rng.standard_normal((n, ncol))
            e = rng.standard_normal((n, ncol))
            d = 4.0 + np.abs(rng.standard_normal((n, ncol)))
            r = rng.standard_normal((n, ncol))
            dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (c, d, e, r)]
            x_ref = torch.zeros(n * ncol, dtype=torch.float64, device="cuda")
            work = torch.zeros(n * ncol, dtype=torch.float64, device="cuda")
            x_new = torch.full((n * ncol,), np.nan, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            dm._check(lib.game_gpu_op_thomas(n, ncol, *map(p, dev), p(x_ref), p(work)))
            dm._check(lib.game_gpu_op_thomas_noalloc(n, ncol, *map(p, dev), p(x_new)))
            a, b = x_new.cpu().numpy(), x_ref.cpu().numpy()
            assert np.array_equal(a, b), (n, ncol, np.abs(a - b).max())
            # the solve itself, a few columns against LAPACK
            xs = a.reshape(n, ncol)
            for k in (0, ncol // 2, ncol - 1):
                ab = np.zeros((3, n))
                ab[0, 1:] = e[:-1, k]
                ab[1] = d[:, k]
                ab[2, :-1] = c[:-1, k]
                assert np.allclose(xs[:, k], scipy.linalg.solve_banded((1, 1), ab, r[:, k]), rtol=1e-12,
                                   atol=1e-13)
        # beyond the scratch: rejected with a message, nothing launched
        small = torch.zeros(8, dtype=torch.float64, device="cuda")
        rc = lib.game_gpu_op_thomas_noalloc(64, (1 << 23) // 64 + 1, *([p(small)] * 5))
        assert rc == 2, rc
        assert b"scratch" in lib.game_gpu_last_error()
    finally:
        dm.close()
