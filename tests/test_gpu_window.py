"""The gradient of a window of dry RKHEVI steps on the device (the
hand-written adjoint kernels with checkpointing, include/game_rkhevi.h
game_gpu_window_*), against:

* autograd of the PyTorch restatement of the step (tests/torch_step.py,
  itself checked against the C oracle) on a small grid;
* central finite differences of the device's own forward, in 64 seeded
  directions: the assimilation workload, res 6 x 26 over the Gaussian
  mountain, 24 steps at dt = 120 s checkpointed every 4, cost
  J = 1/2 sum (rho_N - rho_target)^2 with rho_target the forward result of
  a second seeded initial perturbation (seed 20261203);
* the tangent-linear dot-product test: <M'(x) dx, y> from central
  differences of the forward against <dx, M'(x)^T y> from the adjoint, for
  random cotangents y of every output field;
* the window wrapped in a torch.autograd.Function (tests/adjoint_harness.py):
  gradients from torch.autograd.grad through it against central differences;
* the scale window: res 8 x 55 dry and flat, 12 steps at the automatic time
  step (45.04 s) checkpointed every 4, against central differences in 4
  directions, with the device memory in use reported.

Control variables: rho, rhotheta_v and wind; theta_v_pert and exner_pert are
diagnosed from rho and rhotheta_v by the equation of state.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import get_grid, oracle_state

pytestmark = pytest.mark.gpu

SEED_A, SEED_B = 20261202, 20261203


class Window:
    def __init__(self, grid):
        import torch
        from game_archived_amd.rkhevi import DeviceModel
        self.torch = torch
        self.g = grid
        self.d = grid.dims
        self.dm = DeviceModel(grid)
        self.lib = self.dm.lib

    def close(self):
        self.dm.close()

    def dev(self, a):
        return self.torch.as_tensor(np.ascontiguousarray(a)).cuda()

    @staticmethod
    def p(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    def forward(self, n, dt, x):
        t = self.torch
        d = self.d
        out = [t.zeros(d.s, dtype=t.float64, device="cuda") for _ in range(3)] + [
            t.zeros(d.v, dtype=t.float64, device="cuda")]
        self.dm._check(self.lib.game_gpu_window_forward(n, dt, *map(self.p, x), *map(self.p, out)))
        return out

    def vjp(self, n, dt, k, x, bars):
        t = self.torch
        d = self.d
        g = [t.zeros(d.s, dtype=t.float64, device="cuda"), t.zeros(d.s, dtype=t.float64, device="cuda"),
             t.zeros(d.v, dtype=t.float64, device="cuda")]
        self.dm._check(self.lib.game_gpu_window_vjp(n, dt, k, *map(self.p, x), *map(self.p, bars), *map(self.p, g)))
        return g

    def gradient(self, n, dt, k, x, target):
        t = self.torch
        d = self.d
        g = [t.zeros(d.s, dtype=t.float64, device="cuda"), t.zeros(d.s, dtype=t.float64, device="cuda"),
             t.zeros(d.v, dtype=t.float64, device="cuda")]
        cost = C.c_double(0.0)
        from game_archived_amd import _native as N
        self.dm._check(self.lib.game_gpu_window_gradient(n, dt, k, *map(self.p, x), self.p(target), C.byref(cost),
                                                         *map(self.p, g)))
        return cost.value, g


def controls(st):
    return st.rho, st.rhotheta_v, st.wind


def directions(grid, x, n, seed):
    """seeded random directions scaled to each control field: rho and
    rhotheta_v relative to themselves, the horizontal wind by its rms, the
    vertical wind by 1 cm/s"""
    d = grid.dims
    rng = np.random.default_rng(seed)
    k = np.arange(d.v) % d.vpl
    is_w = k < d.nsh
    u_rms = float(np.sqrt(np.mean(x[2][~is_w] ** 2)))
    out = []
    for _ in range(n):
        dr = x[0] * rng.standard_normal(d.s)
        dq = x[1] * rng.standard_normal(d.s)
        dw = np.where(is_w, 0.01, u_rms) * rng.standard_normal(d.v)
        out.append((dr, dq, dw))
    return out


def test_window_gradient_matches_torch_autograd():
    """4 steps over two checkpointed segments, res 3 mountain: the device
    vector-Jacobian product equals autograd's to 1e-9 in every control field
    (the two differ only in rounding order)."""
    import torch
    import torch_step as TS
    print("case: window vjp vs autograd, res 3 x 26 gaussian, 4 steps, checkpoint every 2", flush=True)
    g = get_grid(3, "gaussian")
    d = g.dims
    s0 = oracle_state(g)
    dt, n = 300.0, 4
    rng = np.random.default_rng(7)
    bars = [rng.standard_normal(d.s), rng.standard_normal(d.s), rng.standard_normal(d.s), rng.standard_normal(d.v)]
    G = TS.TorchGrid(g)
    wl = G.w_idx.ravel()
    bars[3][wl[:d.nsh]] = 0.0
    bars[3][wl[-d.nsh:]] = 0.0
    rho = torch.tensor(s0.rho.reshape(d.n_layers, d.nsh), requires_grad=True)
    rt = torch.tensor(s0.rhotheta_v.reshape(d.n_layers, d.nsh), requires_grad=True)
    wind = torch.tensor(s0.wind, requires_grad=True)
    u, w = G.split_v(wind)
    out = TS.run(G, TS.eos_state(G, rho, rt, u, w), dt, n)
    loss = ((out["rho"].reshape(-1) * torch.as_tensor(bars[0])).sum()
            + (out["rt"].reshape(-1) * torch.as_tensor(bars[1])).sum()
            + (out["ep"].reshape(-1) * torch.as_tensor(bars[2])).sum()
            + (G.join_v(out["u"], out["w"]) * torch.as_tensor(bars[3])).sum())
    loss.backward()
    W = Window(g)
    try:
        x = [W.dev(a) for a in controls(s0)]
        gr, gt, gw = (a.cpu().numpy() for a in W.vjp(n, dt, 2, x, [W.dev(b) for b in bars]))
    finally:
        W.close()
    for name, a, b in (("rho", gr, rho.grad.numpy().ravel()), ("rhotheta_v", gt, rt.grad.numpy().ravel()),
                       ("wind", gw, wind.grad.numpy())):
        err = float(np.abs(a - b).max() / np.abs(b).max())
        print(name, err, flush=True)
        assert err <= 1e-9, (name, err)


@pytest.mark.timeout(900)
def test_assimilation_window_gradient_matches_finite_differences():
    """res 6 x 26 dry over the Gaussian mountain, 24 steps at 120 s, checkpoints
    every 4 steps: <grad J, dx> against (J(x + eps dx) - J(x - eps dx))/(2 eps),
    eps = 1e-6, in 64 seeded directions, each within 1e-6 relative."""
    import torch
    print("case: window gradient vs central differences, res 6 x 26 gaussian, 24 steps, 64 directions",
          flush=True)
    g = get_grid(6, "gaussian", n_lloyd=2000)
    s0 = oracle_state(g, seed=SEED_A)
    sb = oracle_state(g, seed=SEED_B)
    dt, n, k = 120.0, 24, 4
    W = Window(g)
    try:
        xa = [W.dev(a) for a in controls(s0)]
        target = W.forward(n, dt, [W.dev(a) for a in controls(sb)])[0].clone()
        cost, grad = W.gradient(n, dt, k, xa, target)
        print("J", cost, flush=True)

        def J(x):
            r = W.forward(n, dt, x)[0]
            return 0.5 * float(torch.sum((r - target) ** 2))
        assert abs(J(xa) - cost) <= 1e-12 * cost
        eps = 1e-6
        worst = 0.0
        for i, dx in enumerate(directions(g, controls(s0), 64, 99)):
            dxd = [W.dev(a) for a in dx]
            gd = sum(float(torch.sum(a * b)) for a, b in zip(grad, dxd))
            jp = J([a + eps * b for a, b in zip(xa, dxd)])
            jm = J([a - eps * b for a, b in zip(xa, dxd)])
            fd = (jp - jm) / (2 * eps)
            err = abs(fd - gd) / abs(gd)
            worst = max(worst, err)
            if i < 4 or err > 1e-6:
                print(f"direction {i}: adjoint {gd:.12e} fd {fd:.12e} rel {err:.2e}", flush=True)
        print("worst relative difference over 64 directions", worst, flush=True)
        assert worst <= 1e-6
    finally:
        W.close()


def test_tangent_linear_dot_product():
    """<M'(x) dx, y> = <dx, M'(x)^T y> for random cotangents y of every output
    field (rho, rhotheta_v, exner_pert, wind) after 6 steps at res 4 over the
    mountain, checkpointed every 4 steps (a full and a partial segment); the
    tangent from central differences with eps = 1e-7."""
    import torch
    print("case: tangent-linear dot product, res 4 x 26 gaussian, 6 steps, checkpoint every 4", flush=True)
    g = get_grid(4, "gaussian")
    d = g.dims
    s0 = oracle_state(g)
    dt, n, k = 300.0, 6, 4
    W = Window(g)
    try:
        x = [W.dev(a) for a in controls(s0)]
        rng = np.random.default_rng(17)
        outs0 = W.forward(n, dt, x)
        scale = [float(torch.sqrt(torch.mean(o ** 2))) for o in outs0]
        ys = [W.dev(rng.standard_normal(o.numel()) / max(s, 1e-30)) for o, s in zip(outs0, scale)]
        wl = (np.arange(d.n_levels)[:, None] * d.vpl + np.arange(d.nsh)[None, :])
        ys[3][torch.as_tensor(np.concatenate([wl[0], wl[-1]])).cuda()] = 0.0   # top and surface w: pass-through
        xb = W.vjp(n, dt, k, x, ys)
        eps = 1e-7
        for i, dx in enumerate(directions(g, controls(s0), 3, 5)):
            dxd = [W.dev(a) for a in dx]
            op = W.forward(n, dt, [a + eps * b for a, b in zip(x, dxd)])
            om = W.forward(n, dt, [a - eps * b for a, b in zip(x, dxd)])
            tl = sum(float(torch.sum((p - m) / (2 * eps) * y)) for p, m, y in zip(op, om, ys))
            ad = sum(float(torch.sum(a * b)) for a, b in zip(xb, dxd))
            err = abs(tl - ad) / abs(ad)
            print(f"direction {i}: <M' dx, y> {tl:.12e} <dx, M'^T y> {ad:.12e} rel {err:.2e}", flush=True)
            assert err <= 1e-6, (i, err)
    finally:
        W.close()


def test_window_autograd_function_matches_finite_differences():
    """the window as a torch.autograd.Function: a scalar of all four output
    fields (weighted sums and a square), differentiated by torch.autograd.grad
    through the Function's backward (game_gpu_window_vjp), against central
    differences (eps = 1e-6) of the same scalar in 4 seeded directions,
    after 6 steps at res 4 over the mountain checkpointed every 4"""
    import torch
    from adjoint_harness import make_window
    print("case: window autograd.Function vs central differences, res 4 x 26 gaussian, 6 steps, checkpoint 4",
          flush=True)
    g = get_grid(4, "gaussian")
    d = g.dims
    s0 = oracle_state(g)
    dt, n, k = 300.0, 6, 4
    W = Window(g)
    try:
        x0 = [W.dev(a) for a in controls(s0)]
        rng = np.random.default_rng(23)
        outs0 = W.forward(n, dt, x0)
        scale = [float(torch.sqrt(torch.mean(o ** 2))) for o in outs0]
        ws = [W.dev(rng.standard_normal(o.numel()) / max(sc, 1e-30)) for o, sc in zip(outs0, scale)]
        wl = (np.arange(d.n_levels)[:, None] * d.vpl + np.arange(d.nsh)[None, :])
        ws[3][torch.as_tensor(np.concatenate([wl[0], wl[-1]])).cuda()] = 0.0   # top and surface w: pass-through
        target = outs0[0] * 1.001

        def scalar(outs):
            return sum(torch.sum(o * w) for o, w in zip(outs, ws)) + 0.5 * torch.sum((outs[0] - target) ** 2)

        window = make_window(W, n, dt, k)
        xs = [a.clone().requires_grad_(True) for a in x0]
        J = scalar(window(*xs))
        grads = torch.autograd.grad(J, xs)
        eps = 1e-6
        for i, dx in enumerate(directions(g, controls(s0), 4, 29)):
            dxd = [W.dev(a) for a in dx]
            jp = float(scalar(W.forward(n, dt, [a + eps * b for a, b in zip(x0, dxd)])))
            jm = float(scalar(W.forward(n, dt, [a - eps * b for a, b in zip(x0, dxd)])))
            fd = (jp - jm) / (2 * eps)
            ad = sum(float(torch.sum(a * b)) for a, b in zip(grads, dxd))
            err = abs(fd - ad) / abs(ad)
            print(f"direction {i}: fd {fd:.12e} autograd {ad:.12e} rel {err:.2e}", flush=True)
            assert err <= 1e-6, (i, err)
    finally:
        W.close()


@pytest.mark.timeout(900)
def test_scale_window_res8x55():
    """res 8 x 55 (655362 columns), dry, flat: the gradient of J over 12 steps
    at 45.04 s checkpointed every 4 steps against central differences
    (eps = 1e-6) in 4 seeded directions, each within 1e-6 relative."""
    import torch
    from game_archived_amd.grid import Grid
    print("case: scale window, res 8 x 55 dry flat, 12 steps, checkpoint every 4", flush=True)
    g = Grid(8, n_layers=55, n_oro_layers=49, moisture_on=0, n_lloyd=2000, rhombus_tolerance=1e-9,
             trsk_tolerance=1e-9)
    print("grid built", flush=True)
    s0 = oracle_state(g, seed=SEED_A)
    sb = oracle_state(g, seed=SEED_B)
    dt, n, k = 45.04, 12, 4
    W = Window(g)
    try:
        xa = [W.dev(a) for a in controls(s0)]
        target = W.forward(n, dt, [W.dev(a) for a in controls(sb)])[0].clone()
        cost, grad = W.gradient(n, dt, k, xa, target)
        free, total = torch.cuda.mem_get_info()
        print(f"J {cost}; device memory in use {(total - free) / 2**30:.1f} GiB of {total / 2**30:.0f}; "
              f"library {W.dm.bytes_allocated() / 2**30:.1f} GiB", flush=True)

        def J(x):
            r = W.forward(n, dt, x)[0]
            return 0.5 * float(torch.sum((r - target) ** 2))
        eps = 1e-6
        for i, dx in enumerate(directions(g, controls(s0), 4, 31)):
            dxd = [W.dev(a) for a in dx]
            gd = sum(float(torch.sum(a * b)) for a, b in zip(grad, dxd))
            fd = (J([a + eps * b for a, b in zip(xa, dxd)]) - J([a - eps * b for a, b in zip(xa, dxd)])) / (2 * eps)
            err = abs(fd - gd) / abs(gd)
            print(f"direction {i}: adjoint {gd:.12e} fd {fd:.12e} rel {err:.2e}", flush=True)
            assert err <= 1e-6, (i, err)
    finally:
        W.close()
