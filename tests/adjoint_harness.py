"""torch.autograd.Function wrappers of the device operators and their
hand-written adjoints (verification harness only: the library itself has
no torch types; these pass device pointers through the C ABI).

Each Function's forward calls game_gpu_op_<name>, its backward calls
game_gpu_op_<name>_adjoint.  make_window wraps a whole window of steps the
same way: forward game_gpu_window_forward, backward game_gpu_window_vjp
(the hand-written adjoint of every kernel of the step, with checkpoints).  The library runs on its own stream and returns
after it has finished, so the wrappers synchronise torch's stream before
handing it device memory.
"""
import ctypes as C

import torch


class Ops:
    """the operator-level interface of an initialised DeviceModel (FP64)"""

    def __init__(self, dm):
        self.lib = dm.lib
        self.d = dm.dims

    def _run(self, name, *tensors):
        torch.cuda.synchronize()
        rc = getattr(self.lib, f"game_gpu_op_{name}")(*[C.c_void_p(t.data_ptr()) for t in tensors])
        if rc != 0:
            raise RuntimeError(f"{name}: {rc} {self.lib.game_gpu_last_error().decode()}")

    def new(self, n):
        return torch.zeros(int(n), dtype=torch.float64, device="cuda")

    def sizes(self):
        d = self.d
        return {"s": d.s, "v": d.v, "curl": d.curl, "tri": d.n_layers * d.nt, "edge": d.n_layers * d.nvh}


def make_linear(ops, name, n_out):
    """Function for an operator linear in its single input"""

    class Lin(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            y = ops.new(n_out)
            ops._run(name, x.contiguous(), y)
            ctx.n_in = x.numel()
            return y

        @staticmethod
        def backward(ctx, yb):
            xb = ops.new(ctx.n_in)
            ops._run(name + "_adjoint", yb.contiguous(), xb)
            return xb

    return Lin.apply


def make_bilinear_first(ops, name, n_out):
    """Function for an operator linear in its first input with the second
    held fixed (no gradient to it): vorticity flux for fixed PV, inner
    product for fixed b"""

    class Bil(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, fixed):
            y = ops.new(n_out)
            if name == "vorticity_flux":
                ops._run(name, x.contiguous(), fixed.contiguous(), y)
            else:
                ops._run(name, x.contiguous(), fixed.contiguous(), y)
            ctx.save_for_backward(fixed)
            ctx.n_in = x.numel()
            return y

        @staticmethod
        def backward(ctx, yb):
            (fixed,) = ctx.saved_tensors
            xb = ops.new(ctx.n_in)
            ops._run(name + "_adjoint", fixed.contiguous(), yb.contiguous(), xb)
            return xb, None

    return Bil.apply


def make_thomas(ops, n, ncol):
    """x = A^-1 r with gradients to r and to the three bands"""

    class Thomas(torch.autograd.Function):
        @staticmethod
        def forward(ctx, c, d, e, r):
            x = ops.new(n * ncol)
            w = ops.new(n * ncol)
            torch.cuda.synchronize()
            p = lambda t: C.c_void_p(t.data_ptr())
            rc = ops.lib.game_gpu_op_thomas(n, ncol, p(c), p(d), p(e), p(r), p(x), p(w))
            assert rc == 0, ops.lib.game_gpu_last_error()
            ctx.save_for_backward(c, d, e, x)
            return x

        @staticmethod
        def backward(ctx, xb):
            c, d, e, x = ctx.saved_tensors
            rb, cb, db, eb, w = (ops.new(n * ncol) for _ in range(5))
            torch.cuda.synchronize()
            p = lambda t: C.c_void_p(t.data_ptr())
            rc = ops.lib.game_gpu_op_thomas_adjoint(n, ncol, p(c), p(d), p(e), p(x), p(xb.contiguous()),
                                                    p(rb), p(cb), p(db), p(eb), p(w))
            assert rc == 0, ops.lib.game_gpu_last_error()
            return cb, db, eb, rb

    return Thomas.apply


def make_window(win, n_steps, dt, checkpoint_every):
    """(rho0, rhotheta_v0, wind0) -> (rho, rhotheta_v, exner_pert, wind)
    after n_steps dry steps; `win` has forward(n, dt, x) and
    vjp(n, dt, k, x, bars) over the C ABI (tests/test_gpu_window.py)"""

    class Window(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho0, rt0, wind0):
            x = [rho0.contiguous(), rt0.contiguous(), wind0.contiguous()]
            torch.cuda.synchronize()
            out = win.forward(n_steps, dt, x)
            ctx.save_for_backward(*x)
            return tuple(out)

        @staticmethod
        def backward(ctx, b_rho, b_rt, b_ex, b_wind):
            x = list(ctx.saved_tensors)
            bars = [b.contiguous() if b is not None else torch.zeros_like(o)
                    for b, o in zip((b_rho, b_rt, b_ex, b_wind), (x[0], x[1], x[0], x[2]))]
            torch.cuda.synchronize()
            return tuple(win.vjp(n_steps, dt, checkpoint_every, x, bars))

    return Window.apply
