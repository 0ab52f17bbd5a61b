"""Host mirror of the reference's step interface over the device library.

`DeviceModel` is the device-resident analogue of the reference driver's
structs (src/coordinator.c:36-46): `manage_rkhevi` (src/time_stepping/
manage_rkhevi.c:20-100), `linear_combine_two_states`
(src/spatial_operators/linear_combine_two_states.c:14-43), the time loop
(src/coordinator.c:302-324) and the integrals (src/io/write_output.c:1515-1596)
run on the GPU through include/game_rkhevi.h.  There is no CPU fallback: the
constructor raises when the HIP library or a device is missing.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .state import Diagnostics, Forcings, Irrev, State

FP64, MIXED = 0, 1
OLD, NEW, TENDENCY = 0, 1, 2


class GameError(RuntimeError):
    pass


class DeviceModel:
    def __init__(self, grid, config=None, precision=FP64):
        self.lib = N.gpu_lib()
        self.grid = grid
        self.dims = grid.dims
        self.config = config if config is not None else N.default_config()
        self.first_step = self.config.totally_first_step_bool
        self._check(self.lib.game_gpu_init(self.dims, grid.grid_view(), grid.dualgrid_view(), self.config,
                                           precision))
        self.precision = precision

    def _check(self, rc):
        if rc != 0:
            raise GameError(f"device call failed ({rc}): {self.lib.game_gpu_last_error().decode()}")

    def close(self):
        self.lib.game_gpu_finalize()

    # -- states -----------------------------------------------------------
    def set_state(self, slot, state):
        self._check(self.lib.game_gpu_set_state(slot, state.view()))

    def get_state(self, slot=NEW):
        st = State(self.dims)
        self._check(self.lib.game_gpu_get_state(slot, st.view()))
        return st

    def get_forcings(self):
        f = Forcings(self.dims)
        self._check(self.lib.game_gpu_get_forcings(f.view()))
        return f

    def get_diagnostics(self):
        d = Diagnostics(self.dims)
        self._check(self.lib.game_gpu_get_diagnostics(d.view()))
        return d

    def get_irrev(self):
        i = Irrev(self.dims)
        self._check(self.lib.game_gpu_get_irrev(i.view()))
        return i

    def start(self, state0):
        """The driver's preparation (src/coordinator.c:250, 292-299): old =
        state0, its temperature diagnosed (the diagnostics temperature the
        initial integrals read), new = 1*old + 0*old (which re-derives
        theta_v_pert), first-step flag set."""
        self.set_state(OLD, state0)
        self.temperature_diagnostics(OLD)
        self.linear_combine_two_states(OLD, OLD, NEW, 1.0, 0.0)
        self.first_step = 1

    def temperature_diagnostics(self, slot):
        """temperature_diagnostics (src/constituents/derived_quantities.c:15-40)
        of a slot into the diagnostics temperature"""
        self._check(self.lib.game_gpu_temperature_diagnostics(slot))

    # -- stepping ---------------------------------------------------------
    def manage_rkhevi(self, dt):
        self._check(self.lib.game_gpu_manage_rkhevi(dt, self.first_step))
        self.first_step = 0

    def linear_combine_two_states(self, s0, s1, out, c0, c1):
        self._check(self.lib.game_gpu_linear_combine_two_states(s0, s1, out, c0, c1))

    def run(self, n_steps, dt):
        self._check(self.lib.game_gpu_run_steps(n_steps, dt, self.first_step))
        if n_steps > 0:
            self.first_step = 0

    def run_writing_integrals(self, n_steps, dt, t_start, directory):
        """the time loop with the reference's integral files (src/coordinator.c:302-336)"""
        self._check(self.lib.game_gpu_run_steps_write_integrals(n_steps, dt, self.first_step, t_start,
                                                                 directory.encode()))
        if n_steps > 0:
            self.first_step = 0

    def write_integrals(self, time_since_init, directory, slot=NEW):
        self._check(self.lib.game_gpu_write_integrals(slot, time_since_init, directory.encode()))

    def synchronize(self):
        self._check(self.lib.game_gpu_synchronize())

    def integrals(self, slot=NEW):
        """write_out_integral's sums (src/io/write_output.c:1515-1596) of a slot:
        masses per constituent, rhotheta_v, kinetic, potential and internal
        energy; the internal energy multiplies by the diagnostics temperature
        (of the state the last step started from), as the reference does"""
        out = np.zeros(self.dims.n_const + 4)
        self._check(self.lib.game_gpu_integrals(slot, N.ptr(out)))
        return out

    # -- timing -----------------------------------------------------------
    def enable_timing(self, on=True):
        self._check(self.lib.game_gpu_enable_timing(1 if on else 0))

    def reset_timing(self):
        self._check(self.lib.game_gpu_reset_timing())

    def kernel_stats(self):
        self.synchronize()
        out = {}
        for k in range(self.lib.game_gpu_kernel_count()):
            name, n, ms = C.c_char_p(), C.c_longlong(), C.c_double()
            self._check(self.lib.game_gpu_kernel_stats(k, C.byref(name), C.byref(n), C.byref(ms)))
            out[name.value.decode()] = (n.value, ms.value)
        return out

    def bytes_allocated(self):
        return self.lib.game_gpu_bytes_allocated()

    def balance_terms(self):
        """(v_squared, pot_vort_tend) of a (wind, density) pair on the device,
        for the initial hydrostatic balance."""
        d = self.dims

        def f(wind, density):
            vsq = np.zeros(d.s)
            pvt = np.zeros(d.v)
            self._check(self.lib.game_gpu_init_balance_terms(N.ptr(np.ascontiguousarray(wind)),
                                                             N.ptr(np.ascontiguousarray(density)),
                                                             N.ptr(vsq), N.ptr(pvt)))
            return vsq, pvt
        return f
