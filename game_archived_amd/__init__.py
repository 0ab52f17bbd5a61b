"""MI355X-native RKHEVI dynamical-core step of GAME (MHBalsmeier/GAME_archived).

Host side: grid generation and initial states (C), device side: the RKHEVI
time step as HIP kernels for gfx950 behind the C ABI of include/game_rkhevi.h.
"""
from ._native import make_dims, default_config  # noqa: F401
