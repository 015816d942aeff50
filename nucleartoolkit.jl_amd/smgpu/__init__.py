"""smgpu — MI355X-native M-scheme shell-model eigensolver (host-side API).

The compute path is libsmgpu.so (HIP, gfx950); see include/smgpu.h.
"""
from ._lib import SMGPU_FP64, SMGPU_MIXED, SmgpuError, lib  # noqa: F401
from .server import Server  # noqa: F401
from .shellmodel import (Hamiltonian, Moments, Problem, Transitions, format_moments, main_sm,  # noqa: F401
                         operate_H_variants, time_operate_H_variants, transit_main, trl_variants)
