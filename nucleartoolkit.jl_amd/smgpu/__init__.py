"""smgpu — MI355X-native M-scheme shell-model eigensolver (host-side API).

The compute path is libsmgpu.so (HIP, gfx950); see include/smgpu.h.
"""
from ._lib import SMGPU_FP64, SMGPU_MIXED, SmgpuError, lib  # noqa: F401
from .server import Server  # noqa: F401
from .shellmodel import (Hamiltonian, Moments, Problem, format_moments, main_sm, operate_H_variants,  # noqa: F401
                         time_operate_H_variants, trl_variants)
