"""ctypes binding of libsmgpu.so (include/smgpu.h).

The library is loaded from the in-tree build directory; a missing library is an
error (there is no CPU fallback for any compute entry point).
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "build", "libsmgpu.so")

SMGPU_OK = 0
SMGPU_FP64 = 0
SMGPU_MIXED = 1

SMGPU_ECANCELED = -6
SMGPU_ETIMEDOUT = -7
ERRORS = {-1: "EINVAL", -2: "EIO", -3: "EDEVICE", -4: "ENOMEM", -5: "ENOTCONV", -6: "ECANCELED", -7: "ETIMEDOUT"}

# request states of the serving loop
REQ_STATES = {0: "queued", 1: "preparing", 2: "ready", 3: "running", 4: "done", 5: "failed", 6: "canceled"}


class SmgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"smgpu {ERRORS.get(code, code)}: {msg}")
        self.code = code


class ProblemInfo(ctypes.Structure):
    _fields_ = [("Z", ctypes.c_int32), ("N", ctypes.c_int32), ("A", ctypes.c_int32), ("vp", ctypes.c_int32),
                ("vn", ctypes.c_int32), ("twoM", ctypes.c_int32), ("lp", ctypes.c_int32), ("ln", ctypes.c_int32),
                ("cp", ctypes.c_int32), ("cn", ctypes.c_int32), ("nblocks", ctypes.c_int32),
                ("n_msps_p", ctypes.c_int32), ("n_msps_n", ctypes.c_int32), ("mdim", ctypes.c_int64),
                ("n_psd", ctypes.c_int64), ("n_nsd", ctypes.c_int64), ("n_pp_ops", ctypes.c_int64),
                ("n_nn_ops", ctypes.c_int64), ("n_pn_ops", ctypes.c_int64), ("n_hop_entries", ctypes.c_int64)]


class SolveParams(ctypes.Structure):
    _fields_ = [("num_ev", ctypes.c_int32), ("lm", ctypes.c_int32), ("ls", ctypes.c_int32),
                ("num_history", ctypes.c_int32), ("tol", ctypes.c_double), ("precision", ctypes.c_int32),
                ("refine", ctypes.c_int32), ("max_iter", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("reorth", ctypes.c_int32), ("steps", ctypes.c_int32)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("n_hv", ctypes.c_int32), ("n_hv64", ctypes.c_int32),
                ("n_restart", ctypes.c_int32), ("wall_ms", ctypes.c_double), ("ritz", ctypes.c_double * 64),
                ("n_reorth2", ctypes.c_int32), ("n_jproj", ctypes.c_int32), ("n_reorth", ctypes.c_int32),
                ("n_fallback", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_inflight", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("host_threads", ctypes.c_int32), ("table_threads", ctypes.c_int32), ("keep_bytes", ctypes.c_int64),
                ("max_batch", ctypes.c_int32), ("batch_variants", ctypes.c_int32), ("large_mdim", ctypes.c_int64),
                ("reserve_cus", ctypes.c_int32), ("reserved_", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("num_ev", ctypes.c_int32),
                ("index_shared", ctypes.c_int32), ("slot", ctypes.c_int32), ("lm", ctypes.c_int32),
                ("mdim", ctypes.c_int64), ("footprint", ctypes.c_int64), ("index_bytes", ctypes.c_int64),
                ("submit_ms", ctypes.c_double),
                ("prep_start_ms", ctypes.c_double), ("prep_end_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("solve", SolveStats)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("canceled", ctypes.c_int64), ("queued", ctypes.c_int32), ("preparing", ctypes.c_int32),
                ("ready", ctypes.c_int32), ("running", ctypes.c_int32), ("peak_running", ctypes.c_int32),
                ("slots", ctypes.c_int32), ("hbm_budget", ctypes.c_int64), ("hbm_in_use", ctypes.c_int64),
                ("hbm_peak", ctypes.c_int64), ("index_hits", ctypes.c_int64), ("index_misses", ctypes.c_int64),
                ("index_uploads", ctypes.c_int64), ("index_evictions", ctypes.c_int64), ("clock_ms", ctypes.c_double),
                ("batched", ctypes.c_int64), ("peak_batch", ctypes.c_int32), ("reserved_", ctypes.c_int32)]


# exported symbol -> (restype, argtypes)
P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
SIGNATURES = {
    "smgpu_version": (I32, []),
    "smgpu_last_error": (ctypes.c_char_p, []),
    "smgpu_device_count": (I32, [P]),
    "smgpu_problem_create": (I32, [ctypes.c_char_p, ctypes.c_char_p, I32, I32, I32, P]),
    "smgpu_problem_destroy": (I32, [P]),
    "smgpu_problem_get_info": (I32, [P, P]),
    "smgpu_problem_set_interaction": (I32, [P, ctypes.c_char_p]),
    "smgpu_basis_export": (I32, [P, P, P, P, P, P, P, P]),
    "smgpu_hops_export": (I32, [P, P, P, P, P, P, P, P, P, P]),
    "smgpu_pair_export": (I32, [P, I32, P, P, P, P, P]),
    "smgpu_ham_create": (I32, [P, I32, P]),
    "smgpu_ham_destroy": (I32, [P]),
    "smgpu_ham_create_j2": (I32, [P, I32, P]),
    "smgpu_ham_update_interaction": (I32, [P, P]),
    "smgpu_ham_create_variant": (I32, [P, P, P]),
    "smgpu_table_bytes": (I32, [P, I32, I32, P, P, ctypes.c_char_p, I32]),
    "smgpu_work_terms": (I32, [P, P]),
    "smgpu_lds_bytes": (I32, [P, I32, P]),
    "smgpu_apply_H": (I32, [P, I32, P, P, P]),
    "smgpu_apply_H_host": (I32, [P, I32, P, P]),
    "smgpu_time_apply_H": (I32, [P, I32, I32, P]),
    "smgpu_apply_H_multi": (I32, [P, I32, I32, P, I64, P, I64, P]),
    "smgpu_apply_H_host_multi": (I32, [P, I32, I32, P, P]),
    "smgpu_time_apply_H_multi": (I32, [P, I32, I32, I32, P]),
    "smgpu_apply_H_variants": (I32, [P, I32, I32, P, P, P]),
    "smgpu_apply_H_host_variants": (I32, [P, I32, I32, P, P]),
    "smgpu_time_apply_H_variants": (I32, [P, I32, I32, I32, P]),
    "smgpu_default_params": (None, [P]),
    "smgpu_trl_solve": (I32, [P, P, P, P, P]),
    "smgpu_trl_solve_v0": (I32, [P, P, P, P, P, P]),
    "smgpu_trl_solve_j": (I32, [P, P, I32, P, P, P, P, P]),
    "smgpu_trbl_solve": (I32, [P, P, I32, I32, P, P, P, P, P]),
    "smgpu_trl_solve_batch": (I32, [P, I32, P, P, P, P, P, P]),
    "smgpu_trl_solve_variants": (I32, [P, I32, P, P, P, P, P, P]),
    "smgpu_lanczos_steps": (I32, [P, I32, I32, P, P, P]),
    "smgpu_lanczos_steps_noalloc": (I32, [P, I32, I32, P, P, P]),
    "smgpu_lanczos_scratch_bytes": (I64, []),
    "smgpu_moments_create": (I32, [P, I32, P, P, ctypes.c_double, P]),
    "smgpu_moments_destroy": (I32, [P]),
    "smgpu_moments_apply": (I32, [P, I32, P, P, P]),
    "smgpu_moments_apply_host": (I32, [P, I32, I32, P, P]),
    "smgpu_eval_moment": (I32, [P, I32, P, P, P, P]),
    "smgpu_moments_bpar": (I32, [P, P]),
    "smgpu_transitions_create": (I32, [P, P, I32, P, P, ctypes.c_double, P]),
    "smgpu_transitions_destroy": (I32, [P]),
    "smgpu_transitions_apply": (I32, [P, I32, P, P, P]),
    "smgpu_transitions_apply_host": (I32, [P, I32, I32, P, P]),
    "smgpu_transit": (I32, [P, I32, I32, I32, P, I32, P, P, P]),
    "smgpu_transitions_info": (I32, [P, P, P, P, P]),
    "smgpu_malloc": (I32, [P, I64, P]),
    "smgpu_free": (I32, [P, P]),
    "smgpu_memcpy": (I32, [P, P, P, I64, I32]),
    "smgpu_synchronize": (I32, [P]),
    "smgpu_selftest_tridiag": (I32, [P, P, I32, I32, P, P]),
    "smgpu_selftest_symeig": (I32, [P, I32, I32, P, P]),
    "smgpu_selftest_restart": (I32, [I32, P, P, I32, I32, P, P, P]),
    "smgpu_default_server_config": (None, [P]),
    "smgpu_server_create": (I32, [P, P]),
    "smgpu_server_destroy": (I32, [P]),
    "smgpu_server_submit": (I32, [P, ctypes.c_char_p, ctypes.c_char_p, I32, P, P]),
    "smgpu_server_poll": (I32, [P, I64, P]),
    "smgpu_server_wait": (I32, [P, I64, ctypes.c_double]),
    "smgpu_server_result": (I32, [P, I64, P, P]),
    "smgpu_server_info": (I32, [P, I64, P]),
    "smgpu_server_cancel": (I32, [P, I64]),
    "smgpu_server_get_stats": (I32, [P, P]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsmgpu.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(code):
    if code != SMGPU_OK:
        raise SmgpuError(code, lib().smgpu_last_error().decode(errors="replace"))
    return code


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)
