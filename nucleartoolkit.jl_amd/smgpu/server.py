"""Persistent serving loop of libsmgpu (include/smgpu.h, smgpu_server_*).

Each request is one `main_sm` eigenvalue solve (shellmodel_main.jl:89-281,
target_J = []): interaction file + nucleus + solver parameters in, the lowest
eigenvalues (`en[1]`) out.  The library prepares requests on host threads,
admits them by HBM footprint as device slots free up, and runs admitted solves
concurrently, one HIP stream per slot; requests on the same basis share one
device copy of the index tables.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import (REQ_STATES, SMGPU_ETIMEDOUT, RequestInfo, ServerConfig, ServerStats, SmgpuError, SolveParams,
                   check, lib, ptr)
from .shellmodel import _prec


def solve_params(num_ev, lm=300, ls=20, tol=1e-8, num_history=3, precision="mixed", refine=True, seed=1234,
                 max_iter=0, reorth="default"):
    p = SolveParams()
    lib().smgpu_default_params(ctypes.byref(p))
    p.num_ev = int(num_ev); p.lm = int(lm); p.ls = int(ls); p.tol = float(tol)
    p.num_history = int(num_history); p.precision = _prec(precision); p.refine = int(bool(refine))
    p.seed = int(seed); p.max_iter = int(max_iter)
    p.reorth = {"default": 0, "full": 1, "partial": 2}[reorth]
    return p


def _info_dict(info: RequestInfo):
    s = info.solve
    return dict(state=REQ_STATES.get(info.state, info.state), status=info.status, num_ev=info.num_ev,
                index_shared=bool(info.index_shared), slot=info.slot, lm=info.lm, mdim=info.mdim, footprint=info.footprint,
                index_bytes=info.index_bytes,
                submit_ms=info.submit_ms, prep_start_ms=info.prep_start_ms, prep_end_ms=info.prep_end_ms,
                admit_ms=info.admit_ms, done_ms=info.done_ms, latency_ms=info.latency_ms,
                iterations=s.iterations, n_hv=s.n_hv, n_hv64=s.n_hv64, n_restart=s.n_restart, solve_ms=s.wall_ms,
                n_reorth2=s.n_reorth2, n_reorth=s.n_reorth, n_fallback=s.n_fallback)


class Server:
    """smgpu_server handle.  Use as a context manager or call close()."""

    def __init__(self, device=0, max_inflight=4, hbm_budget=0, host_threads=2, table_threads=0, keep_bytes=0,
                 max_batch=1, batch_variants=True, large_mdim=0, reserve_cus=0):
        """large_mdim > 0: requests of at least that dimension run one at a time, the others
        smallest first; reserve_cus > 0 puts the large ones in slot 0, on the other CUs while
        smaller requests are in the server."""
        c = ServerConfig()
        lib().smgpu_default_server_config(ctypes.byref(c))
        c.device = int(device); c.max_inflight = int(max_inflight); c.hbm_budget = int(hbm_budget)
        c.host_threads = int(host_threads); c.table_threads = int(table_threads); c.keep_bytes = int(keep_bytes)
        c.max_batch = int(max_batch); c.batch_variants = int(bool(batch_variants))
        c.large_mdim = int(large_mdim); c.reserve_cus = int(reserve_cus)
        h = ctypes.c_void_p()
        check(lib().smgpu_server_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h

    def submit(self, sntf, target_nuc, num_ev, twoM=-1, **params):
        """Queue one solve; returns the request id at once."""
        p = solve_params(num_ev, **params)
        rid = ctypes.c_int64()
        check(lib().smgpu_server_submit(self._h, os.fsencode(sntf), target_nuc.encode(), int(twoM),
                                        ctypes.byref(p), ctypes.byref(rid)))
        return int(rid.value)

    def poll(self, rid):
        s = ctypes.c_int32()
        check(lib().smgpu_server_poll(self._h, int(rid), ctypes.byref(s)))
        return REQ_STATES[s.value]

    def wait(self, rid, timeout_ms=-1.0):
        """True when the request has finished, False on timeout."""
        c = lib().smgpu_server_wait(self._h, int(rid), float(timeout_ms))
        if c == SMGPU_ETIMEDOUT:
            return False
        check(c)
        return True

    def info(self, rid):
        info = RequestInfo()
        check(lib().smgpu_server_info(self._h, int(rid), ctypes.byref(info)))
        return _info_dict(info)

    def result(self, rid, raise_on_error=True):
        """(evals, info dict) of a finished request; drops its record."""
        info = RequestInfo()
        ev = np.zeros(64)
        c = lib().smgpu_server_result(self._h, int(rid), ptr(ev), ctypes.byref(info))
        err = lib().smgpu_last_error().decode(errors="replace") if c != 0 else ""
        if c != 0 and raise_on_error:
            raise SmgpuError(c, err)
        d = _info_dict(info)
        d["error"] = err
        return ev[:info.num_ev].copy(), d

    def cancel(self, rid):
        check(lib().smgpu_server_cancel(self._h, int(rid)))

    def stats(self):
        s = ServerStats()
        check(lib().smgpu_server_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().smgpu_server_destroy(h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()
