"""Host-side mirror of NuclearToolkit.jl's shell-model entry points over libsmgpu.

* `main_sm(sntf, target_nuc, num_ev, target_J; lm, ls, tol, ...)` mirrors
  src/ShellModel/shellmodel_main.jl:89-281 for the eigenvalue observable
  (`en[1]`), with the H·v and Lanczos on the GPU.
* `Problem` is the host part (readsmsnt -> occ -> operator and hop tables),
  `Hamiltonian` the device-resident tables; `Hamiltonian.operate_H` mirrors
  operate_H! (src/ShellModel/lanczos_methods.jl:153-248) as w = H v and
  `Hamiltonian.trl` mirrors TRL (lanczos_methods.jl:1-91).
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

from ._lib import (SMGPU_FP64, SMGPU_MIXED, ProblemInfo, SolveParams, SolveStats, check, lib, ptr)

PRECISIONS = {"fp64": SMGPU_FP64, "mixed": SMGPU_MIXED}


def _prec(p):
    if isinstance(p, str):
        return PRECISIONS[p]
    return int(p)


class Problem:
    """Interaction + M-scheme basis + host tables for one nucleus and 2M."""

    def __init__(self, sntf, target_nuc, twoM=-1, ignore_massop=False, nthreads=0):
        h = ctypes.c_void_p()
        check(lib().smgpu_problem_create(os.fsencode(sntf), target_nuc.encode(), int(twoM),
                                         int(bool(ignore_massop)), int(nthreads), ctypes.byref(h)))
        self._h = h
        self.sntf = sntf
        self.nucleus = target_nuc
        self.info = self.get_info()

    def get_info(self):
        info = ProblemInfo()
        check(lib().smgpu_problem_get_info(self._h, ctypes.byref(info)))
        return info

    @property
    def mdim(self):
        return int(self.info.mdim)

    def set_interaction(self, sntf):
        check(lib().smgpu_problem_set_interaction(self._h, os.fsencode(sntf)))
        self.sntf = sntf

    def basis(self):
        """dict(tdims, Mp, Mn, pbits[list per block], nbits[list per block])"""
        i = self.info
        nb = i.nblocks
        tdims = np.zeros(nb + 1, np.int64)
        Mp = np.zeros(nb, np.int32)
        Mn = np.zeros(nb, np.int32)
        pb = np.zeros(max(1, i.n_psd), np.int64)
        nbits = np.zeros(max(1, i.n_nsd), np.int64)
        po = np.zeros(nb + 1, np.int64)
        no = np.zeros(nb + 1, np.int64)
        check(lib().smgpu_basis_export(self._h, ptr(tdims), ptr(Mp), ptr(Mn), ptr(pb), ptr(po), ptr(nbits),
                                       ptr(no)))
        return dict(tdims=tdims, Mp=Mp, Mn=Mn,
                    pbits=[pb[po[k]:po[k + 1]] for k in range(nb)],
                    nbits=[nbits[no[k]:no[k + 1]] for k in range(nb)])

    def hops(self):
        n = ctypes.c_int64()
        L = lib()
        check(L.smgpu_hops_export(self._h, ctypes.byref(n), None, None, None, None, None, None, None, None))
        m = n.value
        arrs = [np.zeros(m, np.int32) for _ in range(7)]
        ph = np.zeros(m, np.int8)
        check(L.smgpu_hops_export(self._h, ctypes.byref(n), *[ptr(a) for a in arrs], ptr(ph)))
        keys = ["species", "bi", "bf", "i", "f", "a", "c"]
        out = dict(zip(keys, arrs))
        out["phase"] = ph
        return out

    def pair_matrix(self, species):
        n = ctypes.c_int64()
        L = lib()
        check(L.smgpu_pair_export(self._h, species, ctypes.byref(n), None, None, None, None))
        m = n.value
        b = np.zeros(m, np.int32); f = np.zeros(m, np.int32); i = np.zeros(m, np.int32)
        v = np.zeros(m, np.float64)
        check(L.smgpu_pair_export(self._h, species, ctypes.byref(n), ptr(b), ptr(f), ptr(i), ptr(v)))
        return b, f, i, v

    def j2_hamiltonian(self, device=0):
        """Device tables of J^2 on this basis (prep_J / operate_J!,
        shellmodel_main.jl:149-157, 226-230): its operate_H applies J^2."""
        return Hamiltonian(self, device, _j2=True)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().smgpu_problem_destroy(h)
            self._h = None


class Hamiltonian:
    """Device-resident H for a Problem on one GPU."""

    def __init__(self, problem: Problem, device=0, _j2=False):
        h = ctypes.c_void_p()
        create = lib().smgpu_ham_create_j2 if _j2 else lib().smgpu_ham_create
        check(create(problem._h, int(device), ctypes.byref(h)))
        self._h = h
        self.problem = problem
        self.mdim = problem.mdim

    def update_interaction(self, problem: Problem | None = None):
        check(lib().smgpu_ham_update_interaction(self._h, (problem or self.problem)._h))

    def operate_H(self, v, precision="fp64"):
        """w = H v (host float64 in/out, computed on the device in `precision`)."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.mdim,)
        w = np.zeros(self.mdim, np.float64)
        check(lib().smgpu_apply_H_host(self._h, _prec(precision), ptr(v), ptr(w)))
        return w

    def operate_H_multi(self, V, precision="fp64"):
        """W[b] = H V[b] for the rows of V (nvec x mdim), several vectors per launch
        (bl_operate_H!)."""
        V = np.ascontiguousarray(np.atleast_2d(V), dtype=np.float64)
        assert V.shape[1] == self.mdim
        W = np.zeros_like(V)
        check(lib().smgpu_apply_H_host_multi(self._h, _prec(precision), V.shape[0], ptr(V), ptr(W)))
        return W

    def time_operate_H(self, precision="mixed", reps=20, nvec=1):
        """Milliseconds per H·v launch (nvec > 1: per call on nvec vectors)."""
        ms = ctypes.c_double()
        if nvec == 1:
            check(lib().smgpu_time_apply_H(self._h, _prec(precision), int(reps), ctypes.byref(ms)))
        else:
            check(lib().smgpu_time_apply_H_multi(self._h, _prec(precision), int(nvec), int(reps), ctypes.byref(ms)))
        return ms.value

    def table_bytes(self, precision="mixed"):
        b = np.zeros(64, np.int64)
        n = ctypes.c_int32()
        buf = ctypes.create_string_buffer(4096)
        check(lib().smgpu_table_bytes(self._h, _prec(precision), 64, ptr(b), ctypes.byref(n), buf, 4096))
        names = buf.value.decode().strip().split("\n")
        return dict(zip(names, [int(x) for x in b[:n.value]]))

    def lds_bytes(self, precision="mixed"):
        """LDS bytes of one one-vector H·v launch, by part (smgpu_lds_bytes)."""
        t = np.zeros(5, np.int64)
        check(lib().smgpu_lds_bytes(self._h, _prec(precision), ptr(t)))
        return dict(own_rows=int(t[0]), diag=int(t[1]), nn=int(t[2]), pn_stage=int(t[3]), pn_walk=int(t[4]))

    def work_terms(self):
        t = np.zeros(5, np.int64)
        check(lib().smgpu_work_terms(self._h, ptr(t)))
        return dict(pn_padded=int(t[0]), pn=int(t[1]), pp=int(t[2]), nn=int(t[3]), diag=int(t[4]))

    def lanczos_steps(self, nsteps, v0, precision="fp64"):
        """(alpha, beta) of `nsteps` steps of the device Lanczos recurrence from v0."""
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        a = np.zeros(nsteps)
        b = np.zeros(nsteps)
        check(lib().smgpu_lanczos_steps(self._h, _prec(precision), int(nsteps), ptr(v0), ptr(a), ptr(b)))
        return a, b

    def trl(self, num_ev, lm=300, ls=20, tol=1e-8, num_history=3, precision="fp64", refine=True, seed=1234,
            want_vectors=False, max_iter=0, v0=None, j2=None, twoJ=None, reorth="default", steps="auto"):
        """Thick-restart Lanczos; returns (evals, evecs or None, stats dict).
        With j2 (a J^2 Hamiltonian on this basis) and twoJ: J-projected TRL
        (doubleLanczos), the lowest states with J = twoJ/2.  reorth: "default"
        (full for fp64, partial for mixed), "full" or "partial".  steps: "graph"
        replays each Lanczos step as a captured hipGraph, "launch" launches its
        kernels, "auto" leaves the choice to the library (results bit-identical)."""
        p = SolveParams()
        lib().smgpu_default_params(ctypes.byref(p))
        p.num_ev = int(num_ev); p.lm = int(lm); p.ls = int(ls); p.tol = float(tol)
        p.num_history = int(num_history); p.precision = _prec(precision); p.refine = int(bool(refine))
        p.seed = int(seed); p.max_iter = int(max_iter)
        p.reorth = {"default": 0, "full": 1, "partial": 2}[reorth]
        p.steps = {"auto": 0, "launch": 1, "graph": 2}[steps]
        nev = min(int(num_ev), self.mdim)
        ev = np.zeros(nev, np.float64)
        vec = np.zeros((nev, self.mdim), np.float64) if want_vectors else None
        st = SolveStats()
        v0p = None
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
            assert v0.shape == (self.mdim,)
            v0p = ptr(v0)
        if j2 is not None:
            check(lib().smgpu_trl_solve_j(self._h, j2._h, int(twoJ), ctypes.byref(p), v0p, ptr(ev),
                                          ptr(vec) if vec is not None else None, ctypes.byref(st)))
        else:
            check(lib().smgpu_trl_solve_v0(self._h, ctypes.byref(p), v0p, ptr(ev),
                                           ptr(vec) if vec is not None else None, ctypes.byref(st)))
        stats = dict(iterations=st.iterations, n_hv=st.n_hv, n_hv64=st.n_hv64, n_restart=st.n_restart,
                     wall_ms=st.wall_ms, ritz=np.array(st.ritz[:nev]), n_reorth2=st.n_reorth2,
                     n_jproj=st.n_jproj, n_reorth=st.n_reorth, n_fallback=st.n_fallback)
        return ev, vec, stats

    def trl_batch(self, num_ev, seeds, lm=300, ls=20, tol=1e-8, num_history=3, precision="fp64", refine=True,
                  want_vectors=False, max_iter=0, reorth="default"):
        """len(seeds) TRL solves in lockstep with shared H·v launches
        (smgpu_trl_solve_batch); returns [(evals, evecs or None, stats)] per seed."""
        return _solve_batch([self] * len(seeds), False, num_ev, seeds, lm, ls, tol, num_history, precision, refine,
                            want_vectors, max_iter, reorth)

    def variant(self, problem: Problem):
        """A Hamiltonian for `problem`'s interaction on this handle's basis that shares
        this handle's device index tables (smgpu_ham_create_variant): only the
        coefficients are uploaded, as for prepEC's samples on one basis
        (eigenvector_continuation.jl:140-175)."""
        h = ctypes.c_void_p()
        check(lib().smgpu_ham_create_variant(self._h, problem._h, ctypes.byref(h)))
        obj = Hamiltonian.__new__(Hamiltonian)
        obj._h = h
        obj.problem = problem
        obj.mdim = self.mdim
        return obj

    def trbl(self, num_ev, q, lm=300, ls=20, tol=1e-8, num_history=3, precision="fp64", refine=True, seed=1234,
             want_vectors=False, max_iter=0, v0=None, j2=None, twoJ=None):
        """Thick-restart block Lanczos with block size q (TRBL); returns (evals,
        evecs or None, stats dict).  v0: q start vectors [q, mdim].  With j2 and
        twoJ: J-projected block Lanczos, the lowest states with J = twoJ/2."""
        p = SolveParams()
        lib().smgpu_default_params(ctypes.byref(p))
        p.num_ev = int(num_ev); p.lm = int(lm); p.ls = int(ls); p.tol = float(tol)
        p.num_history = int(num_history); p.precision = _prec(precision); p.refine = int(bool(refine))
        p.seed = int(seed); p.max_iter = int(max_iter)
        nev = min(int(num_ev), self.mdim)
        ev = np.zeros(nev, np.float64)
        vec = np.zeros((nev, self.mdim), np.float64) if want_vectors else None
        st = SolveStats()
        v0p = None
        if v0 is not None:
            v0 = np.ascontiguousarray(np.atleast_2d(v0), dtype=np.float64)
            assert v0.shape == (int(q), self.mdim)
            v0p = ptr(v0)
        check(lib().smgpu_trbl_solve(self._h, j2._h if j2 is not None else None, int(twoJ or 0), int(q),
                                     ctypes.byref(p), v0p, ptr(ev), ptr(vec) if vec is not None else None,
                                     ctypes.byref(st)))
        stats = dict(iterations=st.iterations, n_hv=st.n_hv, n_hv64=st.n_hv64, n_restart=st.n_restart,
                     wall_ms=st.wall_ms, ritz=np.array(st.ritz[:nev]), n_reorth2=st.n_reorth2,
                     n_jproj=st.n_jproj, n_reorth=st.n_reorth, n_fallback=st.n_fallback)
        return ev, vec, stats

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().smgpu_ham_destroy(h)
            self._h = None


def _solve_batch(hams, variants, num_ev, seeds, lm, ls, tol, num_history, precision, refine, want_vectors,
                 max_iter, reorth):
    nb = len(seeds)
    mdim = hams[0].mdim
    params = (SolveParams * nb)()
    nev = min(int(num_ev), mdim)
    evs = [np.zeros(nev, np.float64) for _ in range(nb)]
    vecs = [np.zeros((nev, mdim), np.float64) for _ in range(nb)] if want_vectors else None
    for b, sd in enumerate(seeds):
        p = params[b]
        lib().smgpu_default_params(ctypes.byref(p))
        p.num_ev = int(num_ev); p.lm = int(lm); p.ls = int(ls); p.tol = float(tol)
        p.num_history = int(num_history); p.precision = _prec(precision); p.refine = int(bool(refine))
        p.seed = int(sd); p.max_iter = int(max_iter)
        p.reorth = {"default": 0, "full": 1, "partial": 2}[reorth]
    ev_ptrs = (ctypes.c_void_p * nb)(*[ptr(e) for e in evs])
    vec_ptrs = (ctypes.c_void_p * nb)(*[ptr(v) for v in vecs]) if vecs is not None else None
    stats = (SolveStats * nb)()
    status = (ctypes.c_int32 * nb)()
    if variants:
        hs = (ctypes.c_void_p * nb)(*[h._h.value for h in hams])
        check(lib().smgpu_trl_solve_variants(hs, nb, params, None, ev_ptrs, vec_ptrs, stats, status))
    else:
        check(lib().smgpu_trl_solve_batch(hams[0]._h, nb, params, None, ev_ptrs, vec_ptrs, stats, status))
    out = []
    for b in range(nb):
        st = stats[b]
        sd = dict(iterations=st.iterations, n_hv=st.n_hv, n_hv64=st.n_hv64, n_restart=st.n_restart,
                  wall_ms=st.wall_ms, ritz=np.array(st.ritz[:nev]), n_reorth2=st.n_reorth2,
                  n_jproj=st.n_jproj, n_reorth=st.n_reorth, n_fallback=st.n_fallback)
        out.append((evs[b], vecs[b] if vecs is not None else None, sd))
    return out


def _handles(hams):
    assert len(hams) >= 1 and all(h.mdim == hams[0].mdim for h in hams)
    return (ctypes.c_void_p * len(hams))(*[h._h.value for h in hams])


def operate_H_variants(hams, V, precision="fp64"):
    """W[b] = H_b V[b] for Hamiltonians on one index (Hamiltonian.variant; a handle may
    repeat), the vectors of different interactions sharing index loads in one launch
    (smgpu_apply_H_host_variants); each row bit-identical to hams[b].operate_H(V[b])."""
    V = np.ascontiguousarray(np.atleast_2d(V), dtype=np.float64)
    assert V.shape == (len(hams), hams[0].mdim)
    W = np.zeros_like(V)
    check(lib().smgpu_apply_H_host_variants(_handles(hams), _prec(precision), len(hams), ptr(V), ptr(W)))
    return W


def time_operate_H_variants(hams, precision="mixed", reps=20):
    """Milliseconds per smgpu_apply_H_variants call on len(hams) resident vectors."""
    ms = ctypes.c_double()
    check(lib().smgpu_time_apply_H_variants(_handles(hams), _prec(precision), len(hams), int(reps),
                                            ctypes.byref(ms)))
    return ms.value


def trl_variants(hams, num_ev, seeds, lm=300, ls=20, tol=1e-8, num_history=3, precision="fp64", refine=True,
                 want_vectors=False, max_iter=0, reorth="default"):
    """Lockstep TRL solves of interaction variants on one basis: member b solves
    hams[b] from seed seeds[b] (smgpu_trl_solve_variants; prepEC's samples,
    eigenvector_continuation.jl:140-175, solved together).  Returns
    [(evals, evecs or None, stats)] per member, each bit-identical to hams[b].trl."""
    assert len(hams) == len(seeds)
    _handles(hams)
    return _solve_batch(list(hams), True, num_ev, seeds, lm, ls, tol, num_history, precision, refine,
                        want_vectors, max_iter, reorth)


def j_from_jj1(jj, tol=1e-6):
    """J_from_JJ1 (shellmodel_main.jl:1105-1113): the J in 0, 1/2, ..., 50 with
    |J(J+1) - jj| < tol, else -1.0."""
    for J2 in range(0, 101):
        hJ = 0.5 * J2
        if abs(hJ * (hJ + 1.0) - jj) < tol:
            return hJ
    return -1.0


def expect_j2(problem: Problem, vecs, device=0):
    """<v|J^2|v> of each row of `vecs` (main_sm's Js, shellmodel_main.jl:226-230)."""
    hj = problem.j2_hamiltonian(device)
    V = np.atleast_2d(vecs)
    W = hj.operate_H_multi(V, "fp64")
    return np.einsum("ij,ij->i", V, W)


def mdim_print(target_nuc, info, tJ=-1):
    """The dimension line main_sm prints (mdim_print, shellmodel_main.jl:1316-1321)."""
    s0 = "%6s %14s %10s %10s" % (target_nuc, f"Z,N=({info.Z},{info.N})", f"c({info.cp},{info.cn})",
                                 f"v({info.vp},{info.vn})")
    if tJ != -1:
        s0 += " 2J = %3i" % tJ
    s1 = "%s %12i %s %5.2f %s " % ("  mdim:", info.mdim, "(", np.log10(max(info.mdim, 1)), ")")
    print(s0 + s1, end="")          # the J / En lines follow on the same line, as in main_sm


class Moments:
    """The M1 and E2 one-body operators of main_sm's calc_moment on a problem's basis,
    on one GPU (prep_obtr_diag, transit.jl:82-245): `apply` is op_obtr_diag
    (:247-283), `eval` is eval_moment (:702-765)."""

    def __init__(self, problem: Problem, device=0, gfactors=(1.0, 0.0, 5.586, -3.826), effcharge=(1.5, 0.5),
                 bpar=0.0):
        g = np.ascontiguousarray(gfactors, dtype=np.float64)
        e = np.ascontiguousarray(effcharge, dtype=np.float64)
        if g.shape != (4,) or e.shape != (2,):
            raise ValueError("gfactors needs 4 values and effcharge 2")
        h = ctypes.c_void_p()
        check(lib().smgpu_moments_create(problem._h, int(device), ptr(g), ptr(e), float(bpar), ctypes.byref(h)))
        self._h = h
        self.mdim = problem.mdim
        b = ctypes.c_double()
        check(lib().smgpu_moments_bpar(h, ctypes.byref(b)))
        self.bpar = b.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().smgpu_moments_destroy(self._h)
            self._h = None

    def apply(self, V, which):
        """O V for the rows of V; which = "M1" or "E2"."""
        V = np.ascontiguousarray(np.atleast_2d(V), dtype=np.float64)
        W = np.zeros_like(V)
        check(lib().smgpu_moments_apply_host(self._h, {"M1": 0, "E2": 1}[which], V.shape[0], ptr(V), ptr(W)))
        return W

    def eval(self, vecs, totJs):
        """(mu, Q) per state; NaN where eval_moment prints nothing."""
        V = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
        J = np.ascontiguousarray(totJs, dtype=np.float64)
        mu = np.zeros(V.shape[0])
        Q = np.zeros(V.shape[0])
        check(lib().smgpu_eval_moment(self._h, V.shape[0], ptr(V), ptr(J), ptr(mu), ptr(Q)))
        return mu, Q


class Transitions:
    """The M1 and E2 operators of transit_main between two problems on one interaction
    and nucleus (prep_obtr_nondiag, transit.jl:286-445): the components mu = (2M_l -
    2M_r) / 2 from the right basis into the left one, on one GPU.  `apply` is
    op_obtr_nondiag (:447-499), `transit` the reduced matrix elements of :598-697."""

    def __init__(self, left: Problem, right: Problem, device=0, gfactors=(1.0, 0.0, 5.586, -3.826),
                 effcharge=(1.5, 0.5), bpar=0.0):
        g = np.ascontiguousarray(gfactors, dtype=np.float64)
        e = np.ascontiguousarray(effcharge, dtype=np.float64)
        if g.shape != (4,) or e.shape != (2,):
            raise ValueError("gfactors needs 4 values and effcharge 2")
        h = ctypes.c_void_p()
        check(lib().smgpu_transitions_create(left._h, right._h, int(device), ptr(g), ptr(e), float(bpar),
                                             ctypes.byref(h)))
        self._h = h
        mu2 = ctypes.c_int32(); b = ctypes.c_double(); ml = ctypes.c_int64(); mr = ctypes.c_int64()
        check(lib().smgpu_transitions_info(h, ctypes.byref(mu2), ctypes.byref(b), ctypes.byref(ml), ctypes.byref(mr)))
        self.mu2, self.bpar, self.mdim_l, self.mdim_r = mu2.value, b.value, ml.value, mr.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().smgpu_transitions_destroy(self._h)
            self._h = None

    def apply(self, V, which):
        """O V for the rows of V (right basis) into the left basis; which = "M1" or "E2"."""
        V = np.ascontiguousarray(np.atleast_2d(V), dtype=np.float64)
        if V.shape[1] != self.mdim_r:
            raise ValueError("vectors of the right basis expected")
        W = np.zeros((V.shape[0], self.mdim_l))
        check(lib().smgpu_transitions_apply_host(self._h, {"M1": 0, "E2": 1}[which], V.shape[0], ptr(V), ptr(W)))
        return W

    def transit(self, jl2, jr2, left_vecs, right_vecs):
        """(Mred M1, Mred E2), each [n_left, n_right]; NaN where not allowed."""
        L = np.ascontiguousarray(np.atleast_2d(left_vecs), dtype=np.float64)
        R = np.ascontiguousarray(np.atleast_2d(right_vecs), dtype=np.float64)
        if L.shape[1] != self.mdim_l or R.shape[1] != self.mdim_r:
            raise ValueError("vectors of the left and right bases expected")
        m1 = np.zeros((L.shape[0], R.shape[0]))
        e2 = np.zeros((L.shape[0], R.shape[0]))
        check(lib().smgpu_transit(self._h, int(jl2), int(jr2), L.shape[0], ptr(L), R.shape[0], ptr(R), ptr(m1),
                                  ptr(e2)))
        return m1, e2


def _w3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol of doubled arguments (Racah's formula, exact integers)."""
    from fractions import Fraction
    from math import factorial as F
    if m1 + m2 + m3 != 0 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2 or (j1 + j2 + j3) % 2 or (j1 + m1) % 2 or (j2 + m2) % 2 or (j3 + m3) % 2:
        return 0.0
    a, b, c = (j1 + j2 - j3) // 2, (j1 - j2 + j3) // 2, (-j1 + j2 + j3) // 2
    tri = Fraction(F(a) * F(b) * F(c), F((j1 + j2 + j3) // 2 + 1))
    pre = tri * F((j1 + m1) // 2) * F((j1 - m1) // 2) * F((j2 + m2) // 2) * F((j2 - m2) // 2) * \
        F((j3 + m3) // 2) * F((j3 - m3) // 2)
    s = Fraction(0)
    for k in range(0, a + 1):
        d = [k, a - k, (j1 - m1) // 2 - k, (j2 + m2) // 2 - k, (j3 - j2 + m1) // 2 + k, (j3 - j1 - m2) // 2 + k]
        if min(d) < 0:
            continue
        den = 1
        for x in d:
            den *= F(x)
        s += Fraction((-1) ** k, den)
    sign = -1 if ((j1 - j2 - m3) // 2) % 2 else 1
    return sign * float(s) * float(pre) ** 0.5


def transit_main(sntf, target_nuc, jl2, jr2, in_wfs, *, num_ev_l=100, num_ev_r=100, calc_EM=True,
                 gfactors=(1.0, 0.0, 5.586, -3.826), eff_charge=(1.5, 0.5), device=0, verbose=True):
    """transit_main (transit.jl:561-700): M1 and E2 transitions between the states of two
    .wav files, the left one at 2M = 2J = jl2 and the right one at jr2 (the files are
    swapped when jl2 < jr2, as the reference does).  For every right state r and left state
    l: Mred = <l||O||r> (reduced matrix element), B(l <= r) = Mred^2 / (2J_r + 1) and
    B(l => r) = Mred^2 / (2J_l + 1); with one file given twice, <gm> (mu_N) and <eQ> (e fm^2)
    of each state.  calc_EM=False gives only the latter.  Prints the reference's lines when
    verbose and returns them as dicts."""
    from .wav import read_wav
    if len(in_wfs) != 2:
        raise ValueError("input wfs must be specified: [left .wav, right .wav]")
    widx = [0, 1]
    if jl2 < jr2:
        jl2, jr2, widx = jr2, jl2, [1, 0]
    left = Problem(sntf, target_nuc, twoM=jl2)
    right = Problem(sntf, target_nuc, twoM=jr2)
    _, _, Vl, _ = read_wav(in_wfs[widx[0]], left.mdim)
    _, _, Vr, _ = read_wav(in_wfs[widx[1]], right.mdim)
    Vl, Vr = Vl[:num_ev_l], Vr[:num_ev_r]
    if (jl2 - jr2) // 2 > 2:
        raise ValueError("only M1 and E2 transitions are supported")
    T = Transitions(left, right, device, gfactors, eff_charge)
    m1, e2 = T.transit(jl2, jr2, Vl, Vr)
    same = in_wfs[0] == in_wfs[1]
    omu2 = jl2 - jr2
    out = []
    for mode, lam2, red, cmom, tag in (("M1", 2, m1, np.sqrt(4 * np.pi / 3), "<gm>"),
                                       ("E2", 4, e2, np.sqrt(16 * np.pi / 5), "<eQ>")):
        if np.all(np.isnan(red)):
            continue
        cg = _w3j(jr2, lam2, jl2, -jr2, omu2, jl2)
        sign = -1.0 if ((jl2 - jl2) // 2) % 2 else 1.0      # (-1)^(J_l - M_l), M_l = J_l
        for r in range(red.shape[1]):
            for l in range(red.shape[0]):
                if not calc_EM and l != r:
                    continue
                mr = float(red[l, r])
                rec = {"mode": mode, "l": l + 1, "r": r + 1, "Mred": mr, "B_l_from_r": mr ** 2 / (jr2 + 1),
                       "B_r_from_l": mr ** 2 / (jl2 + 1), "moment": None}
                if (same or not calc_EM) and l == r:
                    rec["moment"] = mr * cg * cmom * sign
                if verbose:
                    if calc_EM:
                        tx = (f"Mred.({mode}) ({l + 1},{r + 1}) {mr:9.3f}   B({mode}) ({l + 1}<={r + 1}) "
                              f"{rec['B_l_from_r']:9.3f}  B({mode}) ({l + 1}=>{r + 1}) {rec['B_r_from_l']:9.3f} ")
                        if rec["moment"] is not None:
                            tx += f" {tag}= {rec['moment']:9.3f} "
                    else:
                        tx = f" {tag}= {rec['moment']:9.3f} "
                    print(tx)
                out.append(rec)
    return out


def format_moments(mu, Q):
    """eval_moment's text (transit.jl:730-763): per state 'mu (n=k) ' / 'Q (n=k) '."""
    out = ""
    for n, (m, q) in enumerate(zip(mu, Q)):
        tx = ""
        if not np.isnan(m):
            tx += f"mu (n={n + 1}) " + f"{m:12.5f} " + "\t"
        if not np.isnan(q):
            tx += f"Q (n={n + 1}) " + f"{q:12.5f} "
        if tx:
            out += tx + "\n"
    return out


def _print_levels(ev, totJs, tJ):
    """main_sm's result lines (shellmodel_main.jl:256-266)."""
    if tJ is None:
        print(f"J {list(totJs)}")
        print("En. " + "".join("%10.4f " % x for x in ev))
        print("Ex. " + "".join("%10.4f " % (x - ev[0]) for x in ev))
    else:
        print(f"2J= {int(round(2 * totJs[0]))}  En." + "".join("%12.5f " % x for x in ev))


def main_sm(sntf, target_nuc, num_ev, target_J=(), *, lm=300, ls=20, tol=1e-8, num_history=3,
            precision="fp64", refine=True, device=0, seed=1234, return_stats=False, return_J=False,
            save_wav=False, wav_path=None, in_wf="", q=1, is_block=False, mdimmode=False, verbose=False,
            is_show=False, calc_moment=False, gfactors=(1.0, 0.0, 5.586, -3.826), effcharge=(1.5, 0.5),
            return_moments=False):
    """Lowest `num_ev` eigenvalues (MeV) of `target_nuc` with the interaction `sntf`.

    Mirrors main_sm (shellmodel_main.jl:89-281): with target_J = [] the plain
    TRL at 2M = A % 2 (:110); with target_J = [tJ] (2J, one value, :109-116) the
    J-projected TRL at 2M = tJ (doubleLanczos).  mdim == 0 returns an empty
    array (:127); num_ev is clamped to mdim (:128).  return_J adds the J of every state from <J^2> of
    its Ritz vector (totJs, :226-232); save_wav writes the Ritz vectors in the
    reference's .wav layout (writeRitzvecs, :1215-1229), by default to
    ./<nucleus>_<interaction>.wav (:241-246; ..._j<tJ>.wav with target_J).  in_wf
    starts the Lanczos from the first vector of a .wav file (read_appwav, :194-198).
    is_block=True runs the thick-restart block Lanczos with block size q instead
    (TRBL, :167-191; in_wf then supplies the first q vectors).  mdimmode=True
    prints the dimension line and returns None after the basis (:129-130);
    verbose=True prints main_sm's dimension and J / En / Ex lines; is_show=True
    prints the wall time of the table build, upload and solve (main_sm's
    TimerOutput summary, :279).  calc_moment=True computes the magnetic dipole and
    electric quadrupole moments of the states (eval_moment, :242-246, 267-269) with
    main_sm's gfactors / effcharge keywords and prints them as main_sm does
    (with verbose) or returns (mu, Q) with return_moments=True.
    """
    if len(target_J) > 1:
        raise ValueError("Multiple J is not supported (main_sm asserts length(target_J) <= 1)")
    tJ = int(target_J[0]) if len(target_J) else None
    t0 = time.perf_counter()
    prob = Problem(sntf, target_nuc, twoM=-1 if tJ is None else tJ)
    t1 = time.perf_counter()
    if verbose or mdimmode:
        mdim_print(target_nuc, prob.info, -1 if tJ is None else tJ)
    if mdimmode:
        print()
        return None
    if prob.mdim == 0:
        return (np.zeros(0), {}) if return_stats else np.zeros(0)
    ham = Hamiltonian(prob, device)
    t2 = time.perf_counter()
    nev = min(int(num_ev), prob.mdim)
    need_vecs = return_J or save_wav or verbose or calc_moment or return_moments
    v0 = None
    if in_wf:
        from .wav import read_appwav
        v0 = read_appwav(in_wf, prob.mdim, int(q) if is_block else 1)
    j2 = prob.j2_hamiltonian(device) if tJ is not None else None
    if is_block:
        ev, vecs, st = ham.trbl(nev, int(q), lm=lm, ls=ls, tol=tol, num_history=num_history, precision=precision,
                                refine=refine, seed=seed, want_vectors=need_vecs, v0=v0, j2=j2, twoJ=tJ)
    else:
        ev, vecs, st = ham.trl(nev, lm=lm, ls=ls, tol=tol, num_history=num_history, precision=precision,
                               refine=refine, seed=seed, want_vectors=need_vecs, v0=v0, j2=j2, twoJ=tJ)
    t3 = time.perf_counter()
    out = [ev]
    if need_vecs:
        totJs = np.array([j_from_jj1(x) for x in expect_j2(prob, vecs, device)])
        if verbose:
            _print_levels(ev, totJs, tJ)
        if save_wav:
            from .wav import write_ritzvecs
            if wav_path is None:
                csnt = os.path.basename(sntf).split(".")[0]
                wav_path = f"./{target_nuc}_{csnt}.wav" if tJ is None else f"./{target_nuc}_{csnt}_j{tJ}.wav"
            write_ritzvecs(wav_path, prob.info.twoM, ev, totJs, vecs)
        if return_J:
            out.append(totJs)
        if calc_moment or return_moments:
            mom = Moments(prob, device, gfactors, effcharge)
            mu, Qm = mom.eval(vecs, totJs)
            del mom
            tx_mom = format_moments(mu, Qm)
            if verbose and tx_mom:
                print(tx_mom)
            if return_moments:
                out.append((mu, Qm))
    if is_show:
        print(f"tables {1e3 * (t1 - t0):.1f} ms, upload {1e3 * (t2 - t1):.1f} ms, solve {1e3 * (t3 - t2):.1f} ms "
              f"({st['iterations']} iterations, {st['n_hv']} H.v, {st['n_restart']} restarts)")
    if return_stats:
        out.append(st)
    return out[0] if len(out) == 1 else tuple(out)
