"""ORACLE — test infrastructure only, never shipped or measured as the product.

CPU restatement of NuclearToolkit.jl's M-scheme shell-model path, used only by
tests/, `__graft_entry__.smoke()` and bench.py's `cpu_baseline` leg as the
checker.  Every function cites the reference file:line it restates (paths are
relative to the reference repository root).

Parts:
  * `read_snt`              src/ShellModel/shellmodel_main.jl:293-365 (+ util_func.jl:1,26-46)
  * `cg`                    exact Clebsch-Gordan (Racah formula, rational arithmetic),
                            replacing WignerSymbols.clebschgordan (call sites :622-623, :705, :711)
  * `construct_msps`        shellmodel_main.jl:373-384
  * `occ`                   shellmodel_main.jl:769-848 (+ all_perm! :401-408, Mcount! :410-418,
                            bitarr_to_int :732-739, count_jocc! :899-907)
  * `prep_hamil_t1`         shellmodel_main.jl:580-649 (operator list part)
  * `prep_hamil_pn`         shellmodel_main.jl:658-730
  * hop tables + operate_H! sm_oracle.c (shellmodel_main.jl:850-897, 1068-1099, 1115-1199;
                            lanczos_methods.jl:153-248)
  * `trl`                   lanczos_methods.jl:1-91, 250-307; shellmodel_main.jl:1005-1012
  * `main_sm`               shellmodel_main.jl:89-281 (eigenvalue observable only)
  * moments                 transit.jl:49-283, 702-765 (init_ho_by_mass, radius_power,
                            prep_obtr_diag, op_obtr_diag, eval_moment; exact 3j / 6j)
  * transitions             transit.jl:286-534, 561-700 (prep_obtr_nondiag, op_obtr_nondiag,
                            calc_bifs, transit_main)

Pinning (tests/test_oracle_*.py): the ckpot Be8 known answers
(test/ShellModel_test.jl:10-14), the Mg24 dimension (example/log_sample_script.txt:323),
closed-form dimensions, sympy's Clebsch-Gordan values, and an independent
Jordan-Wigner brute-force Hamiltonian (tests/brute_force.py).  The moments are pinned
by the single-particle closed forms (Schmidt values, Q = -(2j-1)/(2j+2) <r^2> e),
oscillator <r^2> values and 3j / 6j orthogonality (tests/test_moments.py).  The
transitions are pinned by single-nucleon matrix elements built independently from
Clebsch-Gordan expansions, angular-momentum ladder operators and Gaunt integrals (every
m-component, mu = 0, 1, 2), by their B values' closed form, and by <gm>, <eQ> of one state
given twice equal to eval_moment's (tests/test_transit.py).
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import re
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
            "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni",
            "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb",
            "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn"]


# --------------------------------------------------------------------------- L0 parse
def _rm_comment(lines):
    """util_func.jl:26-38: strip, drop lines starting with '!' or '#'."""
    out = []
    for line in lines:
        line = line.strip()
        if len(line) > 0 and (line.startswith("!") or line.startswith("#")):
            continue
        out.append(line)
    return out


def read_snt(path, nuc, ignore_massop=False):
    """readsmsnt (shellmodel_main.jl:293-365).

    labels[ch] are [a,b,c,d,J,line] with channel 0=pp, 1=nn, 2=pn; a reversed
    copy [c,d,a,b,J,line] is pushed *before* the original when {a,b} != {c,d}.
    """
    A = int(re.search(r"[0-9]+", nuc).group())
    if A <= 0:
        raise ValueError(f"invalid nucleus {nuc}")
    with open(path) as f:
        lines = _rm_comment(f.read().splitlines())
    lp, ln, cp, cn = [int(x) for x in lines[0].split()[:4]]
    p_sps = [None] * lp
    n_sps = [None] * ln
    for i in range(lp):
        ith, n, l, j, tz = [int(x) for x in lines[1 + i].split()[:5]]
        p_sps[ith - 1] = [n, l, j, tz]
    for i in range(ln):
        ith, n, l, j, tz = [int(x) for x in lines[1 + i + ln].split()[:5]]
        n_sps[ith - lp - 1] = [n, l, j, tz]
    nsp = int(lines[1 + ln + lp].split()[0])
    spes = [[0.0] * lp, [0.0] * ln]
    for i in range(nsp):
        t = lines[1 + ln + lp + 1 + i].split()
        j = int(t[0])
        if j <= lp:
            spes[0][j - 1] = float(t[2])
        else:
            spes[1][j - lp - 1] = float(t[2])
    t = lines[1 + ln + lp + 1 + nsp].split()
    ntbme, massop, Aref = int(t[0]), int(t[1]), float(t[2])
    p = float(t[3]) if len(t) > 3 else 0.0
    labels = [[], [], []]
    tbmes = [[], [], []]
    for ith in range(1, ntbme + 1):
        t = lines[1 + ln + lp + 1 + nsp + ith].split()
        i, j, k, l, J = [int(x) for x in t[:5]]
        if i <= lp and j <= lp:
            ch = 0
        elif i > lp and j > lp:
            ch = 1
        elif i <= lp and j > lp:
            ch = 2
        else:
            raise ValueError(f"unordered snt line {i} {j} {k} {l}")
        v = float(t[5])
        if massop == 1 and not ignore_massop:
            v *= (A / Aref) ** p
        if list(dict.fromkeys([i, j])) != list(dict.fromkeys([k, l])):
            labels[ch].append([k, l, i, j, J, ith])
            tbmes[ch].append(v)
        labels[ch].append([i, j, k, l, J, ith])
        tbmes[ch].append(v)
    return dict(lp=lp, ln=ln, cp=cp, cn=cn, massop=massop, Aref=Aref, p=p, p_sps=p_sps,
                n_sps=n_sps, SPEs=spes, labels=labels, TBMEs=tbmes, A=A)


def get_zna(nuc, cp, cn):
    """getZNA (shellmodel_main.jl:908-916) with the element table of util_func.jl:14-24."""
    A = int(re.search(r"[0-9]+", nuc).group())
    el = nuc.replace(str(A), "")
    Z = ELEMENTS.index(el) + 1
    N = A - Z
    return Z, N, Z - cp, N - cn


# --------------------------------------------------------------------------- CG
_CG_CACHE = {}


def _fact(n):
    return math.factorial(n)


def cg_fraction(j1, m1, j2, m2, J, M):
    """Exact CG^2 with sign: returns (sign, Fraction value of CG^2). Doubled arguments."""
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0, Fraction(0)
    if J > j1 + j2 or J < abs(j1 - j2):
        return 0, Fraction(0)
    if (j1 + m1) % 2 or (j2 + m2) % 2 or (J + M) % 2 or (j1 + j2 + J) % 2:
        return 0, Fraction(0)
    a = (j1 + j2 - J) // 2
    b = (j1 - j2 + J) // 2
    c = (-j1 + j2 + J) // 2
    d = (j1 + j2 + J) // 2 + 1
    pref = Fraction((J + 1) * _fact(a) * _fact(b) * _fact(c), _fact(d))
    pref *= (_fact((j1 + m1) // 2) * _fact((j1 - m1) // 2) * _fact((j2 + m2) // 2)
             * _fact((j2 - m2) // 2) * _fact((J + M) // 2) * _fact((J - M) // 2))
    s = Fraction(0)
    for k in range(0, 64):
        args = [k, a - k, (j1 - m1) // 2 - k, (j2 + m2) // 2 - k,
                (J - j2 + m1) // 2 + k, (J - j1 - m2) // 2 + k]
        if any(x < 0 for x in args[:4]):
            if args[1] < 0 or args[2] < 0 or args[3] < 0:
                break
            continue
        if args[4] < 0 or args[5] < 0:
            continue
        den = 1
        for x in args:
            den *= _fact(x)
        s += Fraction((-1) ** k, den)
    if s == 0:
        return 0, Fraction(0)
    return (1 if s > 0 else -1), pref * s * s


def cg(j1, m1, j2, m2, J, M):
    """<j1/2 m1/2; j2/2 m2/2 | J/2 M/2> as the correctly rounded float of the exact value."""
    key = (j1, m1, j2, m2, J, M)
    r = _CG_CACHE.get(key)
    if r is not None:
        return r
    sgn, sq = cg_fraction(j1, m1, j2, m2, J, M)
    if sgn == 0:
        val = 0.0
    else:
        K = 200
        num = sq.numerator << (2 * K)
        root = math.isqrt(num // sq.denominator)
        val = sgn * float(Fraction(root, 1 << K))
    _CG_CACHE[key] = val
    return val


# --------------------------------------------------------------------------- L1 basis
def construct_msps(p_sps, n_sps):
    """construct_msps (shellmodel_main.jl:373-384): [n,l,j,tz,mz,orbit(1-based)]."""
    msps_p, msps_n = [], []
    for idx, (n, l, j, tz) in enumerate(p_sps, start=1):
        for mz in range(-j, j + 1, 2):
            msps_p.append([n, l, j, tz, mz, idx])
    for idx, (n, l, j, tz) in enumerate(n_sps, start=1):
        for mz in range(-j, j + 1, 2):
            msps_n.append([n, l, j, tz, mz, idx])
    return msps_p, msps_n


def _sds(msps, nv):
    """all_perm!/bitarr_to_int/Mcount! (:401-418, :732-739): lexicographic combinations."""
    L = len(msps)
    out = []
    for comb in itertools.combinations(range(L), nv):
        bit = 0
        M = 0
        for k in comb:
            bit |= 1 << (L - 1 - k)
            M += msps[k][4]
        out.append((bit, M, comb))
    return out


def occ(p_sps, msps_p, vp, n_sps, msps_n, vn, Mtot):
    """occ (shellmodel_main.jl:769-848), no truncation scheme."""
    sp = _sds(msps_p, vp)
    sn = _sds(msps_n, vn)
    Mps = sorted(set(m for _, m, _ in sp))
    Mns = sorted(set(m for _, m, _ in sn), reverse=True)
    poss = [(i, j) for i in range(len(Mps)) for j in range(len(Mns)) if Mps[i] + Mns[j] == Mtot]
    pbits, nbits, jp, jn = [], [], [], []
    tdims = [0]
    for (i, j) in poss:
        Mp, Mn = Mps[i], Mns[j]
        pb = [(b, c) for b, m, c in sp if m == Mp]
        nb = [(b, c) for b, m, c in sn if m == Mn]
        pbits.append([b for b, _ in pb])
        nbits.append([b for b, _ in nb])
        jp.append([_jocc(p_sps, msps_p, c) for _, c in pb])
        jn.append([_jocc(n_sps, msps_n, c) for _, c in nb])
        tdims.append(tdims[-1] + len(pb) * len(nb))
    return pbits, nbits, jp, jn, [Mps[i] for i, _ in poss], [Mns[j] for _, j in poss], tdims


def _jocc(sps, msps, comb):
    """count_jocc! (:899-907)."""
    t = [0.0] * len(sps)
    for k in comb:
        t[msps[k][5] - 1] += 1.0
    return t


# --------------------------------------------------------------------------- L2 operator lists
def _possible_mz(orb, msps):
    """possible_mz (:552-565): (j, [mz], [0-based m-state index])."""
    n, l, j, tz = orb
    mzs, idx = [], []
    for mz in range(-j, j + 1, 2):
        mzs.append(mz)
        for k, ms in enumerate(msps):
            if ms[:5] == [n, l, j, tz, mz]:
                idx.append(k)
                break
    return j, mzs, idx


def prep_hamil_t1(p_sps, n_sps, msps_p, msps_n, labels, tbmes):
    """Operator-list part of prep_Hamil_T1 (:580-649).

    Returns for pp and nn: (ops[n,4] int64 bit quadruple a,b,c,d; V[n] float64),
    deduplicated on the exact quadruple in first-seen order.
    """
    res = []
    sps = [p_sps, n_sps]
    mst = [msps_p, msps_n]
    loffs = [0, len(p_sps)]
    for vr in range(2):
        L = len(mst[vr])
        seen = {}
        ops, Vs = [], []
        for (a, b, c, d, totJ, _), ME in zip(labels[vr], tbmes[vr]):
            J2 = 2 * totJ
            lo = loffs[vr]
            ja, ma_s, ma_i = _possible_mz(sps[vr][a - lo - 1], mst[vr])
            jb, mb_s, mb_i = _possible_mz(sps[vr][b - lo - 1], mst[vr])
            jc, mc_s, mc_i = _possible_mz(sps[vr][c - lo - 1], mst[vr])
            jd, md_s, md_i = _possible_mz(sps[vr][d - lo - 1], mst[vr])
            nfac = math.sqrt((1.0 + (1.0 if a == b else 0.0)) * (1.0 + (1.0 if c == d else 0.0)))
            for ic, mc in enumerate(mc_s):
                for id_, md in enumerate(md_s):
                    if c == d and mc >= md:
                        continue
                    if abs(mc + md) > J2:
                        continue
                    M = mc + md
                    bc = 1 << (L - 1 - mc_i[ic])
                    bd = 1 << (L - 1 - md_i[id_])
                    for ia, ma in enumerate(ma_s):
                        for ib, mb in enumerate(mb_s):
                            if a == b and ma >= mb:
                                continue
                            if ma + mb != M:
                                continue
                            ba = 1 << (L - 1 - ma_i[ia])
                            bb = 1 << (L - 1 - mb_i[ib])
                            cg1 = cg(ja, ma, jb, mb, J2, M)
                            cg2 = cg(jc, mc, jd, md, J2, M)
                            key = (ba, bb, bc, bd)
                            val = ME * nfac * cg1 * cg2
                            if key not in seen:
                                seen[key] = len(ops)
                                ops.append(key)
                                Vs.append(val)
                            else:
                                Vs[seen[key]] += val
        res.append((np.array(ops, dtype=np.int64).reshape(-1, 4), np.array(Vs, dtype=np.float64)))
    return res


def prep_hamil_pn(p_sps, n_sps, msps_p, msps_n, labels, tbmes):
    """prep_Hamil_pn (:658-730): per ΔMp in Mzs (descending) a list of ops and V."""
    lp, ln = len(msps_p), len(msps_n)
    loff = len(p_sps)
    Mzs = sorted(set(msps_p[i][4] - msps_p[j][4] for j in range(lp) for i in range(lp)),
                 reverse=True)
    pos = {m: k for k, m in enumerate(Mzs)}
    bV = [dict() for _ in Mzs]
    ops = [[] for _ in Mzs]
    Vs = [[] for _ in Mzs]
    for (a, b, c, d, totJ, _), ME in zip(labels, tbmes):
        J2 = 2 * totJ
        ja, ma_s, ma_i = _possible_mz(p_sps[a - 1], msps_p)
        jc, mc_s, mc_i = _possible_mz(p_sps[c - 1], msps_p)
        jb, mb_s, mb_i = _possible_mz(n_sps[b - loff - 1], msps_n)
        jd, md_s, md_i = _possible_mz(n_sps[d - loff - 1], msps_n)
        for ic, mc in enumerate(mc_s):
            bc = 1 << (lp - 1 - mc_i[ic])
            for ia, ma in enumerate(ma_s):
                ba = 1 << (lp - 1 - ma_i[ia])
                Mp = ma - mc
                k = pos[Mp]
                for id_, md in enumerate(md_s):
                    if abs(mc + md) > J2:
                        continue
                    bd = 1 << (ln - 1 - md_i[id_])
                    cg1 = cg(jc, mc, jd, md, J2, mc + md)
                    tfac = ME * cg1
                    for ib, mb in enumerate(mb_s):
                        if mb - md + Mp != 0:
                            continue
                        bb = 1 << (ln - 1 - mb_i[ib])
                        fac = tfac * cg(ja, ma, jb, mb, J2, ma + mb)
                        key = (ba, bb, bc, bd)
                        if key not in bV[k]:
                            bV[k][key] = len(ops[k])
                            ops[k].append(key)
                            Vs[k].append(fac)
                        else:
                            Vs[k][bV[k][key]] += fac
    return ([np.array(o, dtype=np.int64).reshape(-1, 4) for o in ops],
            [np.array(v, dtype=np.float64) for v in Vs], Mzs)


# --------------------------------------------------------------------------- C part
class _Lib:
    _lib = None

    @classmethod
    def get(cls):
        if cls._lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"oracle C library missing: build it with `make -C {HERE}`")
            lib = ctypes.CDLL(LIB_PATH)
            P = ctypes.c_void_p
            lib.orc_build.restype = P
            lib.orc_build.argtypes = [ctypes.c_int] + [P] * 6 + [ctypes.c_int, P, P, ctypes.c_int, P, P,
                                      ctypes.c_int, P, P, P, P, P, P, P, P, ctypes.c_int, ctypes.c_int]
            lib.orc_free.argtypes = [P]
            lib.orc_operate_H.argtypes = [P, P, P]
            lib.orc_counts.argtypes = [P, P]
            lib.orc_export_pp.argtypes = [P, ctypes.c_int, P, P, P, P]
            lib.orc_export_pn.argtypes = [P, ctypes.c_int, P, P, P, P, P, P]
            lib.orc_set_threads.argtypes = [ctypes.c_int]
            cls._lib = lib
        return cls._lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleHamiltonian:
    """Reference-layout hop tables built in C (prep_2bjumps, prep_pn) and operate_H!."""

    def __init__(self, snt_path, nuc, Mtot=None):
        sm = read_snt(snt_path, nuc)
        self.sm = sm
        A = sm["A"]
        if Mtot is None:
            Mtot = A % 2
        Z, N, vp, vn = get_zna(nuc, sm["cp"], sm["cn"])
        self.Z, self.N, self.vp, self.vn, self.Mtot = Z, N, vp, vn, Mtot
        msps_p, msps_n = construct_msps(sm["p_sps"], sm["n_sps"])
        self.msps_p, self.msps_n = msps_p, msps_n
        (self.pbits, self.nbits, self.jocc_p, self.jocc_n, self.Mps, self.Mns,
         self.tdims) = occ(sm["p_sps"], msps_p, vp, sm["n_sps"], msps_n, vn, Mtot)
        self.mdim = self.tdims[-1]
        (pp_ops, pp_V), (nn_ops, nn_V) = prep_hamil_t1(sm["p_sps"], sm["n_sps"], msps_p, msps_n,
                                                       sm["labels"], sm["TBMEs"])
        pn_ops, pn_V, Mzs = prep_hamil_pn(sm["p_sps"], sm["n_sps"], msps_p, msps_n,
                                          sm["labels"][2], sm["TBMEs"][2])
        self.pp_ops, self.pp_V, self.nn_ops, self.nn_V = pp_ops, pp_V, nn_ops, nn_V
        self.pn_ops, self.pn_V, self.delMs = pn_ops, pn_V, Mzs
        self._h = None
        if self.mdim > 0:
            self._build()

    def _build(self):
        lib = _Lib.get()
        nb = len(self.pbits)
        self._keep = []

        def keep(a):
            self._keep.append(a)
            return _ptr(a)
        poff = np.cumsum([0] + [len(x) for x in self.pbits]).astype(np.int64)
        noff = np.cumsum([0] + [len(x) for x in self.nbits]).astype(np.int64)
        pb = np.array([b for x in self.pbits for b in x], dtype=np.int64)
        nbv = np.array([b for x in self.nbits for b in x], dtype=np.int64)
        tdims = np.array(self.tdims, dtype=np.int64)
        Mps = np.array(self.Mps, dtype=np.int64)
        pp_ops = np.ascontiguousarray(self.pp_ops)
        nn_ops = np.ascontiguousarray(self.nn_ops)
        pp_V = np.ascontiguousarray(self.pp_V)
        nn_V = np.ascontiguousarray(self.nn_V)
        ndm = len(self.delMs)
        dms = np.array(self.delMs, dtype=np.int64)
        pnoff = np.cumsum([0] + [len(v) for v in self.pn_V]).astype(np.int64)
        pn_ops = np.ascontiguousarray(np.concatenate([o.reshape(-1, 4) for o in self.pn_ops]).astype(np.int64))
        pn_V = np.ascontiguousarray(np.concatenate(self.pn_V).astype(np.float64))
        sp = np.array(self.sm["SPEs"][0], dtype=np.float64)
        sn = np.array(self.sm["SPEs"][1], dtype=np.float64)
        jp = np.array([r for blk in self.jocc_p for r in blk], dtype=np.float64).reshape(-1)
        jn = np.array([r for blk in self.jocc_n for r in blk], dtype=np.float64).reshape(-1)
        self._h = lib.orc_build(nb, keep(pb), keep(poff), keep(nbv), keep(noff), keep(tdims), keep(Mps),
                                len(pp_V), keep(pp_ops), keep(pp_V), len(nn_V), keep(nn_ops), keep(nn_V),
                                ndm, keep(dms), keep(pnoff), keep(pn_ops), keep(pn_V),
                                keep(sp), keep(sn), keep(jp), keep(jn),
                                len(self.sm["p_sps"]), len(self.sm["n_sps"]))
        if not self._h:
            raise RuntimeError("oracle table build failed")

    def __del__(self):
        if getattr(self, "_h", None):
            _Lib.get().orc_free(self._h)
            self._h = None

    def operate_H(self, wf, twf=None):
        """twf += H wf (operate_H!, lanczos_methods.jl:153-248); returns twf."""
        wf = np.ascontiguousarray(wf, dtype=np.float64)
        if twf is None:
            twf = np.zeros(self.mdim, dtype=np.float64)
        _Lib.get().orc_operate_H(self._h, _ptr(wf), _ptr(twf))
        return twf

    def counts(self):
        """[pp entries, nn entries, pn pairs*ops, p-hop entries, n-hop entries, W]"""
        out = np.zeros(8, dtype=np.int64)
        _Lib.get().orc_counts(self._h, _ptr(out))
        return out

    def export_pp(self, species):
        """All T1info entries: (block, i, f, coef) 0-based, reference order."""
        c = self.counts()
        n = int(c[species])
        b = np.zeros(n, np.int64); i = np.zeros(n, np.int64); f = np.zeros(n, np.int64)
        v = np.zeros(n, np.float64)
        _Lib.get().orc_export_pp(self._h, species, _ptr(b), _ptr(i), _ptr(f), _ptr(v))
        return b, i, f, v

    def export_pn(self, species):
        """All pn one-body hop entries of the species: (bi, bf, op_global, i, f, phase_bit)."""
        c = self.counts()
        n = int(c[3 + species])
        arrs = [np.zeros(n, np.int64) for _ in range(5)]
        ph = np.zeros(n, np.int8)
        _Lib.get().orc_export_pn(self._h, species, *[_ptr(a) for a in arrs], _ptr(ph))
        return arrs[0], arrs[1], arrs[2], arrs[3], arrs[4], ph

    def dense(self):
        """Dense H by applying operate_H! to unit vectors (the calc_Hflat pattern,
        variational_ansatz.jl:481-585).  Small spaces only."""
        H = np.zeros((self.mdim, self.mdim))
        e = np.zeros(self.mdim)
        for k in range(self.mdim):
            e[:] = 0.0
            e[k] = 1.0
            H[:, k] = self.operate_H(e)
        return H


def set_threads(n):
    _Lib.get().orc_set_threads(int(n))


# --------------------------------------------------------------------------- L4 TRL
def _eigvals_lowest(T, n):
    return np.linalg.eigvalsh(T)[:n]


def trl(ham: OracleHamiltonian, num_ev, lm=300, ls=20, tol=1e-8, num_history=3, v0=None,
        seed=1234, max_restart=10000):
    """TRL (lanczos_methods.jl:1-91) with diagonalize_T! (:250-258), ReORTH
    (shellmodel_main.jl:1005-1012, MGS newest->oldest) and ThickRestart (:260-307).

    The start vector is a seeded normal vector (Julia's randn stream is not
    reproducible here; converged eigenvalues do not depend on it).
    Returns (en[1] lowest num_ev Ritz values, elit, number of H·v, number of restarts, vks, Tmat).
    """
    mdim = ham.mdim
    num_ev = min(num_ev, mdim)
    vks = np.zeros((lm, mdim))
    if v0 is None:
        rng = np.random.default_rng(seed)
        v0 = rng.standard_normal(mdim)
    vks[0] = v0 / math.sqrt(np.dot(v0, v0))
    uks = np.zeros((ls, mdim))
    T = np.zeros((lm, lm))
    en = [np.full(num_ev, 1.0e4) for _ in range(num_history)]
    k = 0
    conv = False
    elit = 1
    nhv = 0
    nres = 0
    while not conv:
        for it in range(k, lm - 1):
            vk = vks[it]
            vkp1 = vks[it + 1]
            vkp1[:] = 0.0
            ham.operate_H(vk, vkp1)
            nhv += 1
            T[it, it] = np.dot(vk, vkp1)
            # diagonalize_T!
            for h in range(num_history - 1, 0, -1):
                en[h][:] = en[h - 1]
            n = min(num_ev, it + 1)
            en[0][:n] = _eigvals_lowest(T[:it + 1, :it + 1], n)
            if np.all(en[1] - en[0] < tol):
                conv = True
            if it + 1 == mdim:
                conv = True
            if conv:
                elit = it + 1
                break
            for l in range(it, -1, -1):          # ReORTH newest -> oldest
                vkp1 -= np.dot(vks[l], vkp1) * vks[l]
            beta = math.sqrt(np.dot(vkp1, vkp1))
            vkp1 *= 1.0 / beta
            T[it + 1, it] = beta
            T[it, it + 1] = beta
        if not conv:
            nres += 1
            if nres > max_restart:
                raise RuntimeError("TRL did not converge")
            vals, vecs = np.linalg.eigh(T[:lm - 1, :lm - 1])
            beta = T[lm - 2, lm - 1]
            T[:, :] = 0.0
            for kk in range(ls):
                T[kk, kk] = vals[kk]
            for kk in range(ls):
                t = beta * vecs[lm - 2, kk]
                T[ls, kk] = t
                T[kk, ls] = t
                uks[kk] = vecs[:lm - 1, kk] @ vks[:lm - 1]
            for kk in range(ls):
                if kk > 0:
                    for l in range(kk - 1, -1, -1):
                        uks[kk] -= np.dot(uks[l], uks[kk]) * uks[l]
                vks[kk] = uks[kk] / math.sqrt(np.dot(uks[kk], uks[kk]))
            vks[ls] = vks[lm - 1]
            for l in range(ls - 1, -1, -1):
                vks[ls] -= np.dot(vks[l], vks[ls]) * vks[l]
            vks[ls + 1:] = 0.0
            k = ls
    return en[0].copy(), elit, nhv, nres, vks, T


def ritz_vectors(vks, T, elit, num_ev):
    """Ritz assembly of main_sm (:208-224)."""
    vals, vecs = np.linalg.eigh(T[:elit, :elit])
    R = vecs[:, :num_ev].T @ vks[:elit]
    R /= np.sqrt(np.sum(R * R, axis=1))[:, None]
    return vals[:num_ev], R


def main_sm(snt_path, nuc, num_ev, lm=300, ls=20, tol=1e-8):
    """Eigenvalue observable of main_sm (:89-281) with target_J=[] (plain TRL)."""
    ham = OracleHamiltonian(snt_path, nuc)
    if ham.mdim == 0:
        return np.zeros(0)
    en, elit, nhv, nres, vks, T = trl(ham, min(num_ev, ham.mdim), lm=lm, ls=ls, tol=tol)
    return en


# --------------------------------------------------------------------------- moments
# transit.jl: init_ho_by_mass :49-60, radius_power :62-80, prep_obtr_diag :82-245,
# op_obtr_diag :247-283, eval_moment :702-765; main_sm's mass formula :104-107.
def init_ho_by_mass(A, hwprm=1):
    """(hw, bpar) of the oscillator for mass number A (transit.jl:49-60)."""
    amass = (938.27231 + 939.56563) / 2
    hc = 197.32696
    if hwprm == 1:
        hw = 41.0 * A ** (-1.0 / 3.0)
    else:
        hw = 45.0 * A ** (-1.0 / 3.0) - 25.0 * A ** (-2.0 / 3.0)
    return hw, hc / math.sqrt(hw * amass)


def bpar_main_sm(A):
    """main_sm's choice (shellmodel_main.jl:104-107): formula 2 for 16 <= A <= 40."""
    return init_ho_by_mass(A, 2 if 16 <= A <= 40 else 1)[1]


def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def _binom(n, k):
    """Julia's binomial, including negative n."""
    if k < 0:
        return 0
    if n < 0:
        return (-1) ** k * _binom(k - n - 1, k)
    return math.comb(n, k)


def radius_power(k, n1, l1, n2, l2, bpar):
    """<n1 l1|r^k|n2 l2> of oscillator orbits (transit.jl:62-80), exact integers then float."""
    ll = l1 + l2 + k
    ll1 = int((l2 - l1 + k) / 2)      # Julia div: truncation toward zero
    ll2 = int((l1 - l2 + k) / 2)
    s = Fraction(0)
    for i in range(max(0, n1 - ll1, n2 - ll2), min(n1, n2) + 1):
        s += _binom(ll1, n1 - i) * _binom(ll2, n2 - i) * Fraction(_dfact(ll + 2 * i + 1), _dfact(2 * i))
    s = float(s)
    s *= math.sqrt(_dfact(2 * n1) / _dfact(2 * n1 + 2 * l1 + 1))
    s *= math.sqrt(_dfact(2 * n2) / _dfact(2 * n2 + 2 * l2 + 1))
    return s * (1.0 / math.sqrt(2 ** k)) * (-1) ** (n1 - n2) * bpar ** k


def wigner3j(j1, j2, j3, m1, m2, m3):
    """(j1 j2 j3; m1 m2 m3), doubled arguments: (-1)^(j1-j2-m3) / sqrt(2 j3 + 1) CG."""
    if m1 + m2 + m3 != 0 or (j1 - j2 - m3) % 2:
        return 0.0
    return (-1) ** ((j1 - j2 - m3) // 2) / math.sqrt(j3 + 1) * cg(j1, m1, j2, m2, j3, -m3)


def _delta2(a, b, c):
    """Delta(abc)^2 of doubled arguments (None if the triangle fails)."""
    if (a + b + c) % 2 or a + b < c or a + c < b or b + c < a:
        return None
    return Fraction(_fact((a + b - c) // 2) * _fact((a - b + c) // 2) * _fact((-a + b + c) // 2),
                    _fact((a + b + c) // 2 + 1))


def wigner6j(j1, j2, j3, j4, j5, j6):
    """{j1 j2 j3; j4 j5 j6}, doubled arguments: Racah's formula in rationals."""
    ds = [_delta2(j1, j2, j3), _delta2(j1, j5, j6), _delta2(j4, j2, j6), _delta2(j4, j5, j3)]
    if any(d is None for d in ds):
        return 0.0
    a = [(j1 + j2 + j3) // 2, (j1 + j5 + j6) // 2, (j4 + j2 + j6) // 2, (j4 + j5 + j3) // 2]
    b = [(j1 + j2 + j4 + j5) // 2, (j2 + j3 + j5 + j6) // 2, (j3 + j1 + j6 + j4) // 2]
    s = Fraction(0)
    for t in range(max(a), min(b) + 1):
        den = 1
        for x in a:
            den *= _fact(t - x)
        for x in b:
            den *= _fact(x - t)
        s += Fraction((-1) ** t * _fact(t + 1), den)
    sq = ds[0] * ds[1] * ds[2] * ds[3] * s * s
    if sq == 0:
        return 0.0
    K = 200
    root = math.isqrt((sq.numerator << (2 * K)) // sq.denominator)
    return (1 if s > 0 else -1) * float(Fraction(root, 1 << K))


def _phase_1(phi, cre, ani):
    """TF_connectable_1 + calc_phase_1! (shellmodel_main.jl:863-868, 884-894): (ok, sign, new)."""
    if phi & ani == 0:
        return False, 0, 0
    mid = phi & ~ani
    if mid & cre:
        return False, 0, 0
    e = 1 + bin(phi & ~(ani - 1)).count("1")
    e += bin(mid & ~(cre - 1)).count("1")
    return True, (-1) ** e, mid | cre


def prep_obtr_diag(p_sps, n_sps, msps_p, msps_n, pbits, nbits, bpar, gfactors, eff_charge):
    """prep_obtr_diag (transit.jl:82-245): M1 and E2 one-body operator lists
    {"M1": (proton, neutron), "E2": (proton, neutron)}, entries (block, Npi, Npf,
    phase, fac) with 0-based SD indices; phase means the term enters negated."""
    c_mu = math.sqrt(3.0 / (4.0 * math.pi))
    c_Q = math.sqrt(5.0 / (16.0 * math.pi))
    out = {"M1": ([], []), "E2": ([], [])}
    for sp, (sps, msps, bits) in enumerate(((p_sps, msps_p, pbits), (n_sps, msps_n, nbits))):
        gl, gs = (gfactors[0], gfactors[2]) if sp == 0 else (gfactors[1], gfactors[3])
        ec = eff_charge[sp]
        L = len(msps)
        for i in range(len(sps)):
            ni, li, ji, _ = sps[i]
            for f in range(i, len(sps)):
                nf, lf, jf, _ = sps[f]
                ex_r2 = radius_power(2, ni, li, nf, lf, bpar)
                l6j = wigner6j(2 * lf, 2 * li, 2, ji, jf, 1)
                s6j = wigner6j(1, 1, 2, ji, jf, 2 * lf)
                Q3j = wigner3j(ji, 4, jf, 1, 0, -1)
                _, mc_s, mc_idx = _possible_mz(sps[i], msps)
                _, ma_s, ma_idx = _possible_mz(sps[f], msps)
                pfac = math.sqrt((ji + 1) * (jf + 1))
                for mode in ("M1", "E2"):
                    if mode == "M1" and (ni != nf or li != lf):
                        continue
                    for ic, mc in enumerate(mc_s):
                        bit_c = 1 << (L - 1 - mc_idx[ic])
                        for ia, ma in enumerate(ma_s):
                            if ma != mc:
                                continue
                            bit_a = 1 << (L - 1 - ma_idx[ia])
                            if mode == "M1":
                                sign1 = (-1) ** ((jf - ma) // 2 + lf + (ji + 3) // 2)
                                sign2 = (-1) ** ((jf - ma) // 2 + lf + (jf + 3) // 2)
                                fac1 = sign1 * l6j * math.sqrt(lf * (lf + 1) * (2 * lf + 1))
                                fac2 = sign2 * s6j * math.sqrt(1.5)
                                coeff = pfac * wigner3j(jf, 2, ji, -ma, 0, mc) * (fac1 * gl + fac2 * gs) * c_mu
                            else:
                                signQ = (-1) ** (jf + (1 - ma) // 2) * (1 + (-1) ** (li + lf))
                                if signQ == 0:
                                    continue
                                coeff = signQ * wigner3j(jf, 4, ji, -ma, 0, mc) * Q3j * pfac * ec * c_Q * ex_r2
                                if sp == 1 and coeff == 0.0:
                                    continue
                            for bi, blist in enumerate(bits):
                                where = {b: k for k, b in enumerate(blist)}
                                for Ni, phi in enumerate(blist):
                                    ok, sgn, new = _phase_1(phi, bit_a, bit_c)
                                    if not ok or new not in where:
                                        continue
                                    Nf = where[new]
                                    out[mode][sp].append((bi, Ni, Nf, sgn == -1, coeff * (0.5 if Ni == Nf else 1.0)))
    return out


def op_obtr_diag(p_ops, n_ops, pbits, nbits, tdims, wf):
    """op_obtr_diag (transit.jl:247-283): twf = O wf (each entry applied both ways)."""
    wf = np.asarray(wf, dtype=np.float64)
    twf = np.zeros_like(wf)
    for bi, Ni, Nf, ph, fac in p_ops:
        c = -fac if ph else fac
        lN = len(nbits[bi])
        a, b = tdims[bi] + Ni * lN, tdims[bi] + Nf * lN
        twf[b:b + lN] += c * wf[a:a + lN]
        twf[a:a + lN] += c * wf[b:b + lN]
    for bi, Ni, Nf, ph, fac in n_ops:
        c = -fac if ph else fac
        lN = len(nbits[bi])
        nP = len(pbits[bi])
        a = tdims[bi] + Ni + lN * np.arange(nP)
        b = tdims[bi] + Nf + lN * np.arange(nP)
        twf[b] += c * wf[a]
        twf[a] += c * wf[b]
    return twf


def eval_moment(Mtot, Rvecs, totJs, ops, pbits, nbits, tdims):
    """eval_moment (transit.jl:702-765): per state n, (mu or None, Q or None), and the
    text main_sm prints.  totJs are J values (J < 0 skipped), Mtot is 2M."""
    c_mu = math.sqrt(4.0 * math.pi / 3.0)
    c_Q = math.sqrt(16.0 * math.pi / 5.0)
    res, tx_mom = [], ""
    M2 = int(Mtot)
    for nth, wf in enumerate(Rvecs):
        t1 = op_obtr_diag(ops["M1"][0], ops["M1"][1], pbits, nbits, tdims, wf)
        t2 = op_obtr_diag(ops["E2"][0], ops["E2"][1], pbits, nbits, tdims, wf)
        J2 = int(round(2 * totJs[nth]))
        mu = Q = None
        if J2 >= 0:
            tx = ""
            sel = wigner3j(J2, 2, J2, -M2, 0, M2)
            if sel != 0.0:
                cgv = wigner3j(J2, 2, J2, -J2, 0, J2) / sel
                mu = (-1) ** ((J2 - M2) // 2) * float(np.dot(t1, wf)) * cgv * c_mu
                tx += f"mu (n={nth + 1}) " + f"{mu:12.5f} " + "\t"
            sel = wigner3j(J2, 4, J2, -M2, 0, M2)
            if sel != 0.0:
                cgv = wigner3j(J2, 4, J2, -J2, 0, J2) / sel
                Q = float(np.dot(t2, wf)) * cgv * c_Q * (-1) ** ((J2 - M2) // 2)
                tx += f"Q (n={nth + 1}) " + f"{Q:12.5f} "
            if tx:
                tx_mom += tx + "\n"
        res.append((mu, Q))
    return res, tx_mom


def calc_bifs(Mps_l, Mns_l, Mps_r, Mns_r):
    """calc_bifs (transit.jl:501-534) from the blocks' (Mp, Mn) lists of occ: (p_bifs, n_bifs)
    as (left block, right block) pairs with equal Mn (proton operators) or equal Mp (neutron)."""
    p_bifs, n_bifs = [], []
    for ir in range(len(Mps_r)):
        for il in range(len(Mps_l)):
            if Mns_r[ir] == Mns_l[il]:
                p_bifs.append((il, ir))
            if Mps_r[ir] == Mps_l[il]:
                n_bifs.append((il, ir))
    return p_bifs, n_bifs


def prep_obtr_nondiag(jl2, jr2, p_sps, n_sps, msps_p, msps_n, pbits_l, nbits_l, pbits_r, nbits_r, p_bifs, n_bifs,
                      omu, bpar, gfactors, eff_charge):
    """prep_obtr_nondiag (transit.jl:286-445): the M1 and E2 operators' component omu from the
    right basis to the left one, {"M1": (proton, neutron), "E2": (...)} with entries
    (bl, Nl, br, Nr, phase, fac) (0-based; phase means the term enters negated)."""
    c_mu = math.sqrt(3.0 / (4.0 * math.pi))
    c_Q = math.sqrt(5.0 / (16.0 * math.pi))
    out = {"M1": ([], []), "E2": ([], [])}
    for sp, (sps, msps, bits_l, bits_r, bifs) in enumerate(((p_sps, msps_p, pbits_l, pbits_r, p_bifs),
                                                           (n_sps, msps_n, nbits_l, nbits_r, n_bifs))):
        gl, gs = (gfactors[0], gfactors[2]) if sp == 0 else (gfactors[1], gfactors[3])
        ec = eff_charge[sp]
        L = len(msps)
        for i in range(len(sps)):
            ni, li, ji, _ = sps[i]
            for f in range(len(sps)):
                nf, lf, jf, _ = sps[f]
                ex_r2 = radius_power(2, ni, li, nf, lf, bpar)
                l6j = wigner6j(2 * lf, 2 * li, 2, ji, jf, 1)
                s6j = wigner6j(1, 1, 2, ji, jf, 2 * lf)
                Q3j = wigner3j(ji, 4, jf, 1, 0, -1)
                _, mc_s, mc_idx = _possible_mz(sps[i], msps)
                _, ma_s, ma_idx = _possible_mz(sps[f], msps)
                for mode in ("M1", "E2"):
                    if mode == "M1":
                        if ni != nf or li != lf or (jl2 - jr2) // 2 > 1:
                            continue
                    elif (jl2 - jr2) // 2 > 2:
                        continue
                    for ic, mc in enumerate(mc_s):
                        bit_c = 1 << (L - 1 - mc_idx[ic])
                        for ia, ma in enumerate(ma_s):
                            if (ma - mc) // 2 != omu:
                                continue
                            bit_a = 1 << (L - 1 - ma_idx[ia])
                            pfac = math.sqrt((ji + 1) * (jf + 1))
                            if mode == "M1":
                                sign1 = (-1) ** ((jf - ma) // 2 + lf + (ji + 3) // 2)
                                sign2 = (-1) ** ((jf - ma) // 2 + lf + (jf + 3) // 2)
                                fac1 = sign1 * l6j * math.sqrt(lf * (lf + 1) * (2 * lf + 1))
                                fac2 = sign2 * s6j * math.sqrt(1.5)
                                coeff = pfac * wigner3j(jf, 2, ji, -ma, 2 * omu, mc) * (fac1 * gl + fac2 * gs) * c_mu
                            else:
                                signQ = (-1) ** (jf + (1 - ma) // 2) * (1 + (-1) ** (li + lf))
                                if signQ == 0:
                                    continue
                                coeff = signQ * wigner3j(jf, 4, ji, -ma, 2 * omu, mc) * Q3j * pfac * ec * c_Q * ex_r2
                            for bl, br in bifs:
                                where = {b: k for k, b in enumerate(bits_l[bl])}
                                for Nr, phi in enumerate(bits_r[br]):
                                    ok, sgn, new = _phase_1(phi, bit_a, bit_c)
                                    if not ok or new not in where:
                                        continue
                                    out[mode][sp].append((bl, where[new], br, Nr, sgn == -1, coeff))
    return out


def op_obtr_nondiag(p_ops, n_ops, pbits_l, nbits_l, tdims_l, pbits_r, nbits_r, tdims_r, wf):
    """op_obtr_nondiag (transit.jl:447-499): twf (left basis) = O wf (right basis)."""
    wf = np.asarray(wf, dtype=np.float64)
    twf = np.zeros(tdims_l[-1])
    for bl, Nl, br, Nr, ph, fac in p_ops:
        c = -fac if ph else fac
        nr_, nl_ = nbits_r[br], nbits_l[bl]
        where = {b: k for k, b in enumerate(nl_)}
        for k_r, n_r in enumerate(nr_):
            k_l = where.get(n_r)
            if k_l is None:
                continue
            twf[tdims_l[bl] + Nl * len(nl_) + k_l] += c * wf[tdims_r[br] + Nr * len(nr_) + k_r]
    for bl, Nl, br, Nr, ph, fac in n_ops:
        c = -fac if ph else fac
        pr_, pl_ = pbits_r[br], pbits_l[bl]
        lNr, lNl = len(nbits_r[br]), len(nbits_l[bl])
        where = {b: k for k, b in enumerate(pl_)}
        for k_r, p_r in enumerate(pr_):
            k_l = where.get(p_r)
            if k_l is None:
                continue
            twf[tdims_l[bl] + k_l * lNl + Nl] += c * wf[tdims_r[br] + k_r * lNr + Nr]
    return twf


def transitions(snt_path, nuc, jl2, jr2, gfactors=(1.0, 0.0, 5.586, -3.826), eff_charge=(1.5, 0.5), bpar=None):
    """transit_main's set-up (transit.jl:561-612): both bases (left at 2M = jl2, right at
    jr2 <= jl2), calc_bifs and prep_obtr_nondiag.  Returns a dict of the pieces."""
    snt = read_snt(snt_path, nuc)
    p_sps, n_sps = snt["p_sps"], snt["n_sps"]
    Z, N, vp, vn = get_zna(nuc, snt["cp"], snt["cn"])
    A = Z + N
    msps_p, msps_n = construct_msps(p_sps, n_sps)
    if bpar is None:
        bpar = bpar_main_sm(A)
    omu = (jl2 - jr2) // 2
    pl, nl, _, _, Mpl, Mnl, tdl = occ(p_sps, msps_p, vp, n_sps, msps_n, vn, jl2)
    pr, nr, _, _, Mpr, Mnr, tdr = occ(p_sps, msps_p, vp, n_sps, msps_n, vn, jr2)
    p_bifs, n_bifs = calc_bifs(Mpl, Mnl, Mpr, Mnr)
    ops = prep_obtr_nondiag(jl2, jr2, p_sps, n_sps, msps_p, msps_n, pl, nl, pr, nr, p_bifs, n_bifs, omu, bpar,
                            gfactors, eff_charge)
    return dict(ops=ops, pbits_l=pl, nbits_l=nl, tdims_l=tdl, pbits_r=pr, nbits_r=nr, tdims_r=tdr, omu=omu,
                jl2=jl2, jr2=jr2, bpar=bpar, msps_p=msps_p, msps_n=msps_n, p_sps=p_sps, n_sps=n_sps)


def transit(T, Vl, Vr, same_file=False, calc_EM=True):
    """transit_main's evaluation (transit.jl:612-697) on the pieces of `transitions`:
    records (mode, l, r, Mred, B(l<=r), B(l=>r), moment or None), 1-based l and r."""
    jl2, jr2, omu = T["jl2"], T["jr2"], T["omu"]
    c_mom = {"M1": math.sqrt(4.0 * math.pi / 3.0), "E2": math.sqrt(16.0 * math.pi / 5.0)}
    recs = []
    for mode, lam in (("M1", 1), ("E2", 2)):
        if not (abs(jl2 - jr2) <= 2 * lam <= jl2 + jr2):
            continue
        sel = wigner3j(jl2, 2 * lam, jr2, -jl2, 2 * omu, jr2)
        signmu = 1                                   # (-1)^(J_l - M_l) with M_l = J_l
        cgv = wigner3j(jr2, 2 * lam, jl2, -jr2, 2 * omu, jl2)
        if sel == 0:
            continue
        for r, wr in enumerate(Vr):
            twf = op_obtr_nondiag(T["ops"][mode][0], T["ops"][mode][1], T["pbits_l"], T["nbits_l"], T["tdims_l"],
                                  T["pbits_r"], T["nbits_r"], T["tdims_r"], wr)
            for l, wl in enumerate(Vl):
                if not calc_EM and l != r:
                    continue
                Mred = float(np.dot(wl, twf)) / sel
                mom = Mred * cgv * c_mom[mode] * signmu if ((same_file or not calc_EM) and l == r) else None
                recs.append((mode, l + 1, r + 1, Mred, Mred ** 2 / (jr2 + 1), Mred ** 2 / (jl2 + 1), mom))
    return recs

