"""Independent brute-force M-scheme Hamiltonian (test helper, small spaces only).

Built straight from the interaction file with second quantisation on one
Fock space (proton modes first, then neutron modes; Jordan-Wigner signs from
the number of occupied modes before the target mode), with sympy's
Clebsch-Gordan coefficients.  It shares no code with the oracle: two-body terms
are  V_J(ab,cd) * sum_M A+_JM(ab) A_JM(cd)  with normalised pair operators
A+_JM(ab) = (1+d_ab)^(-1/2) sum_{ma,mb} <ja ma jb mb|J M> a+_a,ma a+_b,mb  for
like particles and no normalisation for proton-neutron pairs; the
Hermitian-conjugate term is added for every line with {a,b} != {c,d}
(readsmsnt, shellmodel_main.jl:354-359).  Basis states are ordered like the
reference (proton M ascending; within a block proton then neutron
determinants in descending bit-integer order, first m-state = most
significant bit; shellmodel_main.jl:732-848).
"""
from __future__ import annotations

import itertools
import re
from functools import lru_cache

import numpy as np
from sympy import Rational
from sympy.physics.quantum.cg import CG


@lru_cache(maxsize=None)
def cg_sympy(j1, m1, j2, m2, J, M):
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J or J > j1 + j2 or J < abs(j1 - j2):
        return 0.0
    return float(CG(Rational(j1, 2), Rational(m1, 2), Rational(j2, 2), Rational(m2, 2),
                    Rational(J, 2), Rational(M, 2)).doit())


def _parse(path, nuc):
    A = int(re.search(r"[0-9]+", nuc).group())
    rows = []
    for line in open(path):
        s = line.strip()
        if not s or s[0] in "!#":
            continue
        rows.append(s.split())
    lp, ln, cp, cn = map(int, rows[0][:4])
    orbs = [tuple(map(int, rows[1 + i][:5])) for i in range(lp + ln)]   # idx n l j tz
    nsp = int(rows[1 + lp + ln][0])
    spe = {}
    for i in range(nsp):
        r = rows[2 + lp + ln + i]
        spe[int(r[0])] = float(r[2])
    hdr = rows[2 + lp + ln + nsp]
    ntb, massop, aref = int(hdr[0]), int(hdr[1]), float(hdr[2])
    pw = float(hdr[3]) if len(hdr) > 3 else 0.0
    lines = []
    for i in range(ntb):
        r = rows[3 + lp + ln + nsp + i]
        v = float(r[5])
        if massop == 1:
            v *= (A / aref) ** pw
        lines.append((int(r[0]), int(r[1]), int(r[2]), int(r[3]), int(r[4]), v))
    return lp, ln, cp, cn, orbs, spe, lines, A


def brute_hamiltonian(path, nuc, Z, N, twoM=None):
    """Dense H (reference basis order) and the list of (pbits, nbits) per state."""
    lp, ln, cp, cn, orbs, spe, lines, A = _parse(path, nuc)
    vp, vn = Z - cp, N - cn
    if twoM is None:
        twoM = A % 2
    modes = []          # (orbit index 1-based, 2m); protons first
    for (idx, n, l, j, tz) in orbs:
        for m in range(-j, j + 1, 2):
            modes.append((idx, m))
    Lp = sum(orbs[i][3] + 1 for i in range(lp))
    Ln = len(modes) - Lp
    mode_of = {mk: k for k, mk in enumerate(modes)}
    jof = {o[0]: o[3] for o in orbs}

    states = []
    for pc in itertools.combinations(range(Lp), vp):
        for nc in itertools.combinations(range(Lp, Lp + Ln), vn):
            M = sum(modes[k][1] for k in pc) + sum(modes[k][1] for k in nc)
            if M != twoM:
                continue
            occ = 0
            for k in pc + nc:
                occ |= 1 << k
            pint = sum(1 << (Lp - 1 - k) for k in pc)
            nint = sum(1 << (Ln - 1 - (k - Lp)) for k in nc)
            Mp = sum(modes[k][1] for k in pc)
            states.append((Mp, -pint, -nint, occ, pint, nint))
    states.sort()
    index = {s[3]: i for i, s in enumerate(states)}
    dim = len(states)
    H = np.zeros((dim, dim))

    def annihilate(occ, k):
        if not (occ >> k) & 1:
            return None, 0
        sgn = -1 if bin(occ & ((1 << k) - 1)).count("1") % 2 else 1
        return occ & ~(1 << k), sgn

    def create(occ, k):
        if (occ >> k) & 1:
            return None, 0
        sgn = -1 if bin(occ & ((1 << k) - 1)).count("1") % 2 else 1
        return occ | (1 << k), sgn

    def apply(ops, occ):
        """ops applied right-to-left: list of ('c'|'a', mode)."""
        s = 1
        for kind, k in reversed(ops):
            occ, t = (create if kind == "c" else annihilate)(occ, k)
            if occ is None:
                return None, 0
            s *= t
        return occ, s

    for (idx, n, l, j, tz) in orbs:
        e = spe.get(idx, 0.0)
        for m in range(-j, j + 1, 2):
            k = mode_of[(idx, m)]
            for i, st in enumerate(states):
                if (st[3] >> k) & 1:
                    H[i, i] += e

    def pair_terms(a, b, J):
        """(coef, mode_a, mode_b) of A+_JM(ab) for all M, normalisation included."""
        ja, jb = jof[a], jof[b]
        like = (a <= lp) == (b <= lp)
        norm = 1.0 / np.sqrt(2.0) if (like and a == b) else 1.0
        out = {}
        for ma in range(-ja, ja + 1, 2):
            for mb in range(-jb, jb + 1, 2):
                M = ma + mb
                if abs(M) > 2 * J:
                    continue
                c = cg_sympy(ja, ma, jb, mb, 2 * J, M)
                if c == 0.0:
                    continue
                out.setdefault(M, []).append((norm * c, mode_of[(a, ma)], mode_of[(b, mb)]))
        return out

    def add_term(a, b, c, d, J, V):
        Aab = pair_terms(a, b, J)
        Acd = pair_terms(c, d, J)
        for M, lst in Aab.items():
            for (c1, ka, kb) in lst:
                for (c2, kc, kd) in Acd.get(M, []):
                    # A+(ab) A(cd) = a+_a a+_b a_d a_c
                    ops = [("c", ka), ("c", kb), ("a", kd), ("a", kc)]
                    for i, st in enumerate(states):
                        f, s = apply(ops, st[3])
                        if f is None:
                            continue
                        fi = index.get(f)
                        if fi is None:
                            continue
                        H[fi, i] += V * c1 * c2 * s

    for (a, b, c, d, J, V) in lines:
        add_term(a, b, c, d, J, V)
        if sorted({a, b}) != sorted({c, d}):
            add_term(c, d, a, b, J, V)
    return H, [(s[4], s[5]) for s in states]


def _sector(orbs, lp, vp, vn, twoM):
    """Modes and reference-ordered states of one M sector (as brute_hamiltonian)."""
    modes = []
    for (idx, n, l, j, tz) in orbs:
        for m in range(-j, j + 1, 2):
            modes.append((idx, m))
    Lp = sum(orbs[i][3] + 1 for i in range(lp))
    Ln = len(modes) - Lp
    states = []
    for pc in itertools.combinations(range(Lp), vp):
        for nc in itertools.combinations(range(Lp, Lp + Ln), vn):
            M = sum(modes[k][1] for k in pc) + sum(modes[k][1] for k in nc)
            if M != twoM:
                continue
            occ = 0
            for k in pc + nc:
                occ |= 1 << k
            pint = sum(1 << (Lp - 1 - k) for k in pc)
            nint = sum(1 << (Ln - 1 - (k - Lp)) for k in nc)
            Mp = sum(modes[k][1] for k in pc)
            states.append((Mp, -pint, -nint, occ))
    states.sort()
    return modes, [s[3] for s in states]


def brute_j2(path, nuc, Z, N, twoM=None):
    """Dense J^2 = J- J+ + Jz^2 + Jz in the reference basis order, from the ladder
    operator J+ = sum_{orbit, m} sqrt(j(j+1) - m(m+1)) a+_{m+1} a_m (Jordan-Wigner
    signs as in brute_hamiltonian); independent of any two-body bookkeeping."""
    lp, ln, cp, cn, orbs, spe, lines, A = _parse(path, nuc)
    vp, vn = Z - cp, N - cn
    if twoM is None:
        twoM = A % 2
    modes, lo = _sector(orbs, lp, vp, vn, twoM)
    _, hi = _sector(orbs, lp, vp, vn, twoM + 2)
    mode_of = {mk: k for k, mk in enumerate(modes)}
    jof = {o[0]: o[3] for o in orbs}
    hidx = {occ: i for i, occ in enumerate(hi)}
    Jp = np.zeros((len(hi), len(lo)))
    for i, occ in enumerate(lo):
        for k, (idx, m) in enumerate(modes):
            if not (occ >> k) & 1 or m + 2 > jof[idx]:
                continue
            k2 = mode_of[(idx, m + 2)]
            if (occ >> k2) & 1:
                continue
            # a+_{k2} a_k |occ>
            s1 = -1 if bin(occ & ((1 << k) - 1)).count("1") % 2 else 1
            o1 = occ & ~(1 << k)
            s2 = -1 if bin(o1 & ((1 << k2) - 1)).count("1") % 2 else 1
            f = o1 | (1 << k2)
            j, mm = jof[idx] / 2.0, m / 2.0
            Jp[hidx[f], i] += s1 * s2 * np.sqrt(j * (j + 1) - mm * (mm + 1))
    Mz = twoM / 2.0
    return Jp.T @ Jp + (Mz * Mz + Mz) * np.eye(len(lo))


def write_j2_snt(model_path, out_path, nuc):
    """The J^2 operator as an interaction file on the orbits of model_path:
    SPE j(j+1); diagonal TBMEs J(J+1) - ja(ja+1) - jb(jb+1) for pp, nn, pn."""
    lp, ln, cp, cn, orbs, spe, lines, A = _parse(model_path, nuc)
    jof = {o[0]: o[3] for o in orbs}
    jj = lambda j2: 0.5 * j2 * (0.5 * j2 + 1)
    tb = []
    for lo_, hi_ in ((1, lp), (lp + 1, lp + ln)):
        for a in range(lo_, hi_ + 1):
            for b in range(a, hi_ + 1):
                for J in range(abs(jof[a] - jof[b]) // 2, (jof[a] + jof[b]) // 2 + 1):
                    if a == b and J % 2:
                        continue
                    tb.append((a, b, a, b, J, J * (J + 1) - jj(jof[a]) - jj(jof[b])))
    for a in range(1, lp + 1):
        for b in range(lp + 1, lp + ln + 1):
            for J in range(abs(jof[a] - jof[b]) // 2, (jof[a] + jof[b]) // 2 + 1):
                tb.append((a, b, a, b, J, J * (J + 1) - jj(jof[a]) - jj(jof[b])))
    out = [f"{lp} {ln} {cp} {cn}"]
    out += [" ".join(map(str, o)) for o in orbs]
    out.append(f"{lp + ln} 0")
    out += [f"{o[0]} {o[0]} {jj(o[3]):.12e}" for o in orbs]
    out.append(f"{len(tb)} 0 {A}")
    out += [f"{a} {b} {c} {d} {J} {v:.12e}" for a, b, c, d, J, v in tb]
    with open(out_path, "w") as f:
        f.write("\n".join(out) + "\n")
    return out_path
