"""Transitions (transit_main): the oracle's restatement of prep_obtr_nondiag /
op_obtr_nondiag / transit_main (transit.jl:286-534, 561-700), pinned on the CPU.

* One valence nucleon (F17, O17 in the sd space of usdb): every m-component
  <a m_a| O_mu |c m_c> of the M1 and E2 operators (mu = 0, 1, 2, both bases at any
  2M) is rebuilt independently of the reference's 3j / 6j formulas: the orbit
  states expanded into |l m_l> |1/2 m_s> by Clebsch-Gordan coefficients, the
  M1 operator sqrt(3 / 4pi) (g_l l_mu + g_s s_mu) from angular-momentum ladder
  matrix elements, and E2 = e r^2 Y_2mu from Gaunt integrals times the radial
  <r^2> (radius_power, pinned by the oscillator values in test_moments.py).  Per
  orbit pair the oracle's matrix elements equal these times one sign, the same for
  every m (a convention of the reduced matrix element's phase), and the sign is +1
  for a = c.
* B(E2; j_i -> j_f) of the stretched single-particle states from transit_main's
  records equals e^2 <r^2>^2 (5 / 4pi) (2j_f + 1) (j_f 2 j_i; 1/2 0 -1/2)^2.
* mu = 0 on a many-particle basis (Ne20, sd 2p2n, 2M = 2): the non-diagonal
  operators equal prep_obtr_diag's (the moments operators, pinned by the Schmidt
  values) on random vectors, so the fermion phases of many-particle SDs agree, and
  <gm>, <eQ> of one state given twice equal eval_moment's.
"""
import math

import numpy as np
import pytest
from sympy import Rational, sqrt as ssqrt
from sympy.physics.wigner import clebsch_gordan, gaunt

import sm_oracle as orc

G = [1.0, 0.0, 5.586, -3.826]
EC = [1.5, 0.5]


def _unit(T, side, sd_bit, sp):
    """The basis vector of one valence nucleon in the m-state bit `sd_bit` (side 'l' or 'r')."""
    pb, nb, td = T["pbits_" + side], T["nbits_" + side], T["tdims_" + side]
    v = np.zeros(td[-1])
    for b in range(len(pb)):
        for ip, p in enumerate(pb[b]):
            for jn, n in enumerate(nb[b]):
                if (p if sp == 0 else n) == sd_bit:
                    v[td[b] + ip * len(nb[b]) + jn] = 1.0
    assert v.sum() == 1.0
    return v


def _ladder(j2, m2a, m2c, mu):
    """<j m_a| J_mu |j m_c> of the spherical components J_{+1} = -J_+/sqrt2, J_0 = J_z,
    J_{-1} = J_-/sqrt2 (doubled j, m)."""
    j, ma, mc = Rational(j2, 2), Rational(m2a, 2), Rational(m2c, 2)
    if mu == 0:
        return mc if ma == mc else 0
    if mu == 1 and ma == mc + 1:
        return -ssqrt(j * (j + 1) - mc * (mc + 1)) / ssqrt(2)
    if mu == -1 and ma == mc - 1:
        return ssqrt(j * (j + 1) - mc * (mc - 1)) / ssqrt(2)
    return 0


def _brute(orb_a, orb_c, m2a, m2c, mu, mode, gl, gs, e, r2):
    """<a m_a| O_mu |c m_c> from the orbit states' |l m_l>|1/2 m_s> expansions."""
    na, la, ja = orb_a[:3]
    nc, lc, jc = orb_c[:3]
    tot = 0
    for msa in (-1, 1):
        mla = m2a - msa
        if abs(mla) > 2 * la:
            continue
        cga = clebsch_gordan(la, Rational(1, 2), Rational(ja, 2), mla // 2, Rational(msa, 2), Rational(m2a, 2))
        for msc in (-1, 1):
            mlc = m2c - msc
            if abs(mlc) > 2 * lc:
                continue
            cgc = clebsch_gordan(lc, Rational(1, 2), Rational(jc, 2), mlc // 2, Rational(msc, 2), Rational(m2c, 2))
            if cga == 0 or cgc == 0:
                continue
            if mode == "M1":
                if na != nc or la != lc:
                    continue
                x = 0
                if msa == msc:
                    x += gl * _ladder(2 * la, mla, mlc, mu)
                if mla == mlc:
                    x += gs * _ladder(1, msa, msc, mu)
                tot += cga * cgc * x * ssqrt(Rational(3, 4) / math.pi)
            else:
                if msa != msc:
                    continue
                # <l_a m_a| Y_2mu |l_c m_c> = (-1)^m_a * int Y_{l_a,-m_a} Y_{2,mu} Y_{l_c,m_c}
                y = (-1) ** (abs(mla // 2) % 2) * gaunt(la, 2, lc, -(mla // 2), mu, mlc // 2)
                tot += cga * cgc * y * e * r2
    return float(tot)


@pytest.mark.parametrize("nuc,sp", [("F17", 0), ("O17", 1)])
def test_single_nucleon_matrix_elements(golden, nuc, sp):
    snt = f"{golden}/usdb.snt"
    b = orc.bpar_main_sm(17)
    T0 = orc.transitions(snt, nuc, 1, 1, G, EC, b)
    sps = T0["p_sps"] if sp == 0 else T0["n_sps"]
    msps = T0["msps_p"] if sp == 0 else T0["msps_n"]
    L = len(msps)
    gl, gs = (G[0], G[2]) if sp == 0 else (G[1], G[3])
    e = EC[sp]
    signs = {}
    checked = 0
    for ia, ma_state in enumerate(msps):
        for ic, mc_state in enumerate(msps):
            m2a, m2c = ma_state[4], mc_state[4]
            mu2 = m2a - m2c
            if mu2 < 0 or mu2 > 4:
                continue
            oa, oc = ma_state[5] - 1, mc_state[5] - 1
            T = orc.transitions(snt, nuc, m2a, m2c, G, EC, b)
            wl = _unit(T, "l", 1 << (L - 1 - ia), sp)
            wr = _unit(T, "r", 1 << (L - 1 - ic), sp)
            for mode in ("M1", "E2"):
                ops = T["ops"][mode]
                got = float(wl @ orc.op_obtr_nondiag(ops[0], ops[1], T["pbits_l"], T["nbits_l"], T["tdims_l"],
                                                      T["pbits_r"], T["nbits_r"], T["tdims_r"], wr))
                r2 = orc.radius_power(2, sps[oa][0], sps[oa][1], sps[oc][0], sps[oc][1], b)
                want = _brute(sps[oa], sps[oc], m2a, m2c, mu2 // 2, mode, gl, gs, e, r2)
                if mode == "M1" and mu2 > 2:
                    assert got == 0.0
                    continue
                if abs(want) < 1e-12:
                    assert abs(got) < 1e-12, (nuc, mode, oa, oc, m2a, m2c, got)
                    continue
                s = got / want
                assert abs(abs(s) - 1.0) < 1e-10, (nuc, mode, oa, oc, m2a, m2c, got, want)
                key = (mode, oa, oc)
                signs.setdefault(key, round(s))
                assert round(s) == signs[key], (key, m2a, m2c, s)
                checked += 1
    assert checked > 100
    for (mode, oa, oc), s in signs.items():
        if oa == oc:
            assert s == 1, (mode, oa)


@pytest.mark.parametrize("nuc,sp", [("F17", 0), ("O17", 1)])
def test_single_nucleon_be2_closed_form(golden, nuc, sp):
    snt = f"{golden}/usdb.snt"
    b = orc.bpar_main_sm(17)
    probe = orc.transitions(snt, nuc, 1, 1, G, EC, b)
    sps = probe["p_sps"] if sp == 0 else probe["n_sps"]
    msps = probe["msps_p"] if sp == 0 else probe["msps_n"]
    L = len(msps)
    n_pairs = 0
    for oa, (na, la, ja, _) in enumerate(sps):
        for oc, (nc, lc, jc, _) in enumerate(sps):
            if ja < jc or ja - jc > 4 or ja + jc < 4 or (la + lc) % 2:
                continue
            T = orc.transitions(snt, nuc, ja, jc, G, EC, b)
            ia = next(k for k, m in enumerate(msps) if m[5] - 1 == oa and m[4] == ja)
            ic = next(k for k, m in enumerate(msps) if m[5] - 1 == oc and m[4] == jc)
            wl = _unit(T, "l", 1 << (L - 1 - ia), sp)
            wr = _unit(T, "r", 1 << (L - 1 - ic), sp)
            recs = [x for x in orc.transit(T, [wl], [wr]) if x[0] == "E2"]
            assert len(recs) == 1
            r2 = orc.radius_power(2, na, la, nc, lc, b)
            want = EC[sp] ** 2 * r2 ** 2 * 5 / (4 * math.pi) * (ja + 1) * orc.wigner3j(ja, 4, jc, 1, 0, -1) ** 2
            assert abs(recs[0][4] - want) <= 1e-12 * max(1.0, want), (oa, oc, recs[0][4], want)
            n_pairs += 1
    assert n_pairs >= 4


def test_mu0_equals_moment_operators(golden):
    snt = f"{golden}/usdb.snt"
    h = orc.OracleHamiltonian(snt, "Ne20", Mtot=2)
    b = orc.bpar_main_sm(20)
    T = orc.transitions(snt, "Ne20", 2, 2, G, EC, b)
    assert T["tdims_l"] == h.tdims and T["pbits_l"] == h.pbits and T["nbits_l"] == h.nbits
    diag = orc.prep_obtr_diag(h.sm["p_sps"], h.sm["n_sps"], h.msps_p, h.msps_n, h.pbits, h.nbits, b, G, EC)
    rng = np.random.default_rng(5)
    for _ in range(3):
        v = rng.standard_normal(h.mdim)
        v /= np.linalg.norm(v)
        for mode in ("M1", "E2"):
            a = orc.op_obtr_diag(diag[mode][0], diag[mode][1], h.pbits, h.nbits, h.tdims, v)
            o = T["ops"][mode]
            c = orc.op_obtr_nondiag(o[0], o[1], T["pbits_l"], T["nbits_l"], T["tdims_l"], T["pbits_r"], T["nbits_r"],
                                    T["tdims_r"], v)
            assert np.abs(a - c).max() <= 1e-13 * max(1.0, np.abs(a).max()), mode
        # one state given twice at M = J = 1: transit's <gm>, <eQ> equal eval_moment's
        mom = {x[0]: x[6] for x in orc.transit(T, [v], [v], same_file=True)}
        res, _ = orc.eval_moment(2, [v], [1.0], diag, h.pbits, h.nbits, h.tdims)
        assert abs(mom["M1"] - res[0][0]) < 1e-12 and abs(mom["E2"] - res[0][1]) < 1e-12


def test_create_rejects_mismatched_problems(golden, tmp_path):
    """smgpu_transitions_create checks its problems before any device work: a right problem
    read from an interaction file that lists the same orbit counts in another order, a right
    basis of higher 2M, and |M_left - M_right| above 2 are errors."""
    import smgpu
    snt = f"{golden}/usdb.snt"
    lines = open(snt).read().splitlines()
    # swap the proton orbits 2 (0d3/2) and 3 (0d5/2): same counts, other m-state bits
    lines[2], lines[3] = lines[3].replace("   3   0   2   5", "   2   0   2   5"), lines[2].replace("   2   0   2   3", "   3   0   2   3")
    swapped = tmp_path / "usdb_swapped.snt"
    swapped.write_text("\n".join(lines) + "\n")
    p0, p2, p6 = (smgpu.Problem(snt, "Ne20", twoM=m) for m in (0, 2, 6))
    for left, right, msg in ((p2, smgpu.Problem(str(swapped), "Ne20", twoM=0), "different nuclei or spaces"),
                             (p2, smgpu.Problem(snt, "Ne20", twoM=4), "left basis needs 2M >="),
                             (p6, p0, "only M1 and E2")):
        with pytest.raises(smgpu.SmgpuError, match=msg):
            smgpu.Transitions(left, right, 0, G, EC)
