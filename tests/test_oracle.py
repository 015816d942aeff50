"""Pins the CPU oracle (oracle/) before it is used as the parity checker.

* ckpot Be8 known answers          test/ShellModel_test.jl:10-14 (|dE| < 1e-3)
* Mg24 sd M=0 dimension 28,503     example/log_sample_script.txt:323
* dimensions vs a generating-function count (independent of `occ`)
* Clebsch-Gordan vs sympy
* dense H from the reference-layout tables == independent Jordan-Wigner H
* closed forms: SPE-only spectrum, mirror symmetry, 2M=0 vs 2M=2 spectra
"""
import itertools

import numpy as np
import pytest

import brute_force as bf
import sm_oracle as orc
from smgpu import synth

CKPOT_BE8 = [-31.119, -27.300, -19.162, -18.249, -16.722, -14.925, -14.517, -14.017, -13.951, -13.478]


def test_cg_matches_sympy():
    for j1 in range(1, 8, 2):
        for j2 in range(1, 8, 2):
            for J in range(abs(j1 - j2), j1 + j2 + 1, 2):
                for m1 in range(-j1, j1 + 1, 2):
                    for m2 in range(-j2, j2 + 1, 2):
                        M = m1 + m2
                        if abs(M) > J:
                            continue
                        a = orc.cg(j1, m1, j2, m2, J, M)
                        b = bf.cg_sympy(j1, m1, j2, m2, J, M)
                        assert abs(a - b) <= 2e-16 * max(1.0, abs(b)), (j1, m1, j2, m2, J, M, a, b)


def test_cg_orthogonality():
    # sum_{m1,m2} <j1 m1 j2 m2|J M><j1 m1 j2 m2|J' M> = delta_JJ'
    j1, j2 = 7, 5
    for M in range(-2, 3, 2):
        Js = [J for J in range(abs(j1 - j2), j1 + j2 + 1, 2) if abs(M) <= J]
        for J in Js:
            for Jp in Js:
                s = sum(orc.cg(j1, m1, j2, M - m1, J, M) * orc.cg(j1, m1, j2, M - m1, Jp, M)
                        for m1 in range(-j1, j1 + 1, 2) if abs(M - m1) <= j2)
                assert abs(s - (1.0 if J == Jp else 0.0)) < 1e-14


def test_ckpot_be8_known_answers(golden):
    en = orc.main_sm(f"{golden}/ckpot.snt", "Be8", 10)
    assert np.all((np.array(CKPOT_BE8) - en) ** 2 < 1e-6)


def _count_dim(space, zv, nv, twoM):
    """Generating function prod(1 + y x^m) per species (independent of occ)."""
    ms = [m for (_, _, j) in synth.SPACES[space] for m in range(-j, j + 1, 2)]
    def table(nv_):
        cnt = {}
        for comb in itertools.combinations(ms, nv_):
            cnt[sum(comb)] = cnt.get(sum(comb), 0) + 1
        return cnt
    tp, tn = table(zv), table(nv)
    return sum(c * tn.get(twoM - m, 0) for m, c in tp.items())


@pytest.mark.parametrize("space,zv,nv,expect", [("sd", 4, 4, 28503), ("pf", 4, 4, 1963461),
                                                  ("pf", 5, 5, 18600516), ("p", 2, 2, 51)])
def test_dimensions(space, zv, nv, expect):
    assert _count_dim(space, zv, nv, (zv + nv) % 2) == expect
    sps = [[n, l, j, -1] for (n, l, j) in synth.SPACES[space]]
    nps = [[n, l, j, 1] for (n, l, j) in synth.SPACES[space]]
    mp, mn = orc.construct_msps(sps, nps)
    *_, tdims = orc.occ(sps, mp, zv, nps, mn, nv, (zv + nv) % 2)
    assert tdims[-1] == expect


@pytest.mark.parametrize("fname,nuc,Z,N", [("ckpot.snt", "Be8", 4, 4), ("ckpot.snt", "Li7", 3, 4),
                                           ("usdb.snt", "O20", 8, 12), ("usdb.snt", "F19", 9, 10),
                                           ("random_snts/tmp_2.snt", "Ne20", 10, 10)])
def test_tables_equal_brute_force(golden, fname, nuc, Z, N):
    H, _ = bf.brute_hamiltonian(f"{golden}/{fname}", nuc, Z, N)
    h = orc.OracleHamiltonian(f"{golden}/{fname}", nuc)
    Ho = h.dense()
    assert Ho.shape == H.shape
    assert np.abs(Ho - H).max() < 1e-12 * max(1.0, np.abs(H).max())


def test_synthetic_pf_small_equals_brute_force(snt_dir):
    path, nuc = synth.config_snt(f"{snt_dir}/pf11.snt", "pf", 1, 2, 7, (-8, 0), 1.0)
    Z, N = 21, 22
    H, _ = bf.brute_hamiltonian(path, nuc, Z, N)
    Ho = orc.OracleHamiltonian(path, nuc).dense()
    assert np.abs(Ho - H).max() < 1e-12


def test_spe_only_closed_form(snt_dir):
    inter = synth.make_interaction("sd", 3, (-4, 0), 1.0)
    inter.lines = [(a, b, c, d, J, 0.0) for (a, b, c, d, J, v) in inter.lines]
    path = synth.write_snt(f"{snt_dir}/speonly.snt", inter, 20)
    h = orc.OracleHamiltonian(path, "O20")      # 0 protons, 4 neutrons
    ev = np.linalg.eigvalsh(h.dense())
    ms = [(orb, m) for orb, (_, _, j) in enumerate(synth.SPACES["sd"]) for m in range(-j, j + 1, 2)]
    expect = sorted(sum(inter.spe[ms[k][0]] for k in comb)
                    for comb in itertools.combinations(range(len(ms)), 4)
                    if sum(ms[k][1] for k in comb) == 0)
    assert np.allclose(ev, expect, atol=1e-12)


def test_mirror_and_m_invariance(snt_dir):
    path, _ = synth.config_snt(f"{snt_dir}/sd_mirror.snt", "sd", 2, 3, 11, (-4, 0), 1.5)
    # F21 (vp=1, vn=4) and its mirror Mg21 (vp=4, vn=1): SPE_p = SPE_n, pp = nn
    e1 = np.linalg.eigvalsh(orc.OracleHamiltonian(path, "F21").dense())
    e2 = np.linalg.eigvalsh(orc.OracleHamiltonian(path, "Mg21").dense())
    assert len(e1) == len(e2) and np.allclose(e1, e2, atol=1e-10)
    pathm, _ = synth.config_snt(f"{snt_dir}/sd_m.snt", "sd", 2, 2, 12, (-4, 0), 1.5)
    h0 = orc.OracleHamiltonian(pathm, "Ne20", Mtot=0)
    h2 = orc.OracleHamiltonian(pathm, "Ne20", Mtot=2)
    e0 = np.linalg.eigvalsh(h0.dense())
    e2m = np.linalg.eigvalsh(h2.dense())
    # every 2M=2 level appears in the 2M=0 spectrum
    for x in e2m:
        assert np.min(np.abs(e0 - x)) < 1e-9


def test_trl_matches_dense(golden):
    h = orc.OracleHamiltonian(f"{golden}/usdb.snt", "Ne22")
    ev = np.linalg.eigvalsh(h.dense())[:5]
    en, elit, nhv, nres, *_ = orc.trl(h, 5, lm=40, ls=10)
    assert nres > 0                                    # exercises ThickRestart
    assert np.allclose(en, ev, atol=1e-7)
