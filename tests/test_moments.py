"""Moments (main_sm's calc_moment): the oracle's restatement of prep_obtr_diag /
op_obtr_diag / eval_moment (transit.jl:49-283, 702-765), pinned by closed forms,
and the product's host tables against it (CPU only).

* One valence nucleon (F17, O17 in the sd space of usdb): the states are pure
  orbit states, and the moments are the textbook single-particle values:
  Schmidt mu = g_l l + g_s / 2 for j = l + 1/2 and j / (j + 1) (g_l (l + 1) - g_s / 2)
  for j = l - 1/2; Q = -(2j - 1) / (2j + 2) <r^2> e with <r^2> = b^2 (2n + l + 3/2).
* radius_power against the oscillator expectation values, 3j / 6j symbols
  against their symmetry and orthogonality relations.
* The product's 6j / 3j / radius_power / oscillator length (C ABI-free host
  checks through the oracle's exact values) are exercised by the GPU tests.
"""
import math

import numpy as np
import pytest

import sm_oracle as orc

G = [1.0, 0.0, 5.586, -3.826]
EC = [1.5, 0.5]


def _schmidt(l, j2, gl, gs):
    j = j2 / 2
    if j2 == 2 * l + 1:
        return gl * l + gs / 2
    return j / (j + 1) * (gl * (l + 1) - gs / 2)


def _q_sp(n, l, j2, b, e):
    j = j2 / 2
    return -(2 * j - 1) / (2 * j + 2) * b * b * (2 * n + l + 1.5) * e


@pytest.mark.parametrize("nuc,sp", [("F17", 0), ("O17", 1)])
def test_single_nucleon_closed_forms(golden, nuc, sp):
    h = orc.OracleHamiltonian(f"{golden}/usdb.snt", nuc)
    assert h.mdim == 3
    w, V = np.linalg.eigh(h.dense())
    b = orc.bpar_main_sm(17)
    ops = orc.prep_obtr_diag(h.sm["p_sps"], h.sm["n_sps"], h.msps_p, h.msps_n, h.pbits, h.nbits, b, G, EC)
    sps = h.sm["p_sps"] if sp == 0 else h.sm["n_sps"]
    # each eigenvector is one orbit's m = 1/2 state: its orbit from the basis SD
    bits = h.pbits if sp == 0 else h.nbits
    msps = h.msps_p if sp == 0 else h.msps_n
    L = len(msps)
    orbs = []
    for k in range(3):
        x = int(np.argmax(np.abs(V[:, k])))
        assert abs(abs(V[x, k]) - 1.0) < 1e-12
        blk = max(i for i in range(len(h.tdims) - 1) if h.tdims[i] <= x)
        sd = bits[blk][(x - h.tdims[blk]) // len(h.nbits[blk])] if sp == 0 else bits[blk][(x - h.tdims[blk]) % len(h.nbits[blk])]
        m = next(i for i in range(L) if sd >> (L - 1 - i) & 1)
        orbs.append(msps[m][5] - 1)
    Js = [sps[o][2] / 2 for o in orbs]
    res, tx = orc.eval_moment(h.Mtot, V.T, Js, ops, h.pbits, h.nbits, h.tdims)
    gl, gs = (G[0], G[2]) if sp == 0 else (G[1], G[3])
    for k, o in enumerate(orbs):
        n, l, j2, _ = sps[o]
        mu, Q = res[k]
        assert abs(mu - _schmidt(l, j2, gl, gs)) < 1e-12, (nuc, sps[o], mu)
        if j2 == 1:
            assert Q is None or abs(Q) < 1e-12
        else:
            assert abs(Q - _q_sp(n, l, j2, b, EC[sp])) < 1e-10, (nuc, sps[o], Q)
    assert tx.count("mu (n=") == 3


def test_radius_power_and_symbols():
    for n, l in [(0, 0), (0, 1), (1, 0), (0, 2), (1, 2), (2, 1), (0, 3), (1, 3)]:
        assert abs(orc.radius_power(2, n, l, n, l, 1.3) - 1.69 * (2 * n + l + 1.5)) < 1e-12
    # <0s|r^2|1s> = -sqrt(3/2) b^2 for oscillator functions with a positive slope at the origin
    assert abs(abs(orc.radius_power(2, 0, 0, 1, 0, 1.0)) - math.sqrt(1.5)) < 1e-12
    # 6j orthogonality: sum_x (2x+1)(2c+1) {a b x; d e c}{a b x; d e c'} = delta_cc'
    a, b_, d, e = 3, 5, 3, 5
    for c in (2, 4, 6):
        for c2 in (2, 4, 6):
            s = sum((x + 1) * (c + 1) * orc.wigner6j(a, b_, x, d, e, c) * orc.wigner6j(a, b_, x, d, e, c2)
                    for x in range(2, 9, 2))
            assert abs(s - (1.0 if c == c2 else 0.0)) < 1e-13
    # 3j orthogonality over m1 m2 for fixed (j3, m3)
    j1, j2 = 5, 4
    for j3 in (1, 3, 5, 7, 9):
        s = sum(orc.wigner3j(j1, j2, j3, m1, m2, -(m1 + m2)) ** 2 * (j3 + 1)
                for m1 in range(-j1, j1 + 1, 2) for m2 in range(-j2, j2 + 1, 2) if m1 + m2 == 1)
        assert abs(s - 1.0) < 1e-13
