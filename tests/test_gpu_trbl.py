"""GPU: thick-restart block Lanczos (main_sm with is_block=true, TRBL).

* The reference's own known-answer test (test/ShellModel_test.jl:9-15): ckpot
  Be8, 10 states with q=2 against its rounded Eref, (E - Eref)^2 < 1e-6, and
  against the dense eigenvalues of the brute-force H (1e-7 MeV FP64, 1e-5
  mixed).  Be8's 51-dimensional space is exhausted by the block Krylov basis
  before the ten levels meet the convergence rule, so this also covers the
  rank-loss stop.
* q = 3, 4 on usdb Ne20 (mdim 640) and Na21 against dense eigenvalues.
* lm small enough to force thick restarts (n_restart >= 1): Ne20 within
  1e-6 MeV (the restarted runs stop on the same per-step change rule as TRL).
* A synthetic sd case (Mg24, mdim 28503) against the single-vector TRL's
  FP64 energies (1e-7 MeV), with Ritz-vector residuals |H v - E v| and
  orthonormality.
* J-projected block Lanczos (target_J with is_block): Be8 2J=2 and Ne20 2J=4
  against the dense levels of that J.
* in_wf: a block start from the first q vectors of a .wav file reproduces
  the energies in fewer H.v products.
"""
import numpy as np
import pytest

import brute_force as bf
import smgpu
from smgpu import synth
from smgpu.shellmodel import j_from_jj1

pytestmark = pytest.mark.gpu

EREF_BE8 = [-31.119, -27.300, -19.162, -18.249, -16.722, -14.925, -14.517, -14.017, -13.951, -13.478]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("mixed", 1e-5)])
def test_trbl_ckpot_be8_known_answer(golden, precision, tol):
    path = f"{golden}/ckpot.snt"
    H, _ = bf.brute_hamiltonian(path, "Be8", 4, 4)
    ex = np.linalg.eigvalsh(H)[:10]
    en, st = smgpu.main_sm(path, "Be8", 10, [], q=2, is_block=True, precision=precision, return_stats=True)
    for i in range(10):
        assert (EREF_BE8[i] - en[i]) ** 2 < 1e-6
    np.testing.assert_allclose(en, ex, atol=tol)
    assert st["n_hv"] <= H.shape[0]


@pytest.mark.parametrize("q", [3, 4])
@pytest.mark.parametrize("fname,nuc,Z,N", [("usdb.snt", "Ne20", 10, 10), ("usdb.snt", "Na21", 11, 10)])
def test_trbl_block_sizes(golden, fname, nuc, Z, N, q):
    path = f"{golden}/{fname}"
    H, _ = bf.brute_hamiltonian(path, nuc, Z, N)
    ex = np.linalg.eigvalsh(H)[:8]
    en = smgpu.main_sm(path, nuc, 8, [], q=q, is_block=True, ls=4 * q)
    np.testing.assert_allclose(en, ex, atol=1e-7)
    en32 = smgpu.main_sm(path, nuc, 8, [], q=q, is_block=True, ls=4 * q, precision="mixed")
    np.testing.assert_allclose(en32, ex, atol=1e-5)


@pytest.mark.parametrize("q,lm,ls", [(2, 40, 10), (2, 30, 12), (3, 36, 12)])
def test_trbl_restarts(golden, q, lm, ls):
    path = f"{golden}/usdb.snt"
    H, _ = bf.brute_hamiltonian(path, "Ne20", 10, 10)
    ex = np.linalg.eigvalsh(H)[:10]
    en, st = smgpu.main_sm(path, "Ne20", 10, [], q=q, is_block=True, lm=lm, ls=ls, return_stats=True)
    assert st["n_restart"] >= 1
    np.testing.assert_allclose(en, ex, atol=1e-6)


def test_trbl_matches_trl_sd(snt_dir):
    path, nuc = synth.config_snt(f"{snt_dir}/mg24b.snt", "sd", 4, 4, synth.BASE_SEED + 7, (-4, 0), 1.5)
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    e1, _, _ = h.trl(6, precision="fp64")
    for q, prec, tol in [(2, "fp64", 1e-7), (3, "fp64", 1e-7), (2, "mixed", 1e-5)]:
        eb, R, st = h.trbl(6, q, precision=prec, want_vectors=True)
        np.testing.assert_allclose(eb, e1, atol=tol)
        assert np.abs(R @ R.T - np.eye(6)).max() < 1e-6
        for i in range(6):
            r = h.operate_H(R[i], "fp64") - eb[i] * R[i]
            assert np.linalg.norm(r) < 1e-3, (q, prec, i, np.linalg.norm(r))


def _dense_levels(path, nuc, Z, N, twoJ, k):
    H, _ = bf.brute_hamiltonian(path, nuc, Z, N, twoM=twoJ)
    J2 = bf.brute_j2(path, nuc, Z, N, twoM=twoJ)
    ev, U = np.linalg.eigh(H)
    out = [ev[i] for i in range(len(ev)) if j_from_jj1(float(U[:, i] @ J2 @ U[:, i]), 1e-6) == 0.5 * twoJ]
    return np.array(out[:k])


@pytest.mark.parametrize("fname,nuc,Z,N,twoJ,k", [("ckpot.snt", "Be8", 4, 4, 2, 2), ("usdb.snt", "Ne20", 10, 10, 4, 3)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("mixed", 1e-5)])
def test_trbl_j_projected(golden, fname, nuc, Z, N, twoJ, k, precision, tol):
    path = f"{golden}/{fname}"
    ref = _dense_levels(path, nuc, Z, N, twoJ, k)
    en, J, st = smgpu.main_sm(path, nuc, k, [twoJ], q=2, is_block=True, precision=precision, return_J=True,
                              return_stats=True, ls=10, lm=60)
    np.testing.assert_allclose(en, ref, atol=tol)
    assert all(x == 0.5 * twoJ for x in J), J
    assert st["n_jproj"] >= 2


def test_trbl_in_wf_block_start(snt_dir, tmp_path):
    path, nuc = synth.config_snt(f"{snt_dir}/mg24c.snt", "sd", 4, 4, synth.BASE_SEED + 8, (-4, 0), 1.5)
    wpath = str(tmp_path / "start.wav")
    en, st = smgpu.main_sm(path, nuc, 4, [], q=2, is_block=True, save_wav=True, wav_path=wpath, return_stats=True)
    en2, st2 = smgpu.main_sm(path, nuc, 2, [], q=2, is_block=True, in_wf=wpath, return_stats=True, ls=4)
    np.testing.assert_allclose(en2, en[:2], atol=1e-7)
    assert st2["n_hv"] < st["n_hv"]
