"""GPU: the J^2 operator (smgpu_ham_create_j2), main_sm's J values and .wav I/O.

* J^2 H.v (FP64) against the ladder-operator J^2 of tests/brute_force.py on
  ckpot Be8/Li7 and usdb Ne21 (relative 1e-12), and against the oracle's
  operate_H! with the J^2 interaction on a synthetic pf case (1e-12).
* main_sm(return_J=True): the J of each of ckpot Be8's ten lowest states equals
  the J of the dense eigenvector (brute-force H and J^2), FP64 and mixed.
* save_wav / in_wf: the written file reads back bit-exactly, and a restart from
  the ground-state vector reproduces the energies (1e-7 MeV) in fewer steps.
"""
import numpy as np
import pytest

import brute_force as bf
import sm_oracle as orc
import smgpu
from smgpu import synth, wav
from smgpu.shellmodel import expect_j2, j_from_jj1

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fname,nuc,Z,N", [("ckpot.snt", "Be8", 4, 4), ("ckpot.snt", "Li7", 3, 4),
                                           ("usdb.snt", "Ne21", 10, 11)])
def test_j2_matches_ladder_operators(golden, fname, nuc, Z, N):
    path = f"{golden}/{fname}"
    J2 = bf.brute_j2(path, nuc, Z, N)
    p = smgpu.Problem(path, nuc)
    hj = p.j2_hamiltonian()
    rng = np.random.default_rng(2)
    for _ in range(3):
        v = rng.standard_normal(p.mdim)
        ref = J2 @ v
        w = hj.operate_H(v, "fp64")
        assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max(), nuc


def test_j2_matches_oracle_pf(snt_dir):
    path, nuc = synth.config_snt(f"{snt_dir}/pf33.snt", "pf", 3, 3, 4242, (-8, 0), 1.0)
    j2p = bf.write_j2_snt(path, f"{snt_dir}/pf33_j2.snt", nuc)
    o = orc.OracleHamiltonian(j2p, nuc)
    p = smgpu.Problem(path, nuc)
    hj = p.j2_hamiltonian()
    v = np.random.default_rng(3).standard_normal(p.mdim)
    ref = o.operate_H(v)
    w = hj.operate_H(v, "fp64")
    assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max()
    w32 = hj.operate_H(v, "mixed")
    assert np.abs(w32 - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_main_sm_j_values_be8(golden, precision):
    path = f"{golden}/ckpot.snt"
    H, _ = bf.brute_hamiltonian(path, "Be8", 4, 4)
    J2 = bf.brute_j2(path, "Be8", 4, 4)
    ev, U = np.linalg.eigh(H)
    ref_J = [j_from_jj1(float(U[:, i] @ J2 @ U[:, i])) for i in range(10)]
    assert -1.0 not in ref_J
    en, J = smgpu.main_sm(path, "Be8", 10, precision=precision, return_J=True)
    np.testing.assert_allclose(en, ev[:10], atol=1e-7)
    assert list(J) == ref_J, (list(J), ref_J)


def test_wav_save_and_warm_start(snt_dir, tmp_path):
    path, nuc = synth.config_snt(f"{snt_dir}/mg24w.snt", "sd", 4, 4, synth.BASE_SEED + 1, (-4, 0), 1.5)
    wpath = str(tmp_path / "mg24.wav")
    en, J, st = smgpu.main_sm(path, nuc, 5, return_J=True, save_wav=True, wav_path=wpath, return_stats=True)
    es, jj, vecs, mtot = wav.read_wav(wpath, 28503)
    assert np.array_equal(es, en) and mtot == 0
    assert jj.tolist() == [int(round(2 * x)) for x in J]
    assert np.allclose(np.linalg.norm(vecs, axis=1), 1.0, atol=1e-12)
    assert np.allclose(expect_j2(smgpu.Problem(path, nuc), vecs[:1]), J[0] * (J[0] + 1), atol=1e-6)
    en2, st2 = smgpu.main_sm(path, nuc, 1, in_wf=wpath, return_stats=True)
    assert abs(en2[0] - en[0]) < 1e-7
    assert st2["n_hv"] < st["n_hv"]
