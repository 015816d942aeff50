"""GPU: the one-workgroup Lanczos step of small bases (mixed precision, partial
re-orthogonalisation, mdim <= 8192: k_step_small).

* usdb Ne20 (mdim 640) against the dense eigenvalues of the brute-force H
  (1e-5 MeV), with and without thick restarts.
* Synthetic sd Na22 (mdim 6,116) and pf Ti44 (4,000): the mixed solve against
  the plain FP64 solve (1e-6 MeV), default and restarting (lm 60) bases, with
  the omega recurrence's projection passes counted (n_reorth > 0), and the
  Ritz vectors' residuals |H v - E v| from the FP64 H.v.
* Lockstep batches on a small basis equal their solves alone, bit for bit.
"""
import numpy as np
import pytest

import brute_force as bf
import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lm,ls", [(300, 20), (40, 10)])
def test_small_step_ne20_dense(golden, lm, ls):
    path = f"{golden}/usdb.snt"
    H, _ = bf.brute_hamiltonian(path, "Ne20", 10, 10)
    ex = np.linalg.eigvalsh(H)[:8]
    en, st = smgpu.main_sm(path, "Ne20", 8, [], precision="mixed", lm=lm, ls=ls, return_stats=True)
    np.testing.assert_allclose(en, ex, atol=1e-5)
    if lm == 40:
        assert st["n_restart"] >= 1


@pytest.mark.parametrize("space,zv,nv", [("sd", 3, 3), ("pf", 2, 2)])
@pytest.mark.parametrize("lm,ls", [(300, 20), (60, 12)])
def test_small_step_vs_fp64(snt_dir, space, zv, nv, lm, ls):
    path, nuc = synth.config_snt(f"{snt_dir}/small_{space}{zv}{nv}.snt", space, zv, nv, synth.BASE_SEED + 90,
                                 (-6, 0), 1.2)
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    assert h.mdim <= 8192
    e64, _, _ = h.trl(6, precision="fp64", lm=lm, ls=ls, seed=3)
    e32, R, st = h.trl(6, precision="mixed", lm=lm, ls=ls, seed=3, want_vectors=True)
    np.testing.assert_allclose(e32, e64, atol=1e-6)
    assert st["n_reorth"] > 0
    for i in range(6):
        r = h.operate_H(R[i], "fp64") - e32[i] * R[i]
        assert np.linalg.norm(r) < 1e-3, (i, np.linalg.norm(r))


def test_small_step_batch_equals_solo(snt_dir):
    path, nuc = synth.config_snt(f"{snt_dir}/small_b.snt", "sd", 2, 4, synth.BASE_SEED + 91, (-6, 0), 1.2)
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    assert h.mdim <= 8192
    seeds = [5, 6, 7]
    out = h.trl_batch(5, seeds, precision="mixed", want_vectors=True)
    for sd, (ev, vec, st) in zip(seeds, out):
        ev1, vec1, st1 = h.trl(5, precision="mixed", seed=sd, want_vectors=True)
        assert np.array_equal(ev, ev1)
        assert np.array_equal(vec, vec1)
        assert st["iterations"] == st1["iterations"]
