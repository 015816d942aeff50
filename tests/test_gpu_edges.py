"""GPU: the smallest bases and the argument limits.

* He4 (no valence particles, mdim 1), He5 (mdim 2), He6 (5) and Li6 (10) on
  ckpot: num_ev above mdim is clamped to mdim (shellmodel_main.jl:128), the
  Krylov space is exhausted at step mdim, and every level equals the dense
  diagonalisation of the brute-force H: 1e-7 MeV FP64, 1e-5 MeV mixed (the
  one-workgroup step of small bases), by plain TRL, block TRL (q = 2 where
  mdim >= 2) and a lockstep batch.
* He6 at 2J = 0 (a J space of two states): both levels, and an error when
  three are asked for.
* A block larger than the basis, num_ev above 64, ls above 64, ls >= lm - 1 and
  lm above 1000 fail with an error; the handle stays usable afterwards.
* Restart and Ritz sizes past one rotation pass (lanczos_methods.jl:250-307 has
  no limit on ls or num_ev): on a seeded sd 2p2n interaction (Ne20, mdim 640),
  TRL with ls 40 and num_ev 36, TRL with num_ev 12 above ls 8 (a restart then
  keeps 12 vectors), and block TRL (q = 2) with ls 40, and with num_ev 12 above
  ls 6, all give the dense levels (1e-7 MeV FP64, 1e-5 MeV mixed).
"""
import numpy as np
import pytest

import brute_force as bf
import smgpu

pytestmark = pytest.mark.gpu

TINY = [("He4", 2, 2, 1), ("He5", 2, 3, 2), ("He6", 2, 4, 5), ("Li6", 3, 3, 10)]
TOL = {"fp64": 1e-7, "mixed": 1e-5}
# |H x - E x| of the Ritz vectors: the mixed basis is kept semi-orthogonal (FP32
# vectors, partial re-orthogonalisation), so its vectors carry errors of order
# sqrt(eps_32) |H| while its eigenvalues are accurate to eps_32 |H|
RES = {"fp64": 1e-7, "mixed": 1e-3}


@pytest.mark.parametrize("nuc,Z,N,mdim", TINY)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiny_bases_all_levels(golden, nuc, Z, N, mdim, precision):
    path = f"{golden}/ckpot.snt"
    print("case", nuc, precision, flush=True)
    H, _ = bf.brute_hamiltonian(path, nuc, Z, N)
    assert H.shape == (mdim, mdim)
    ex = np.linalg.eigvalsh(H)
    en = smgpu.main_sm(path, nuc, 10, [], precision=precision)
    assert len(en) == mdim
    np.testing.assert_allclose(en, ex, atol=TOL[precision])
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    for ev, vec, st in h.trl_batch(10, [3, 4], precision=precision, want_vectors=True):
        np.testing.assert_allclose(ev, ex, atol=TOL[precision])
        assert st["n_restart"] == 0
        # Ritz vectors of an exhausted space are eigenvectors
        for i in range(mdim):
            assert np.linalg.norm(h.operate_H(vec[i], "fp64") - ex[i] * vec[i]) < RES[precision]
    if mdim >= 2:
        eb = smgpu.main_sm(path, nuc, 10, [], precision=precision, is_block=True, q=2)
        np.testing.assert_allclose(eb, ex, atol=TOL[precision])


def test_argument_limits(golden):
    path = f"{golden}/ckpot.snt"
    with pytest.raises(smgpu.SmgpuError):
        smgpu.main_sm(path, "He4", 1, [], is_block=True, q=2)          # q > mdim
    h = smgpu.Hamiltonian(smgpu.Problem(path, "Be8"))
    for kw in (dict(num_ev=65), dict(num_ev=5, lm=20, ls=19), dict(num_ev=5, lm=1001), dict(num_ev=5, lm=100, ls=65)):
        with pytest.raises(smgpu.SmgpuError):
            h.trl(**kw)
    ev, _, _ = h.trl(5)
    H, _ = bf.brute_hamiltonian(path, "Be8", 4, 4)
    np.testing.assert_allclose(ev, np.linalg.eigvalsh(H)[:5], atol=1e-7)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiny_j_space(golden, precision):
    """He6 holds two J = 0 states (dense: -3.909812 and 7.921112 MeV): both are found,
    and asking for three fails (the J space is exhausted, as the reference's
    tbeta < Jtol stop, lanczos_methods.jl:63-64)."""
    path = f"{golden}/ckpot.snt"
    H, _ = bf.brute_hamiltonian(path, "He6", 2, 4, twoM=0)
    ev = np.linalg.eigvalsh(H)
    en, J = smgpu.main_sm(path, "He6", 2, target_J=[0], precision=precision, return_J=True, ls=10, lm=60)
    np.testing.assert_allclose(en, [ev[0], ev[4]], atol=TOL[precision])
    assert list(J) == [0.0, 0.0]
    with pytest.raises(smgpu.SmgpuError):
        smgpu.main_sm(path, "He6", 3, target_J=[0], precision=precision, ls=10, lm=60)


@pytest.fixture(scope="module")
def ne20_dense(tmp_path_factory):
    from smgpu import synth
    path, nuc = synth.config_snt(str(tmp_path_factory.mktemp("ne20") / "ne20.snt"), "sd", 2, 2, 4242, (-6, 0), 1.2)
    H, _ = bf.brute_hamiltonian(path, nuc, 10, 10)
    assert H.shape == (640, 640)
    return path, nuc, np.linalg.eigvalsh(H)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_and_ritz_beyond_one_pass(ne20_dense, precision):
    path, nuc, ex = ne20_dense
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    print("case ls40 nev36", precision, flush=True)
    ev, _, st = h.trl(36, lm=100, ls=40, precision=precision, max_iter=5000)
    assert st["n_restart"] >= 1, st
    np.testing.assert_allclose(ev, ex[:36], atol=TOL[precision])
    print("case nev12 > ls8", precision, flush=True)
    ev, _, st = h.trl(12, lm=80, ls=8, precision=precision, max_iter=5000)
    assert st["n_restart"] >= 1, st
    np.testing.assert_allclose(ev, ex[:12], atol=TOL[precision])
    print("case block q2 ls40", precision, flush=True)
    eb = smgpu.main_sm(path, nuc, 20, [], precision=precision, is_block=True, q=2, lm=100, ls=40)
    np.testing.assert_allclose(eb, ex[:20], atol=TOL[precision])
    print("case block q2 nev12 > ls6", precision, flush=True)
    eb = smgpu.main_sm(path, nuc, 12, [], precision=precision, is_block=True, q=2, lm=80, ls=6)
    np.testing.assert_allclose(eb, ex[:12], atol=TOL[precision])
