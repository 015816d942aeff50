"""CPU checks of the J^2 operator and the .wav format (no GPU).

* J^2 written as an interaction (one-body j(j+1), diagonal TBMEs
  J(J+1) - ja(ja+1) - jb(jb+1) for pp, nn, pn) equals J- J+ + Jz^2 + Jz built
  from ladder operators, on ckpot Be8/Li7 and usdb Ne21 (brute force, 1e-12);
  its spectrum is {J(J+1)}.  This is the operator the library builds in
  smgpu_ham_create_j2 (sm_host.cpp: j2_operator) for main_sm's Js
  (shellmodel_main.jl:149-157, 226-230).
* J_from_JJ1 (shellmodel_main.jl:1105-1113) on exact and perturbed values.
* .wav round trip in writeRitzvecs' layout (shellmodel_main.jl:1215-1229):
  Int32 n, Int32 Mtot, Float64 E[n], Int32 2J[n], n x mdim Float64, byte-exact.
"""
import numpy as np
import pytest

import brute_force as bf
from smgpu import wav
from smgpu.shellmodel import j_from_jj1


@pytest.mark.parametrize("fname,nuc,Z,N", [("ckpot.snt", "Be8", 4, 4), ("ckpot.snt", "Li7", 3, 4),
                                           ("usdb.snt", "Ne21", 10, 11)])
def test_j2_interaction_equals_ladder_operators(golden, tmp_path, fname, nuc, Z, N):
    path = f"{golden}/{fname}"
    J2 = bf.brute_j2(path, nuc, Z, N)
    H, _ = bf.brute_hamiltonian(bf.write_j2_snt(path, str(tmp_path / "j2.snt"), nuc), nuc, Z, N)
    assert np.abs(J2 - H).max() < 1e-12
    ev = np.linalg.eigvalsh(J2)
    J = (-1 + np.sqrt(1 + 4 * np.clip(ev, 0, None))) / 2
    assert np.abs(2 * J - np.round(2 * J)).max() < 1e-9


def test_j_from_jj1():
    assert j_from_jj1(0.0) == 0.0
    assert j_from_jj1(6.0) == 2.0
    assert j_from_jj1(0.75) == 0.5
    assert j_from_jj1(15.75 + 5e-7) == 3.5
    assert j_from_jj1(6.0 + 1e-5) == -1.0


def test_wav_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    vecs = rng.standard_normal((3, 17))
    es = np.array([-10.5, -9.25, -3.0])
    p = wav.write_ritzvecs(str(tmp_path / "x.wav"), 0, es, [0.0, 2.0, 1.5], vecs)
    raw = open(p, "rb").read()
    assert len(raw) == 8 + 3 * 8 + 3 * 4 + 3 * 17 * 8
    assert np.frombuffer(raw[:8], "<i4").tolist() == [3, 0]
    assert np.frombuffer(raw[32:44], "<i4").tolist() == [0, 4, 3]
    e2, jj, v2, mtot = wav.read_wav(p, 17)
    assert np.array_equal(e2, es) and jj.tolist() == [0, 4, 3] and np.array_equal(v2, vecs) and mtot == 0
    assert np.array_equal(wav.read_appwav(p, 17), vecs[0])
    with pytest.raises(ValueError):
        wav.read_wav(p, 18)


def test_snt_params_and_tdmat_layout(golden, tmp_path):
    """The TD mode's file handling (CPU): an interaction file rewritten with its
    mass-scaled TBMEs and massop 0 gives the same host tables as the original;
    the TD file round-trips in the reference's layout (Int64 header, per pair
    i, j, Nij, N_ij, TDs) with idx_from_ij's 1-based pair numbering."""
    import numpy as np
    import smgpu
    from smgpu import ec
    p = ec.read_snt_params(f"{golden}/usdb.snt")
    assert (p.lp, p.ln, len(p.tbme), p.massop) == (3, 3, 158, 1)
    f = ec.write_snt_params(str(tmp_path / "usdb_scaled.snt"), p, p.spe, ec.scaled_tbmes(p, "Na23"))
    a = smgpu.Problem(f"{golden}/usdb.snt", "Na23")
    b = smgpu.Problem(f, "Na23")
    for s in (0, 1):
        xa, xb = a.pair_matrix(s), b.pair_matrix(s)
        for u, v in zip(xa[:3], xb[:3]):
            assert np.array_equal(u, v)
        assert np.abs(xa[3] - xb[3]).max() <= 1e-14 * np.abs(xa[3]).max()
    rng = np.random.default_rng(3)
    dim, P = 4, 7
    N = rng.normal(size=(dim, dim))
    N = N + N.T
    TD = rng.normal(size=(P, dim, dim))
    ec.write_tdmat(str(tmp_path / "x.dat"), N, TD)
    raw = np.fromfile(str(tmp_path / "x.dat"), dtype="<i8", count=3)
    assert list(raw) == [dim, 10, P]
    d, npair, lTD, Nmat, infos, TDs = ec.read_tdmat(str(tmp_path / "x.dat"))
    assert (d, npair, lTD) == (dim, 10, P)
    assert infos[:5] == [(1, 1, 1), (1, 2, 2), (1, 3, 3), (1, 4, 4), (2, 2, 5)]
    for (i, j, _), td in zip(infos, TDs):
        assert Nmat[i - 1, j - 1] == N[i - 1, j - 1] and np.array_equal(td, TD[:, i - 1, j - 1])
