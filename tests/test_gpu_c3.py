"""GPU parity at the largest single-GPU workload: Mn50 (pf 5p5n, 2M = 0,
mdim 18,600,516; neutron blocks up to 1,651 determinants, so the kernel runs
its wide-block loop, NPT = 8) and on small bases with wide neutron blocks.

* H·v through the C ABI against the oracle's C operate_H!
  (lanczos_methods.jl:153-248): FP64 <= 1e-12 and FP32 <= 1e-5 of |H v|_inf,
  plus symmetry <u, H v> = <H u, v> and linearity at full size.
* The lowest 5 eigenvalues with a 100-vector thick-restart basis (the
  configuration's lm) against tests/golden/eigs_mn50.json (oracle operate_H! +
  ARPACK, make_golden.py): FP64 <= 1e-7 MeV, mixed <= 1e-5 MeV.
* Wide neutron blocks on small bases (Sc46 pf 1p5n: widest block 1,651, NPT 8;
  Sc47 pf 1p6n: 3,952, NPT 16; Ti48 pf 2p6n): H·v against the oracle in both
  precisions and TRL eigenvalues against ARPACK over the oracle's operate_H!;
  two-vector launches (one Hamiltonian, two interactions) and a lockstep solve of
  two interactions equal the single-vector ones bit for bit; block Lanczos with
  q = 2 (two-vector H·V) against the same ARPACK levels.
"""
import json

import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, eigsh

import sm_oracle as orc
import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mn50(tmp_path_factory):
    path, nuc = synth.config_snt(str(tmp_path_factory.mktemp("mn50") / "mn50.snt"), "pf", 5, 5,
                                 synth.BASE_SEED + 3, (-8, 0), 1.0)
    assert nuc == "Mn50"
    p = smgpu.Problem(path, nuc)
    assert p.mdim == 18600516
    return path, nuc, p, smgpu.Hamiltonian(p)


def test_hv_parity_mn50(mn50):
    path, nuc, p, h = mn50
    print("oracle tables", nuc, flush=True)
    o = orc.OracleHamiltonian(path, nuc)
    v = np.random.default_rng(50).standard_normal(p.mdim)
    print("oracle operate_H", flush=True)
    ref = o.operate_H(v)
    del o
    w64 = h.operate_H(v, "fp64")
    w32 = h.operate_H(v, "mixed")
    scale = np.abs(ref).max()
    e64 = np.abs(w64 - ref).max() / scale
    e32 = np.abs(w32 - ref).max() / scale
    print("mn50 rel err fp64", e64, "fp32", e32, flush=True)
    assert e64 <= 1e-12 and e32 <= 1e-5
    u = np.random.default_rng(51).standard_normal(p.mdim)
    hu = h.operate_H(u, "fp64")
    assert abs(u @ w64 - hu @ v) <= 1e-11 * (np.abs(u) @ np.abs(w64))
    hsum = h.operate_H(u - 0.5 * v, "fp64")
    assert np.abs(hsum - (hu - 0.5 * w64)).max() <= 1e-11 * scale


def test_trl_mn50_vs_golden(mn50, golden):
    _, nuc, p, h = mn50
    g = json.load(open(f"{golden}/eigs_mn50.json"))
    assert (g["nucleus"], g["mdim"], g["seed"]) == (nuc, p.mdim, synth.BASE_SEED + 3)
    ref = np.array(g["eigenvalues_MeV"])
    k = len(ref)
    for prec, tol in (("fp64", 1e-7), ("mixed", 1e-5)):
        en, _, st = h.trl(k, lm=100, ls=20, precision=prec)
        print("mn50", prec, np.abs(en - ref).max(), st["iterations"], st["n_restart"], st["n_hv"], flush=True)
        assert st["n_restart"] > 0
        assert np.abs(en - ref).max() < tol, (prec, en - ref)


WIDE = [("pf", 1, 5, 8), ("pf", 1, 6, 16), ("pf", 2, 6, 16)]


@pytest.mark.parametrize("space,zv,nv,npt", WIDE)
def test_wide_blocks_hv_and_trl(snt_dir, space, zv, nv, npt):
    path, nuc = synth.config_snt(f"{snt_dir}/wide_{zv}{nv}.snt", space, zv, nv, synth.BASE_SEED + 40 + 10 * zv + nv,
                                 (-6, 0), 1.2)
    p = smgpu.Problem(path, nuc)
    b = p.basis()
    widest = max(len(x) for x in b["nbits"])
    need = (widest + 255) // 256
    assert need <= npt and 2 * need > npt
    h = smgpu.Hamiltonian(p)
    o = orc.OracleHamiltonian(path, nuc)
    rng = np.random.default_rng(zv * 10 + nv)
    for _ in range(2):
        v = rng.standard_normal(p.mdim)
        ref = o.operate_H(v)
        scale = np.abs(ref).max()
        e64 = np.abs(h.operate_H(v, "fp64") - ref).max() / scale
        e32 = np.abs(h.operate_H(v, "mixed") - ref).max() / scale
        print(nuc, "widest", widest, "mdim", p.mdim, "err", e64, e32, flush=True)
        assert e64 <= 1e-12 and e32 <= 1e-5
    # two vectors per launch, of one Hamiltonian and of two interactions on the basis (the
    # 512-thread FP32 shapes of five-neutron blocks run two-vector launches), equal single
    # launches bit for bit, and so does a lockstep solve of the two interactions
    path2, _ = synth.config_snt(f"{snt_dir}/wide_{zv}{nv}_b.snt", space, zv, nv,
                                synth.BASE_SEED + 41 + 10 * zv + nv, (-6, 0), 1.2)
    hv = h.variant(smgpu.Problem(path2, nuc))
    V = rng.standard_normal((2, p.mdim))
    for prec in ("fp64", "mixed"):
        W1 = np.stack([h.operate_H(V[0], prec), h.operate_H(V[1], prec)])
        assert np.array_equal(h.operate_H_multi(V, prec), W1), prec
        Wv = smgpu.operate_H_variants([h, hv], V, prec)
        assert np.array_equal(Wv[0], W1[0]) and np.array_equal(Wv[1], hv.operate_H(V[1], prec)), prec
    out = smgpu.trl_variants([h, hv], 4, [1, 2], precision="mixed")
    for hh, sd, (ev2, _, st2) in zip((h, hv), (1, 2), out):
        ev1, _, st1 = hh.trl(4, precision="mixed", seed=sd)
        assert np.array_equal(ev2, ev1) and st2["n_hv"] == st1["n_hv"], (sd, ev2 - ev1)
    op = LinearOperator((o.mdim, o.mdim), matvec=lambda x: o.operate_H(np.ascontiguousarray(x).ravel()),
                        dtype=np.float64)
    ev = np.sort(eigsh(op, k=4, which="SA", tol=1e-13, v0=np.ones(o.mdim))[0])
    for prec, tol in (("fp64", 1e-7), ("mixed", 1e-5)):
        en, _, st = h.trl(4, precision=prec)
        print(nuc, prec, en - ev, st["iterations"], flush=True)
        assert np.abs(en - ev).max() < tol
    # block Lanczos, q = 2: its H·V runs two-vector launches (six-neutron blocks: at 1024 threads)
    en, _, st = h.trbl(4, 2, precision="mixed")
    print(nuc, "trbl q=2 mixed", en - ev, st["iterations"], flush=True)
    assert np.abs(en - ev).max() < 1e-5
