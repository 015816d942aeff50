"""GPU parity: H·v and TRL through the C ABI against the oracle.

Tolerances: FP64 H·v relative max-norm error <= 1e-12; FP32 H·v <= 1e-5 of
|H v|_inf; eigenvalues vs oracle / dense <= 1e-7 MeV (FP64 path) and
<= 1e-5 MeV (mixed path, the tolerance at which served solves are compared with FP64).
"""
import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, eigsh

import brute_force as bf
import sm_oracle as orc
import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu

CKPOT_BE8 = [-31.119, -27.300, -19.162, -18.249, -16.722, -14.925, -14.517, -14.017, -13.951, -13.478]


def _cases(golden, snt_dir):
    c1, _ = synth.config_snt(f"{snt_dir}/mg24.snt", "sd", 4, 4, synth.BASE_SEED + 1, (-4, 0), 1.5)
    pfs, _ = synth.config_snt(f"{snt_dir}/pf23.snt", "pf", 2, 3, 99, (-8, 0), 1.0)
    return [(f"{golden}/ckpot.snt", "Be8"), (f"{golden}/ckpot.snt", "Li7"), (f"{golden}/usdb.snt", "O20"),
            (f"{golden}/usdb.snt", "Na23"), (c1, "Mg24"), (pfs, "V48")]


def test_hv_parity_fp64_fp32(golden, snt_dir):
    rng = np.random.default_rng(11)
    for path, nuc in _cases(golden, snt_dir):
        print("case", nuc, flush=True)
        p = smgpu.Problem(path, nuc)
        h = smgpu.Hamiltonian(p)
        o = orc.OracleHamiltonian(path, nuc)
        for trial in range(2):
            v = rng.standard_normal(p.mdim)
            ref = o.operate_H(v)
            w64 = h.operate_H(v, "fp64")
            w32 = h.operate_H(v, "mixed")
            scale = np.abs(ref).max()
            assert np.abs(w64 - ref).max() <= 1e-12 * scale, (nuc, np.abs(w64 - ref).max() / scale)
            assert np.abs(w32 - ref).max() <= 1e-5 * scale, (nuc, np.abs(w32 - ref).max() / scale)


def test_hv_parity_cr48(snt_dir):
    path, nuc = synth.config_snt(f"{snt_dir}/cr48.snt", "pf", 4, 4, synth.BASE_SEED + 2, (-8, 0), 1.0)
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    o = orc.OracleHamiltonian(path, nuc)
    v = np.random.default_rng(3).standard_normal(p.mdim)
    ref = o.operate_H(v)
    w64 = h.operate_H(v, "fp64")
    w32 = h.operate_H(v, "mixed")
    scale = np.abs(ref).max()
    assert np.abs(w64 - ref).max() <= 1e-12 * scale
    assert np.abs(w32 - ref).max() <= 1e-5 * scale
    # linearity and symmetry at full size: <u, H v> == <H u, v>
    u = np.random.default_rng(4).standard_normal(p.mdim)
    hu = h.operate_H(u, "fp64")
    assert abs(u @ w64 - hu @ v) <= 1e-11 * abs(u) @ np.abs(w64)
    hsum = h.operate_H(u + 2.0 * v, "fp64")
    assert np.abs(hsum - (hu + 2.0 * w64)).max() <= 1e-11 * scale


@pytest.mark.parametrize("space,zv,nv,seed", [("sd", 4, 4, 11), ("pf", 3, 3, 12), ("pf", 4, 2, 13)])
def test_hv_multi_bit_identical(snt_dir, space, zv, nv, seed):
    """smgpu_apply_H_multi (bl_operate_H!: 2 or 4 vectors per launch sharing the
    table loads) gives each vector bit for bit what the one-vector kernel gives,
    for 1..9 vectors (launches of 4, 2 and 1), FP64 and FP32, on one-row and
    two-row group shapes."""
    path, nuc = synth.config_snt(f"{snt_dir}/multi_{space}{zv}{nv}.snt", space, zv, nv, synth.BASE_SEED + seed,
                                 (-6, 0), 1.2)
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((9, h.mdim))
    for prec in ("fp64", "mixed"):
        ref = np.array([h.operate_H(v, prec) for v in V])
        for k in (1, 2, 3, 4, 5, 9):
            W = h.operate_H_multi(V[:k], prec)
            assert np.array_equal(W, ref[:k]), (prec, k, np.abs(W - ref[:k]).max())


def test_dense_equals_brute_force(golden):
    H, _ = bf.brute_hamiltonian(f"{golden}/ckpot.snt", "Be8", 4, 4)
    p = smgpu.Problem(f"{golden}/ckpot.snt", "Be8")
    h = smgpu.Hamiltonian(p)
    G = np.zeros_like(H)
    e = np.zeros(p.mdim)
    for k in range(p.mdim):
        e[:] = 0
        e[k] = 1
        G[:, k] = h.operate_H(e, "fp64")
    assert np.abs(G - H).max() < 1e-12


def test_trl_ckpot_be8(golden):
    for prec in ("fp64", "mixed"):
        en = smgpu.main_sm(f"{golden}/ckpot.snt", "Be8", 10, precision=prec)
        assert np.all((np.array(CKPOT_BE8) - en) ** 2 < 1e-6), (prec, en)
        o = orc.OracleHamiltonian(f"{golden}/ckpot.snt", "Be8")
        ev = np.linalg.eigvalsh(o.dense())[:10]
        assert np.abs(en - ev).max() < (1e-8 if prec == "fp64" else 1e-5)


def test_trl_restarts_match_dense(golden):
    o = orc.OracleHamiltonian(f"{golden}/usdb.snt", "Ne22")
    ev = np.linalg.eigvalsh(o.dense())[:6]
    p = smgpu.Problem(f"{golden}/usdb.snt", "Ne22")
    h = smgpu.Hamiltonian(p)
    for prec in ("fp64", "mixed"):
        en, vec, st = h.trl(6, lm=40, ls=10, precision=prec, want_vectors=True)
        assert st["n_restart"] > 0
        assert np.abs(en - ev).max() < (1e-7 if prec == "fp64" else 1e-5), (prec, en - ev)
        # Ritz vectors are unit eigenvectors of the oracle's H.  The stopping
        # rule bounds the Ritz-value change (1e-8), so the residual is only
        # bounded by sqrt(dE * gap) ~ 1e-3; the eigenvector overlap is checked too.
        _, evec = np.linalg.eigh(o.dense())
        for j in range(6):
            r = o.operate_H(vec[j]) - ev[j] * vec[j]
            assert abs(np.linalg.norm(vec[j]) - 1) < 1e-10
            assert np.linalg.norm(r) < (3e-3 if prec == "fp64" else 1e-2)
            assert abs(abs(vec[j] @ evec[:, j]) - 1) < 1e-5


def test_trl_mg24_vs_arpack(snt_dir):
    path, nuc = synth.config_snt(f"{snt_dir}/mg24.snt", "sd", 4, 4, synth.BASE_SEED + 1, (-4, 0), 1.5)
    o = orc.OracleHamiltonian(path, nuc)
    op = LinearOperator((o.mdim, o.mdim), matvec=lambda x: o.operate_H(np.ascontiguousarray(x).ravel()),
                        dtype=np.float64)
    ev = np.sort(eigsh(op, k=5, which="SA", tol=1e-13, v0=np.ones(o.mdim))[0])
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    en64, _, st64 = h.trl(5, precision="fp64")
    enmx, _, stmx = h.trl(5, precision="mixed")
    print("fp64", en64 - ev, st64["iterations"], "mixed", enmx - ev, stmx["iterations"], flush=True)
    assert np.abs(en64 - ev).max() < 1e-7
    assert np.abs(enmx - ev).max() < 1e-5


def _golden_cfg(golden, name):
    import json
    g = json.load(open(f"{golden}/eigs_{name}.json"))
    return g


@pytest.mark.parametrize("name,space,zv,nv,spe,sig", [("mg24", "sd", 4, 4, (-4, 0), 1.5),
                                                      ("cr48", "pf", 4, 4, (-8, 0), 1.0)])
def test_trl_vs_golden_arpack(golden, snt_dir, name, space, zv, nv, spe, sig):
    """Eigenvalues of the C1/C2 workloads vs the committed oracle+ARPACK values
    (tests/golden/make_golden.py): FP64 path <= 1e-7 MeV, mixed path <= 1e-5 MeV."""
    g = _golden_cfg(golden, name)
    path, nuc = synth.config_snt(f"{snt_dir}/{name}.snt", space, zv, nv, g["seed"], spe, sig)
    assert nuc == g["nucleus"]
    p = smgpu.Problem(path, nuc)
    assert p.mdim == g["mdim"]
    h = smgpu.Hamiltonian(p)
    ref = np.array(g["eigenvalues_MeV"])
    k = len(ref)
    for prec, tol in (("fp64", 1e-7), ("mixed", 1e-5)):
        en, _, st = h.trl(k, precision=prec)
        print(name, prec, np.abs(en - ref).max(), st["iterations"], st["n_restart"], flush=True)
        assert np.abs(en - ref).max() < tol, (prec, en - ref)
    # a restarted solve (small lm) reaches the same values
    en, _, st = h.trl(min(k, 5), lm=40, ls=max(10, min(k, 5)), precision="mixed")
    assert st["n_restart"] > 0
    assert np.abs(en - ref[:len(en)]).max() < 1e-5


@pytest.mark.parametrize("name,space,zv,nv,spe,sig", [("mg24", "sd", 4, 4, (-4, 0), 1.5),
                                                      ("cr48", "pf", 4, 4, (-8, 0), 1.0)])
def test_partial_reorth(golden, snt_dir, name, space, zv, nv, spe, sig):
    """Partial re-orthogonalisation (omega recurrence): the mixed default runs a
    projection pass on a fraction of the steps, full runs one every step, and both
    reach the ARPACK goldens (1e-5 MeV mixed); FP64 with partial keeps 1e-7 MeV.
    A restarted partial solve (lm 40) reaches the same values."""
    g = _golden_cfg(golden, name)
    path, nuc = synth.config_snt(f"{snt_dir}/{name}_pr.snt", space, zv, nv, g["seed"], spe, sig)
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    ref = np.array(g["eigenvalues_MeV"])
    k = len(ref)
    en_p, _, st_p = h.trl(k, precision="mixed")
    en_f, _, st_f = h.trl(k, precision="mixed", reorth="full")
    print(name, "partial", st_p["n_reorth"], "/", st_p["n_hv"], np.abs(en_p - ref).max(),
          "full", st_f["n_reorth"], "/", st_f["n_hv"], np.abs(en_f - ref).max(), flush=True)
    assert np.abs(en_p - ref).max() < 1e-5 and np.abs(en_f - ref).max() < 1e-5
    assert st_f["n_reorth"] == st_f["n_hv"]
    assert 0 < st_p["n_reorth"] < st_p["n_hv"] // 2
    en64, _, st64 = h.trl(k, precision="fp64", reorth="partial")
    assert np.abs(en64 - ref).max() < 1e-7 and st64["n_reorth"] < st64["n_hv"]
    en_r, _, st_r = h.trl(min(k, 5), lm=40, ls=10, precision="mixed")
    assert st_r["n_restart"] > 0 and st_r["n_reorth"] < st_r["n_hv"]
    assert np.abs(en_r - ref[:len(en_r)]).max() < 1e-5


def test_lanczos_recurrence_matches_numpy(snt_dir):
    """alpha/beta of the device recurrence vs numpy CGS2 Lanczos with the same
    (device) H·v, at the Cr48 size: FP64 <= 1e-10, mixed <= 1e-5 (relative)."""
    path, nuc = synth.config_snt(f"{snt_dir}/cr48.snt", "pf", 4, 4, synth.BASE_SEED + 2, (-8, 0), 1.0)
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    n = p.mdim
    v0 = np.random.default_rng(5).standard_normal(n)
    K = 16
    V = np.zeros((K + 1, n))
    V[0] = v0 / np.linalg.norm(v0)
    al = np.zeros(K)
    be = np.zeros(K)
    for it in range(K):
        w = h.operate_H(V[it], "fp64")
        hc = V[:it + 1] @ w
        al[it] = hc[it]
        w = w - hc @ V[:it + 1]
        w = w - (V[:it + 1] @ w) @ V[:it + 1]
        be[it] = np.linalg.norm(w)
        V[it + 1] = w / be[it]
    a64, b64 = h.lanczos_steps(K, v0, "fp64")
    a32, b32 = h.lanczos_steps(K, v0, "mixed")
    assert np.abs(a64 - al).max() < 1e-10 * np.abs(al).max()
    assert np.abs(b64 - be).max() < 1e-10 * np.abs(be).max()
    assert np.abs(a32 - al).max() < 1e-5 * np.abs(al).max()
    assert np.abs(b32 - be).max() < 1e-5 * np.abs(be).max()


def test_restart_small_problem_device_matches_host():
    """The device restart pipeline (parallel Ritz values, cluster inverse iteration,
    Lanczos on diag(theta), rotation) equals the host build, including clustered
    Ritz values like those of converged shell-model spectra."""
    from smgpu._lib import lib, ptr
    rng = np.random.default_rng(21)
    cases = []
    for m, ls in [(30, 10), (120, 20), (299, 20)]:
        a = rng.normal(size=m) * 10
        b = rng.uniform(0.1, 3, size=m)
        cases.append((a, b, ls))
    m = 80                                          # nearly degenerate lowest values
    a = np.concatenate([np.full(6, -86.0) + rng.normal(size=6) * 1e-4, rng.normal(size=m - 6) * 5])
    b = np.concatenate([np.full(10, 1e-6), rng.uniform(0.5, 2, size=m - 10)])
    cases.append((a, b, 12))
    for a, b, ls in cases:
        m = len(a)
        out = {}
        for dev in (-1, 0):
            S = np.zeros((ls, m)); ao = np.zeros(ls); bo = np.zeros(ls)
            assert lib().smgpu_selftest_restart(dev, ptr(a), ptr(b), m, ls, ptr(S), ptr(ao), ptr(bo)) == 0
            out[dev] = (S, ao, bo)
        T = np.diag(a) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
        S, ao, bo = out[0]
        St = S.T
        Tn = np.diag(ao) + np.diag(bo[:ls - 1], 1) + np.diag(bo[:ls - 1], -1)
        assert np.abs(St.T @ St - np.eye(ls)).max() < 1e-10
        assert np.abs(St.T @ T @ St - Tn).max() < 1e-9
        assert np.abs(np.linalg.eigvalsh(Tn) - np.linalg.eigvalsh(T)[:ls]).max() < 1e-9
        # the rotation is unique only up to invariant subspaces of diag(theta), so
        # host and device are compared through the spectrum of the new block
        Sh, ah, bh = out[-1]
        Th = np.diag(ah) + np.diag(bh[:ls - 1], 1) + np.diag(bh[:ls - 1], -1)
        assert np.abs(np.linalg.eigvalsh(Tn) - np.linalg.eigvalsh(Th)).max() < 1e-9
        assert abs(bo[-1] - bh[-1]) < 1e-9


def test_solve_is_deterministic(golden):
    p = smgpu.Problem(f"{golden}/usdb.snt", "Ne22")
    h = smgpu.Hamiltonian(p)
    r1 = h.trl(5, lm=40, ls=10, precision="mixed", seed=7)
    r2 = h.trl(5, lm=40, ls=10, precision="mixed", seed=7)
    assert np.array_equal(r1[0], r2[0]) and r1[2]["iterations"] == r2[2]["iterations"]
