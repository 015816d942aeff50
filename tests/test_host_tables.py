"""CPU checks of the product's host side against the oracle (no GPU needed).

* every symbol declared in include/smgpu.h is exported by libsmgpu.so
* basis (pbits, nbits, tdims, Mp) bit-exact vs the oracle's `occ`
* one-body hop tables bit-exact vs the oracle's `calc_1b_jumps!` lists
* pp / nn symmetric matrices vs the oracle's `prep_2bjumps` entries
* the shared tridiagonal eigen-solver vs numpy
"""
import os
import re

import numpy as np
import pytest

import sm_oracle as orc
import smgpu
from smgpu import synth
from smgpu._lib import lib, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    text = open(os.path.join(ROOT, "include", "smgpu.h")).read()
    names = set(re.findall(r"\b(smgpu_[a-z0-9_]+)\s*\(", text))
    assert len(names) >= 20
    L = lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def _cases(golden, snt_dir):
    c1, _ = synth.config_snt(f"{snt_dir}/mg24.snt", "sd", 4, 4, synth.BASE_SEED + 1, (-4, 0), 1.5)
    pfs, _ = synth.config_snt(f"{snt_dir}/pf22.snt", "pf", 2, 3, 99, (-8, 0), 1.0)
    return [(f"{golden}/ckpot.snt", "Be8"), (f"{golden}/ckpot.snt", "Li7"), (f"{golden}/usdb.snt", "O20"),
            (f"{golden}/usdb.snt", "Na23"), (f"{golden}/random_snts/tmp_0.snt", "Ne20"), (c1, "Mg24"),
            (pfs, "V48")]


def test_basis_bit_exact(golden, snt_dir):
    for path, nuc in _cases(golden, snt_dir):
        p = smgpu.Problem(path, nuc)
        o = orc.OracleHamiltonian(path, nuc)
        b = p.basis()
        assert list(b["tdims"]) == o.tdims, nuc
        assert list(b["Mp"]) == o.Mps and list(b["Mn"]) == o.Mns
        for x, y in zip(b["pbits"], o.pbits):
            assert np.array_equal(x, np.array(y, dtype=np.int64))
        for x, y in zip(b["nbits"], o.nbits):
            assert np.array_equal(x, np.array(y, dtype=np.int64))


def _bit_to_index(bits, L):
    return (L - 1 - np.log2(bits.astype(np.float64)).round().astype(np.int64))


def test_hop_tables_bit_exact(golden, snt_dir):
    for path, nuc in _cases(golden, snt_dir):
        p = smgpu.Problem(path, nuc)
        o = orc.OracleHamiltonian(path, nuc)
        h = p.hops()
        ops = np.concatenate([x.reshape(-1, 4) for x in o.pn_ops])
        for s in (0, 1):
            L = len(o.msps_p) if s == 0 else len(o.msps_n)
            bi, bf, op, i, f, ph = o.export_pn(s)
            a = _bit_to_index(ops[op, 0 if s == 0 else 1], L)
            c = _bit_to_index(ops[op, 2 if s == 0 else 3], L)
            ref = np.unique(np.stack([bi, bf, i, f, a, c, ph]).T, axis=0)
            m = (h["species"] == s) & (h["bf"] >= h["bi"])
            mine = np.stack([h["bi"][m], h["bf"][m], h["i"][m], h["f"][m], h["a"][m], h["c"][m],
                             h["phase"][m]]).T.astype(np.int64)
            # restrict to the (bi, bf, a, c) operators the reference kept (V != 0)
            keys_ref = {tuple(r) for r in np.unique(ref[:, [0, 1, 4, 5]], axis=0)}
            sel = np.array([tuple(r) in keys_ref for r in mine[:, [0, 1, 4, 5]]], dtype=bool)
            mine = np.unique(mine[sel], axis=0)
            assert mine.shape == ref.shape, (nuc, s, mine.shape, ref.shape)
            assert np.array_equal(mine, ref), (nuc, s)


def test_pair_matrices(golden, snt_dir):
    for path, nuc in _cases(golden, snt_dir):
        p = smgpu.Problem(path, nuc)
        o = orc.OracleHamiltonian(path, nuc)
        for s in (0, 1):
            blk, i, f, cf = o.export_pp(s)
            ref = {}
            for b_, i_, f_, c_ in zip(blk, i, f, cf):
                ref[(b_, f_, i_)] = ref.get((b_, f_, i_), 0.0) + c_
                ref[(b_, i_, f_)] = ref.get((b_, i_, f_), 0.0) + c_
            b2, f2, i2, v2 = p.pair_matrix(s)
            mine = {(int(x), int(y), int(z)): float(w) for x, y, z, w in zip(b2, f2, i2, v2)}
            # entries whose terms cancel may differ by a rounding residue (~1e-17)
            # with the summation order; they count as zero
            nz_ref = {k: v for k, v in ref.items() if abs(v) > 1e-12}
            nz_mine = {k: v for k, v in mine.items() if abs(v) > 1e-12}
            assert set(nz_ref) == set(nz_mine), nuc
            for k, v in nz_ref.items():
                assert abs(nz_mine[k] - v) <= 1e-13 * max(1.0, abs(v)), (nuc, k)


def test_problem_errors(golden, tmp_path):
    with pytest.raises(smgpu.SmgpuError) as e:
        smgpu.Problem(f"{golden}/nonexistent.snt", "Be8")
    assert e.value.code == -2
    bad = tmp_path / "bad.snt"
    bad.write_text("  2  2  2  2\n  1  0  1  1  -1\n")
    with pytest.raises(smgpu.SmgpuError) as e:
        smgpu.Problem(str(bad), "Be8")
    assert e.value.code == -1
    with pytest.raises(smgpu.SmgpuError):
        smgpu.Problem(f"{golden}/ckpot.snt", "Xx8")
    with pytest.raises(smgpu.SmgpuError):           # 2M parity must match A
        smgpu.Problem(f"{golden}/ckpot.snt", "Be8", twoM=1)
    p = smgpu.Problem(f"{golden}/ckpot.snt", "He4")  # no valence particles: mdim 1
    assert p.mdim == 1


def test_tridiag_solver():
    rng = np.random.default_rng(5)
    for n, k in [(1, 1), (2, 2), (7, 5), (60, 20), (299, 20)]:
        a = rng.normal(size=n) * 10
        b = rng.uniform(0.01, 3.0, size=max(n - 1, 1))
        vals = np.zeros(k)
        vecs = np.zeros((k, n))
        assert lib().smgpu_selftest_tridiag(ptr(a), ptr(b), n, k, ptr(vals), ptr(vecs)) == 0
        T = np.diag(a) + (np.diag(b[:n - 1], 1) + np.diag(b[:n - 1], -1) if n > 1 else 0)
        ev, evec = np.linalg.eigh(T)
        assert np.allclose(vals, ev[:k], atol=1e-12 * max(1, abs(ev).max()))
        for j in range(k):
            assert abs(abs(vecs[j] @ evec[:, j]) - 1.0) < 1e-10
    # clustered eigenvalues (converged Ritz values): orthogonality must hold
    n = 80
    a = np.concatenate([np.full(10, -30.0), np.linspace(-20, 20, n - 10)])
    b = np.full(n - 1, 1e-7)
    vals = np.zeros(12); vecs = np.zeros((12, n))
    assert lib().smgpu_selftest_tridiag(ptr(a), ptr(b), n, 12, ptr(vals), ptr(vecs)) == 0
    assert np.abs(vecs @ vecs.T - np.eye(12)).max() < 1e-10


def _block_lanczos_like(rng, q, nb, lsv=0):
    """Block tridiagonal (q x q blocks, upper-triangular couplings) with an
    optional thick-restart arrowhead: diag(theta[lsv]) coupled to the next block."""
    n = q * nb
    T = np.zeros((n, n))
    for b in range(nb):
        A = rng.normal(size=(q, q)) * 5
        T[b * q:(b + 1) * q, b * q:(b + 1) * q] = 0.5 * (A + A.T)
        if b + 1 < nb:
            B = np.triu(rng.normal(size=(q, q)))
            T[(b + 1) * q:(b + 2) * q, b * q:(b + 1) * q] = B
            T[b * q:(b + 1) * q, (b + 1) * q:(b + 2) * q] = B.T
    if lsv:
        T[:lsv, :] = 0; T[:, :lsv] = 0
        T[np.arange(lsv), np.arange(lsv)] = np.sort(rng.normal(size=lsv) * 10)
        r = rng.normal(size=(q, lsv))
        T[lsv:lsv + q, :lsv] = r; T[:lsv, lsv:lsv + q] = r.T
    return T


def test_dense_symeig_host():
    """Host eigen-solver of the block Lanczos matrix (Householder + tridiagonal
    bisection / inverse iteration) against LAPACK (numpy.linalg.eigh): block
    tridiagonal and restarted (arrowhead) shapes, a dense random matrix, an
    already-tridiagonal one, a 1x1, and a matrix with a repeated eigenvalue."""
    rng = np.random.default_rng(17)
    cases = [_block_lanczos_like(rng, 2, 60), _block_lanczos_like(rng, 3, 40, lsv=21),
             _block_lanczos_like(rng, 4, 12, lsv=8), np.array([[3.5]])]
    A = rng.normal(size=(70, 70)); cases.append(A + A.T)
    cases.append(np.diag(rng.normal(size=50)) + np.diag(np.full(49, 0.7), 1) + np.diag(np.full(49, 0.7), -1))
    Q, _ = np.linalg.qr(rng.normal(size=(30, 30)))
    d = np.concatenate([[-5.0, -5.0, -5.0], np.linspace(-4, 4, 27)])
    cases.append(Q @ np.diag(d) @ Q.T)
    for T in cases:
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.shape[0]
        k = min(n, 12)
        ev = np.linalg.eigvalsh(T)
        scale = max(1.0, np.abs(ev).max())
        vals = np.zeros(k)
        assert lib().smgpu_selftest_symeig(ptr(T), n, k, ptr(vals), None) == 0
        assert np.abs(vals - ev[:k]).max() < 1e-12 * scale * max(1, n / 10)
        vecs = np.zeros((k, n))
        assert lib().smgpu_selftest_symeig(ptr(T), n, k, ptr(vals), ptr(vecs)) == 0
        assert np.abs(vecs @ vecs.T - np.eye(k)).max() < 1e-10
        res = T @ vecs.T - vecs.T * vals
        assert np.abs(res).max() < 1e-10 * scale


def test_restart_small_problem_host():
    """tri_restart: S orthonormal, S^T T S tridiagonal with the returned entries,
    coupling only on the last kept vector, eigenvalues preserved."""
    rng = np.random.default_rng(9)
    for m, ls in [(10, 4), (60, 20), (299, 20)]:
        a = rng.normal(size=m) * 10
        b = rng.uniform(0.1, 3, size=m)
        T = np.diag(a) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
        S = np.zeros((ls, m)); ao = np.zeros(ls); bo = np.zeros(ls)
        assert lib().smgpu_selftest_restart(-1, ptr(a), ptr(b), m, ls, ptr(S), ptr(ao), ptr(bo)) == 0
        St = S.T
        Tn = np.diag(ao) + np.diag(bo[:ls - 1], 1) + np.diag(bo[:ls - 1], -1)
        assert np.abs(St.T @ St - np.eye(ls)).max() < 1e-12
        assert np.abs(St.T @ T @ St - Tn).max() < 1e-10
        coup = b[m - 1] * St[m - 1, :]
        assert np.abs(coup[:-1]).max() < 1e-10 and abs(abs(coup[-1]) - bo[-1]) < 1e-10
        assert np.abs(np.linalg.eigvalsh(Tn) - np.linalg.eigvalsh(T)[:ls]).max() < 1e-10


def test_main_sm_mdimmode(golden, capsys):
    """mdimmode: main_sm prints its dimension line and stops after the basis (host
    only).  usdb Mg24 gives the line of the reference's example log
    (example/log_sample_script.txt:323, up to the J list that follows it there)."""
    out = smgpu.main_sm(f"{golden}/usdb.snt", "Mg24", 10, mdimmode=True)
    assert out is None
    line = capsys.readouterr().out
    assert line.startswith("  Mg24    Z,N=(12,12)     c(8,8)     v(4,4)  mdim:        28503 (  4.45 ) "), repr(line)
