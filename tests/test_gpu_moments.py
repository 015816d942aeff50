"""GPU: moments of main_sm's calc_moment (transit.jl:82-283, 702-765) through the
C ABI (smgpu_moments_*), against the oracle's restatement.

* The M1 and E2 operators applied to random vectors equal the oracle's
  op_obtr_diag over prep_obtr_diag's lists (1e-12 relative) on p-, sd- and
  pf-shell bases with both species active, odd and even A.
* eval_moment of the GPU's Ritz vectors equals the oracle's eval_moment of the
  same vectors (1e-10), with the J values from <J^2>.
* One valence nucleon (F17 / O17, usdb): main_sm(calc_moment=True) gives the
  Schmidt moments and the single-particle quadrupole moments (closed forms).
"""
import numpy as np
import pytest

import sm_oracle as orc
import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu

G = [1.0, 0.0, 5.586, -3.826]
EC = [1.5, 0.5]


def _cases(golden, snt_dir):
    pf, _ = synth.config_snt(f"{snt_dir}/mom_pf23.snt", "pf", 2, 3, 71, (-8, 0), 1.0)
    sd, _ = synth.config_snt(f"{snt_dir}/mom_sd33.snt", "sd", 3, 3, 72, (-6, 0), 1.2)
    return [(f"{golden}/ckpot.snt", "Li7"), (f"{golden}/ckpot.snt", "Be8"), (f"{golden}/usdb.snt", "Na23"),
            (f"{golden}/usdb.snt", "O20"), (sd, "Na22"), (pf, "V47")]


def test_operators_vs_oracle(golden, snt_dir):
    for path, nuc in _cases(golden, snt_dir):
        p = smgpu.Problem(path, nuc)
        o = orc.OracleHamiltonian(path, nuc)
        assert p.mdim == o.mdim
        m = smgpu.Moments(p, 0, G, EC)
        b = orc.bpar_main_sm(o.sm["A"])
        assert abs(m.bpar - b) < 1e-12 * b
        ops = orc.prep_obtr_diag(o.sm["p_sps"], o.sm["n_sps"], o.msps_p, o.msps_n, o.pbits, o.nbits, b, G, EC)
        V = np.random.default_rng(7).standard_normal((2, p.mdim))
        for which in ("M1", "E2"):
            W = m.apply(V, which)
            for k in range(2):
                ref = orc.op_obtr_diag(ops[which][0], ops[which][1], o.pbits, o.nbits, o.tdims, V[k])
                err = np.abs(W[k] - ref).max() / max(1e-300, np.abs(ref).max())
                print(nuc, which, "mdim", p.mdim, "rel err", err, flush=True)
                assert err < 1e-12


def test_eval_moment_vs_oracle(golden, snt_dir):
    for path, nuc in _cases(golden, snt_dir)[:5]:
        o = orc.OracleHamiltonian(path, nuc)
        p = smgpu.Problem(path, nuc)
        h = smgpu.Hamiltonian(p)
        nev = min(6, p.mdim)
        en, R, _ = h.trl(nev, want_vectors=True)
        Js = np.array([smgpu.shellmodel.j_from_jj1(x) for x in smgpu.shellmodel.expect_j2(p, R, 0)])
        m = smgpu.Moments(p, 0, G, EC)
        mu, Q = m.eval(R, Js)
        ops = orc.prep_obtr_diag(o.sm["p_sps"], o.sm["n_sps"], o.msps_p, o.msps_n, o.pbits, o.nbits,
                                 orc.bpar_main_sm(o.sm["A"]), G, EC)
        res, tx = orc.eval_moment(o.Mtot, R, Js, ops, o.pbits, o.nbits, o.tdims)
        print(nuc, "J", Js, "mu", mu, "Q", Q, flush=True)
        for k, (rm, rq) in enumerate(res):
            assert (rm is None) == np.isnan(mu[k]) and (rq is None) == np.isnan(Q[k])
            if rm is not None:
                assert abs(mu[k] - rm) < 1e-10
            if rq is not None:
                assert abs(Q[k] - rq) < 1e-10
        import re
        pat = r"(mu|Q) \(n=(\d+)\)"
        assert re.findall(pat, smgpu.format_moments(mu, Q)) == re.findall(pat, tx)


@pytest.mark.parametrize("nuc,sp", [("F17", 0), ("O17", 1)])
def test_main_sm_single_nucleon(golden, nuc, sp, capsys):
    en, J, (mu, Q) = smgpu.main_sm(f"{golden}/usdb.snt", nuc, 3, return_J=True, return_moments=True,
                                   calc_moment=True, verbose=True)
    out = capsys.readouterr().out
    assert "mu (n=1)" in out and "Q (n=1)" in out
    b = orc.bpar_main_sm(17)
    gl, gs = (G[0], G[2]) if sp == 0 else (G[1], G[3])
    # usdb orbits: 0d3/2, 0d5/2, 1s1/2 by SPE order d5/2 < s1/2 < d3/2
    for k, (l, j2, n) in enumerate([(2, 5, 0), (0, 1, 1), (2, 3, 0)]):
        assert abs(J[k] - j2 / 2) < 1e-9
        j = j2 / 2
        schmidt = gl * l + gs / 2 if j2 == 2 * l + 1 else j / (j + 1) * (gl * (l + 1) - gs / 2)
        assert abs(mu[k] - schmidt) < 1e-9, (k, mu[k], schmidt)
        q_sp = -(2 * j - 1) / (2 * j + 2) * b * b * (2 * n + l + 1.5) * EC[sp]
        if j2 == 1:
            assert np.isnan(Q[k]) or abs(Q[k]) < 1e-9
        else:
            assert abs(Q[k] - q_sp) < 1e-9, (k, Q[k], q_sp)
