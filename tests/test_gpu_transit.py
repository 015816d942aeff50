"""GPU: transit_main's M1 and E2 transitions (transit.jl:286-534, 561-700) through
the C ABI (smgpu_transitions_*, smgpu_transit), against the oracle's restatement.

* The operators (component mu = (2M_l - 2M_r) / 2 = 0, 1, 2) applied to random
  right-basis vectors equal the oracle's op_obtr_nondiag over prep_obtr_nondiag's
  lists (1e-12 relative), on sd and pf bases with both species active, odd and
  even A.
* transit_main end to end on Ne20 (usdb): J-projected Ritz vectors of the 2+ (2M = 4)
  and 0+ (2M = 0) states written by main_sm(save_wav=True); every reduced matrix
  element and B(E2) equals the oracle's transit_main on the same vectors (1e-10).
* One file given twice (the 2+ states): <gm> and <eQ> equal eval_moment's.
"""
import numpy as np
import pytest

import sm_oracle as orc
import smgpu
from smgpu import synth
from smgpu.wav import read_wav

pytestmark = pytest.mark.gpu

G = [1.0, 0.0, 5.586, -3.826]
EC = [1.5, 0.5]


def _cases(golden, snt_dir):
    pf, _ = synth.config_snt(f"{snt_dir}/tr_pf23.snt", "pf", 2, 3, 81, (-8, 0), 1.0)
    sd, _ = synth.config_snt(f"{snt_dir}/tr_sd33.snt", "sd", 3, 3, 82, (-6, 0), 1.2)
    # (interaction, nucleus, 2M_left, 2M_right)
    return [(f"{golden}/usdb.snt", "Ne20", 2, 2), (f"{golden}/usdb.snt", "Ne20", 4, 2),
            (f"{golden}/usdb.snt", "Ne20", 4, 0), (f"{golden}/usdb.snt", "Na23", 5, 3),
            (sd, "Na22", 6, 2), (pf, "V47", 3, 1), (pf, "V47", 5, 1)]


def test_operators_vs_oracle(golden, snt_dir):
    for path, nuc, ml2, mr2 in _cases(golden, snt_dir):
        left = smgpu.Problem(path, nuc, twoM=ml2)
        right = smgpu.Problem(path, nuc, twoM=mr2)
        t = smgpu.Transitions(left, right, 0, G, EC)
        T = orc.transitions(path, nuc, ml2, mr2, G, EC)
        assert abs(t.bpar - T["bpar"]) < 1e-12 * T["bpar"]
        assert t.mu2 == ml2 - mr2 and t.mdim_l == T["tdims_l"][-1] and t.mdim_r == T["tdims_r"][-1]
        V = np.random.default_rng(9).standard_normal((2, t.mdim_r))
        for which in ("M1", "E2"):
            W = t.apply(V, which)
            o = T["ops"][which]
            for k in range(2):
                ref = orc.op_obtr_nondiag(o[0], o[1], T["pbits_l"], T["nbits_l"], T["tdims_l"], T["pbits_r"],
                                          T["nbits_r"], T["tdims_r"], V[k])
                scale = max(1e-300, np.abs(ref).max())
                err = np.abs(W[k] - ref).max() / scale
                print(nuc, ml2, mr2, which, "mdim", t.mdim_l, t.mdim_r, "rel err", err, flush=True)
                if which == "M1" and ml2 - mr2 > 2:
                    assert np.abs(W[k]).max() == 0.0
                else:
                    assert np.abs(ref).max() > 0 and err < 1e-12


def test_transit_main_ne20(golden, tmp_path):
    path = f"{golden}/usdb.snt"
    w2 = str(tmp_path / "ne20_j2.wav")
    w0 = str(tmp_path / "ne20_j0.wav")
    smgpu.main_sm(path, "Ne20", 3, [4], save_wav=True, wav_path=w2)
    smgpu.main_sm(path, "Ne20", 2, [0], save_wav=True, wav_path=w0)
    recs = smgpu.transit_main(path, "Ne20", 4, 0, [w2, w0])
    T = orc.transitions(path, "Ne20", 4, 0, G, EC)
    _, jj2, V2, m2 = read_wav(w2, T["tdims_l"][-1])
    _, jj0, V0, m0 = read_wav(w0, T["tdims_r"][-1])
    assert m2 == 4 and m0 == 0 and list(jj2) == [4, 4, 4] and list(jj0) == [0, 0]
    ref = orc.transit(T, V2, V0)
    assert [r["mode"] for r in recs] == ["E2"] * 6 and [x[0] for x in ref] == ["E2"] * 6
    for r, x in zip(recs, ref):
        print(r, flush=True)
        assert (r["l"], r["r"]) == (x[1], x[2])
        assert abs(r["Mred"] - x[3]) <= 1e-10 * max(1.0, abs(x[3]))
        assert abs(r["B_l_from_r"] - x[4]) <= 1e-10 * max(1.0, x[4])
    assert max(r["B_l_from_r"] for r in recs) > 1.0        # a collective 2+ -> 0+ strength in e^2 fm^4
    # one file given twice: <gm> and <eQ> of the 2+ states equal eval_moment's
    same = smgpu.transit_main(path, "Ne20", 4, 4, [w2, w2], verbose=False)
    mom = {(r["mode"], r["l"]): r["moment"] for r in same if r["moment"] is not None}
    mu, Q = smgpu.Moments(smgpu.Problem(path, "Ne20", twoM=4), 0, G, EC).eval(V2, [2.0] * 3)
    for k in range(3):
        assert abs(mom[("M1", k + 1)] - mu[k]) < 1e-10 and abs(mom[("E2", k + 1)] - Q[k]) < 1e-10
