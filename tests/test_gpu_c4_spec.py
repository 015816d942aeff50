"""A seeded subset of the 512-request serving stream at its full spec, on the GPU.

The stream is `synth.serving_stream` with its defaults (sd/pf, 2..6 valence
protons and neutrons, k 1..10, 2M = A % 2, per-request interactions), the one
the benchmark serves.  The subset takes its first pf 6p6n request (mdim
109,954,620), its first pf 5p6n/6p5n request (mdim 44,993,824) and its first six
requests on sd bases or pf bases of at most 4 valence nucleons per species, and serves them through one server in 4 slots with
lockstep batches of two.  Each request is one main_sm eigenvalue solve
(shellmodel_main.jl:89-281).

* alone vs in the mix: every request gives the same eigenvalues served with
  the others (all submitted at once) as served alone (tolerance 1e-9 MeV);
* mixed precision vs plain FP64 (every vector FP64, full reorthogonalisation,
  one request at a time): within 1e-5 MeV;
* the two large bases are admitted on both sides (the FP64 solve of pf 6p6n
  is admitted with its Lanczos basis bounded by the HBM budget).

At the stream's largest basis (pf 6p6n, mdim 109,954,620, too large for the
oracle) H·v is checked through size-independent properties: <u, H v> = <H u, v>
and linearity in FP64, and the FP32 H·v within 1e-5 of the FP64 one.
"""
import numpy as np
import pytest

from smgpu import synth
from smgpu.server import Server

pytestmark = pytest.mark.gpu

TOL_MIX = 1e-9
TOL_FP64 = 1e-5


def _subset(directory):
    reqs = synth.serving_stream(directory, 512, write=False)
    pick, kinds = [], set()
    for r in reqs:
        key = (r.space, r.zv, r.nv)
        if key == ("pf", 6, 6) and "66" not in kinds:
            kinds.add("66"); pick.append(r)
        elif key in (("pf", 5, 6), ("pf", 6, 5)) and "56" not in kinds:
            kinds.add("56"); pick.append(r)
        elif r.space == "sd" or max(r.zv, r.nv) <= 4:
            if sum(1 for q in pick if q.space == "sd" or max(q.zv, q.nv) <= 4) < 6:
                pick.append(r)
        if len(kinds) == 2 and len(pick) == 8:
            break
    assert kinds == {"66", "56"} and len(pick) == 8
    for r in pick:
        synth.config_snt(r.snt, r.space, r.zv, r.nv, r.seed, (-6.0, 0.0), 1.2)
    return pick


def _one_at_a_time(srv, reqs, precision):
    out = []
    for r in reqs:
        rid = srv.submit(r.snt, r.nucleus, r.num_ev, precision=precision, seed=r.seed % (1 << 31))
        assert srv.wait(rid, 300000), (r.nucleus, precision)
        out.append(srv.result(rid))
        print(precision, "alone", r.index, r.nucleus, out[-1][1]["mdim"], r.num_ev, out[-1][1]["solve_ms"],
              flush=True)
    return out


def test_c4_spec_subset_alone_vs_mix_vs_fp64(tmp_path):
    reqs = _subset(str(tmp_path))
    with Server(max_inflight=4, max_batch=2, host_threads=4, keep_bytes=64 << 30) as srv:
        alone = _one_at_a_time(srv, reqs, "mixed")
        ids = [srv.submit(r.snt, r.nucleus, r.num_ev, precision="mixed", seed=r.seed % (1 << 31)) for r in reqs]
        for rid in ids:
            assert srv.wait(rid, 300000)
        mix = [srv.result(rid) for rid in ids]
        st = srv.stats()
    assert st["failed"] == 0 and st["peak_running"] >= 2, st
    assert st["hbm_peak"] <= st["hbm_budget"], st
    with Server(max_inflight=1, max_batch=1, host_threads=4) as seq:
        fp64 = _one_at_a_time(seq, reqs, "fp64")
    for r, (ea, ia), (em, im), (ef, iff) in zip(reqs, alone, mix, fp64):
        print(r.index, r.nucleus, im["mdim"], r.num_ev, "mix", im["latency_ms"], "fp64 lm", iff["lm"], flush=True)
        assert ia["state"] == im["state"] == iff["state"] == "done"
        assert len(ea) == len(em) == len(ef) == min(r.num_ev, im["mdim"])
        assert np.abs(em - ea).max() <= TOL_MIX, (r.nucleus, em - ea)
        assert np.abs(em - ef).max() <= TOL_FP64, (r.nucleus, em - ef)
        assert im["n_fallback"] == 0
    big = [im["mdim"] for _, (_, im) in zip(reqs, mix)]
    assert 109954620 in big and 44993824 in big


def test_hv_properties_fe52(tmp_path):
    path, nuc = synth.config_snt(str(tmp_path / "fe52.snt"), "pf", 6, 6, 7, (-6.0, 0.0), 1.2)
    from smgpu import Hamiltonian, Problem
    p = Problem(path, nuc)
    assert p.mdim == 109954620
    h = Hamiltonian(p)
    rng = np.random.default_rng(52)
    v = rng.standard_normal(p.mdim)
    u = rng.standard_normal(p.mdim)
    w64 = h.operate_H(v, "fp64")
    hu = h.operate_H(u, "fp64")
    scale = np.abs(w64).max()
    sym = abs(u @ w64 - hu @ v) / (np.abs(u) @ np.abs(w64))
    hsum = h.operate_H(u - 0.5 * v, "fp64")
    lin = np.abs(hsum - (hu - 0.5 * w64)).max() / scale
    w32 = h.operate_H(v, "mixed")
    e32 = np.abs(w32 - w64).max() / scale
    print("fe52 symmetry", sym, "linearity", lin, "fp32 vs fp64", e32, flush=True)
    assert sym <= 1e-11 and lin <= 1e-11 and e32 <= 1e-5
