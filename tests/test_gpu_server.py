"""Serving loop on the GPU (smgpu_server_* through the C ABI).

* alone vs in the mix: every request of a mixed stream (sd and pf spaces,
  different particle numbers, k and interactions, several sharing a basis) gives
  the same eigenvalues when served concurrently as when solved alone through
  smgpu_trl_solve with the same parameters and seed (tolerance 1e-9 MeV; the
  kernels are deterministic, so in practice the values are identical), and the
  mixed-precision values agree with a plain FP64 solve to 1e-5 MeV;
* admission by HBM footprint: with a budget that holds one request's
  workspace only, the server never runs two at once and never exceeds the
  budget; a request larger than the budget fails with ENOMEM;
* cancel of a queued and of a running request; errors (missing file) are
  reported per request; the server keeps serving afterwards.
* the large lane: with large_mdim and reserve_cus, the requests at or above
  large_mdim run in slot 0 only (streams on a CU partition) and the others never
  there, each with the eigenvalues of the request solved alone; invalid
  combinations fail at create.
"""
import time

import numpy as np
import pytest

import smgpu
from smgpu import synth
from smgpu._lib import SMGPU_ECANCELED
from smgpu.server import Server

pytestmark = pytest.mark.gpu

TOL_MIX = 1e-9       # alone vs in the mix, same precision and seed
TOL_FP64 = 1e-5      # mixed precision vs FP64 (MeV)


def _stream(snt_dir):
    """12 requests: sd and pf, 2..4 valence particles per species, k 1..10;
    two pairs share a basis (same orbits, Z_v, N_v) with different interactions."""
    spec = [("sd", 2, 2, 3), ("pf", 2, 2, 5), ("sd", 3, 2, 4), ("pf", 2, 3, 2), ("sd", 4, 4, 5),
            ("sd", 2, 4, 10), ("pf", 3, 3, 6), ("sd", 4, 4, 1), ("sd", 3, 3, 7), ("pf", 2, 2, 9),
            ("sd", 2, 3, 8), ("pf", 3, 2, 3)]
    out = []
    for i, (space, zv, nv, k) in enumerate(spec):
        path, nuc = synth.config_snt(f"{snt_dir}/srv_{i}.snt", space, zv, nv, 7000 + i, (-6, 0), 1.2)
        out.append((path, nuc, k, 100 + i))
    return out


def _alone(path, nuc, k, seed, precision):
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    ev, _, _ = h.trl(min(k, p.mdim), precision=precision, seed=seed)
    return ev


def test_server_alone_vs_mix(snt_dir):
    reqs = _stream(snt_dir)
    alone = [_alone(p, n, k, s, "mixed") for p, n, k, s in reqs]
    fp64 = [_alone(p, n, k, s, "fp64") for p, n, k, s in reqs]
    with Server(max_inflight=4) as srv:
        ids = [srv.submit(p, n, k, precision="mixed", seed=s) for p, n, k, s in reqs]
        for rid in ids:
            assert srv.wait(rid, 120000)
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_running"] >= 2, st            # the mix really overlapped
    assert st["hbm_peak"] <= st["hbm_budget"]
    for (path, nuc, k, s), a, f, (ev, info) in zip(reqs, alone, fp64, res):
        print(nuc, info["mdim"], k, info["latency_ms"], flush=True)
        assert info["state"] == "done" and len(ev) == len(a)
        assert np.abs(ev - a).max() <= TOL_MIX, (nuc, ev - a)
        assert np.abs(ev - f).max() <= TOL_FP64, (nuc, ev - f)
        assert info["latency_ms"] >= info["solve_ms"] > 0


def test_server_shared_basis_and_budget(snt_dir):
    """Same basis, different interactions: the index is uploaded once; a budget
    of one request's footprint plus half its private part admits one at a time; a
    request whose Krylov basis does not fit the budget runs with a smaller lm."""
    paths = [synth.config_snt(f"{snt_dir}/bud_{i}.snt", "sd", 3, 3, 8000 + i, (-6, 0), 1.2)[0] for i in range(4)]
    nuc = synth.nucleus_name("sd", 3, 3)
    with Server(max_inflight=1) as srv:
        rid = srv.submit(paths[0], nuc, 4)
        assert srv.wait(rid, 60000)
        ev0, info0 = srv.result(rid)
    F, idx = info0["footprint"], info0["index_bytes"]
    assert 0 < idx < F
    budget = F + (F - idx) // 2
    with Server(max_inflight=3, hbm_budget=budget) as srv:
        ids = [srv.submit(p, nuc, 4) for p in paths]
        for rid in ids:
            assert srv.wait(rid, 60000)
        infos = [srv.result(rid)[1] for rid in ids]
        st = srv.stats()
    assert st["peak_running"] == 1, st
    assert st["hbm_peak"] <= budget
    assert st["index_uploads"] == 1 and st["index_misses"] == 1 and st["index_hits"] == 3, st
    assert sum(i["index_shared"] for i in infos) == 3
    # a budget below one request's footprint at lm 300: the request runs with the
    # largest Krylov basis that fits (thick restarts) and reaches the same levels
    with Server(max_inflight=2, hbm_budget=F // 2) as srv:
        rid = srv.submit(paths[0], nuc, 4)
        assert srv.wait(rid, 60000)
        ev, info = srv.result(rid)
        assert info["state"] == "done" and 40 <= info["lm"] < 300 and info["footprint"] <= F // 2, info
        assert np.abs(ev - ev0).max() <= TOL_FP64
    # a budget that cannot hold even the index tables: ENOMEM for that request only
    with Server(max_inflight=2, hbm_budget=idx // 2) as srv:
        rid = srv.submit(paths[0], nuc, 4)
        assert srv.wait(rid, 60000)
        _, info = srv.result(rid, raise_on_error=False)
        assert info["state"] == "failed" and info["status"] == -4


def test_server_cancel_and_errors(snt_dir, golden):
    path, nuc = synth.config_snt(f"{snt_dir}/cancel.snt", "sd", 4, 4, 9001, (-6, 0), 1.2)
    with Server(max_inflight=1) as srv:
        # tol = -1 never converges (the Ritz values never rise by 1 MeV): it
        # restarts until canceled (the flag is polled once per iteration)
        long_id = srv.submit(path, nuc, 5, tol=-1.0, ls=10, lm=60, max_iter=20000)
        t0 = time.time()
        while srv.poll(long_id) != "running":
            assert time.time() - t0 < 60
            time.sleep(0.005)
        queued_id = srv.submit(path, nuc, 5)      # the only slot is busy: it waits
        srv.cancel(queued_id)
        srv.cancel(long_id)
        assert srv.wait(long_id, 60000) and srv.wait(queued_id, 60000)
        for rid in (long_id, queued_id):
            _, info = srv.result(rid, raise_on_error=False)
            print(rid, info, flush=True)
            assert info["state"] == "canceled" and info["status"] == SMGPU_ECANCELED, info
        bad = srv.submit(f"{golden}/does_not_exist.snt", "Be8", 2)
        assert srv.wait(bad, 60000)
        _, info = srv.result(bad, raise_on_error=False)
        assert info["state"] == "failed" and info["status"] == -2
        # still serving
        ok = srv.submit(f"{golden}/ckpot.snt", "Be8", 5, precision="fp64")
        assert srv.wait(ok, 60000)
        ev, info = srv.result(ok)
        np.testing.assert_allclose(ev, [-31.119, -27.300, -19.162, -18.249, -16.722], atol=1e-3)
        with pytest.raises(smgpu.SmgpuError):
            srv.submit(path, nuc, 0)


def test_server_interaction_variants(snt_dir):
    """Interaction variants on one basis (base + N(0, 0.05) per TBME): served
    concurrently with one shared index upload; each variant's lowest three
    eigenvalues agree with its FP64 solve to 1e-5 MeV."""
    paths, nuc = synth.variant_snts(f"{snt_dir}/variants", 6, zv=3, nv=3)
    with Server(max_inflight=3) as srv:
        ids = [srv.submit(p, nuc, 3, seed=50 + i) for i, p in enumerate(paths)]
        for rid in ids:
            assert srv.wait(rid, 120000)
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    assert st["index_uploads"] == 1 and st["index_misses"] == 1
    prob = smgpu.Problem(paths[0], nuc)
    ham = smgpu.Hamiltonian(prob)
    grounds = []
    for i, p in enumerate(paths):
        prob.set_interaction(p)
        ham.update_interaction(prob)
        ev64, _, _ = ham.trl(3, precision="fp64", seed=50 + i)
        assert np.abs(res[i][0] - ev64).max() <= TOL_FP64, (i, res[i][0] - ev64)
        grounds.append(ev64[0])
    assert len(set(np.round(grounds, 6))) == len(paths)       # the variants differ


def test_server_batches_identical_hamiltonians(snt_dir):
    """max_batch: ready requests with the same basis and identical coefficients (and
    precision) run in one slot in lockstep with shared H·v launches.  Every result
    equals the request solved alone (1e-9 MeV; the batched solve is bit-identical),
    requests on other interactions or bases are served beside them, a request
    canceled while its batch runs ends canceled and its mates still finish."""
    a, nuc_a = synth.config_snt(f"{snt_dir}/bat_a.snt", "sd", 4, 4, 8101, (-6, 0), 1.2)
    b, nuc_b = synth.config_snt(f"{snt_dir}/bat_b.snt", "sd", 4, 4, 8102, (-6, 0), 1.2)   # same basis
    c, nuc_c = synth.config_snt(f"{snt_dir}/bat_c.snt", "pf", 2, 3, 8103, (-6, 0), 1.2)
    reqs = [(a, nuc_a, 5, 300 + i) for i in range(6)] + [(b, nuc_b, 3, 310), (b, nuc_b, 3, 311), (c, nuc_c, 4, 312)]
    alone = [_alone(p, n, k, s, "mixed") for p, n, k, s in reqs]
    with Server(max_inflight=2, max_batch=3) as srv:
        ids = [srv.submit(p, n, k, precision="mixed", seed=s) for p, n, k, s in reqs]
        for rid in ids:
            assert srv.wait(rid, 120000)
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    for (ev, info), ref in zip(res, alone):
        assert info["state"] == "done"
        np.testing.assert_allclose(ev, ref, atol=TOL_MIX)
    assert st["batched"] >= 3 and 2 <= st["peak_batch"] <= 3, st
    # cancel one member of a running batch: the others complete
    with Server(max_inflight=1, max_batch=3) as srv:
        ids = [srv.submit(a, nuc_a, 5, precision="mixed", seed=400 + i) for i in range(3)]
        t0 = time.time()
        while srv.poll(ids[1]) not in ("running", "done", "canceled", "failed") and time.time() - t0 < 60:
            time.sleep(0.002)
        srv.cancel(ids[1])
        for rid in ids:
            assert srv.wait(rid, 120000)
        out = [srv.result(rid, raise_on_error=False) for rid in ids]
        assert out[1][1]["state"] in ("canceled", "done")
        for i in (0, 2):
            assert out[i][1]["state"] == "done"
            np.testing.assert_allclose(out[i][0], _alone(a, nuc_a, 5, 400 + i, "mixed"), atol=TOL_MIX)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_server_batches_interaction_variants(snt_dir, precision):
    """batch_variants: ready requests on one basis and precision with different
    interactions share a slot's lockstep batch, each member with its own
    coefficients on the shared index (admitted together and joining between
    rounds).  Every result equals the request solved alone on its own handle;
    with batch_variants off no two of them share a batch."""
    paths, nuc = synth.variant_snts(snt_dir, 6, space="sd", zv=4, nv=4)
    reqs = [(paths[i % 6], 4 + i % 3, 500 + i) for i in range(10)]
    alone = [_alone(p, nuc, k, s, precision) for p, k, s in reqs]
    with Server(max_inflight=1, max_batch=4) as srv:
        ids = [srv.submit(p, nuc, k, precision=precision, seed=s) for p, k, s in reqs]
        for rid in ids:
            assert srv.wait(rid, 120000)
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    for (ev, info), ref in zip(res, alone):
        assert info["state"] == "done"
        np.testing.assert_allclose(ev, ref, atol=TOL_MIX)
    assert st["batched"] >= 6 and st["peak_batch"] >= 3, st
    assert st["index_uploads"] == 1, st
    distinct = reqs[:6]
    with Server(max_inflight=1, max_batch=4, batch_variants=False) as srv:
        ids = [srv.submit(p, nuc, k, precision=precision, seed=s) for p, k, s in distinct]
        for rid in ids:
            assert srv.wait(rid, 120000)
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    for (ev, info), ref in zip(res, alone[:6]):
        np.testing.assert_allclose(ev, ref, atol=TOL_MIX)
    assert st["batched"] == 0, st


def _zero_like_particle(src, dst):
    """Copy of a synthetic interaction file with every pp and nn TBME set to zero (pn
    and SPEs kept): on the same basis its like-particle pair tables are empty, a
    different sparsity pattern from the original's.  TBME lines are the lines of six
    fields (a b c d J V); the first line without a comment gives lp."""
    lines = open(src).read().splitlines()
    lp = int(next(ln for ln in lines if not ln.lstrip().startswith("!")).split()[0])
    out = []
    for ln in lines:
        f = ln.split()
        if len(f) == 6 and not ln.lstrip().startswith("!"):
            a, b, c, d = (int(x) for x in f[:4])
            if (a <= lp) == (b <= lp) == (c <= lp) == (d <= lp):
                f[5] = "0.0000000000e+00"
                ln = "  " + "  ".join(f)
        out.append(ln)
    open(dst, "w").write("\n".join(out) + "\n")
    return dst


def test_server_interactions_with_other_sparsity_on_one_basis(snt_dir):
    """Two interactions on one basis whose pp / nn tables differ in sparsity (one has
    every like-particle TBME zero): both requests are served, each through index
    tables of its own pattern, and each equals the request solved alone."""
    path, nuc = synth.config_snt(f"{snt_dir}/pat_a.snt", "sd", 2, 3, 8100, (-6, 0), 1.2)
    zero = _zero_like_particle(path, f"{snt_dir}/pat_b.snt")
    pa, pb = smgpu.Problem(path, nuc), smgpu.Problem(zero, nuc)
    assert pa.mdim == pb.mdim
    assert pb.pair_matrix(0)[0].size == 0 and pa.pair_matrix(0)[0].size > 0
    alone = [_alone(p, nuc, 4, 5, "mixed") for p in (path, zero)]
    with Server(max_inflight=2) as srv:
        ids = [srv.submit(p, nuc, 4, precision="mixed", seed=5) for p in (path, zero, zero, path)]
        for rid in ids:
            assert srv.wait(rid, 60000)
        res = [srv.result(rid) for rid in ids]
        st = srv.stats()
    assert st["failed"] == 0 and st["index_misses"] == 2, st
    for (ev, info), a in zip(res, [alone[0], alone[1], alone[1], alone[0]]):
        assert info["state"] == "done"
        assert np.abs(ev - a).max() <= TOL_MIX, ev - a
    assert np.abs(alone[0] - alone[1]).max() > 1e-3


def test_server_large_lane_on_cu_partition(snt_dir):
    reqs = _stream(snt_dir)
    alone = [_alone(p, n, k, s, "mixed") for p, n, k, s in reqs]
    md = [smgpu.Problem(p, n).mdim for p, n, _, _ in reqs]
    large = sorted(md)[len(md) // 2]        # about half the requests are large
    with pytest.raises(smgpu.SmgpuError):
        Server(max_inflight=4, reserve_cus=16)                  # needs large_mdim
    with pytest.raises(smgpu.SmgpuError):
        Server(max_inflight=4, large_mdim=large, reserve_cus=100000)

    def serve_all(srv):
        ids = [srv.submit(p, n, k, precision="mixed", seed=s) for p, n, k, s in reqs]
        for rid in ids:
            assert srv.wait(rid, 120000)
        return [srv.result(rid) for rid in ids]

    with Server(max_inflight=4, max_batch=2, large_mdim=large, reserve_cus=16) as srv:
        res = serve_all(srv)                # slot 0 on the CU partition while smaller requests run
        # the large requests one at a time, nothing smaller in the server: slot 0 on its
        # whole-device streams; then the mix again, back on the partition
        solo = []
        for (p, n, k, s), m, a in zip(reqs, md, alone):
            if m >= large:
                rid = srv.submit(p, n, k, precision="mixed", seed=s)
                assert srv.wait(rid, 120000)
                ev, info = srv.result(rid)
                assert info["slot"] == 0
                solo.append((n, ev, a))
        res2 = serve_all(srv)
        st = srv.stats()
    assert st["completed"] == 2 * len(reqs) + len(solo) and st["failed"] == 0
    assert len(solo) >= 2
    for n, ev, a in solo:
        assert np.abs(ev - a).max() <= TOL_MIX, (n, ev - a)
    for (path, nuc, k, s), a, (ev, info), (ev2, info2) in zip(reqs, alone, res, res2):
        print(nuc, info["mdim"], "slot", info["slot"], info2["slot"], flush=True)
        assert (info["slot"] == 0) == (info["mdim"] >= large), (nuc, info["mdim"], info["slot"])
        assert (info2["slot"] == 0) == (info2["mdim"] >= large), (nuc, info2["mdim"], info2["slot"])
        assert np.abs(ev - a).max() <= TOL_MIX, (nuc, ev - a)
        assert np.abs(ev2 - a).max() <= TOL_MIX, (nuc, ev2 - a)
