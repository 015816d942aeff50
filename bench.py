#!/usr/bin/env python3
"""Benchmark: converged eigensolves/s + p50/p99 latency of the MI355X shell-model
eigensolver against the plain FP64 TRL, one solve at a time.

Default workload (the serving stream the metric is quoted on): 512 eigensolve
requests with Poisson arrivals at 50/s; each request picks the sd or pf space
uniformly, proton and neutron valence numbers uniformly in 2..6, 2M = A % 2,
k uniformly in 1..10, and carries its own seeded synthetic interaction (SPEs
U[-6,0] MeV, TBMEs N(0,1.2 MeV)); bases range from mdim 640 (sd 2p2n) to
109,954,620 (pf 6p6n).  Each request is one main_sm solve (lm 300, ls 20,
tol 1e-8).

  served leg : the serving loop (smgpu_server_*) with FP32 H·v and FP64
               recurrence / small eigensolves / Rayleigh-quotient refinement,
               requests submitted at their arrival times, `--inflight` slots
               overlapping on the device, each running up to `--batch` requests
               on one basis in lockstep.  `value` = requests whose every
               eigenvalue agrees with the FP64 leg within 1e-5 MeV, divided by
               the time from the first arrival to the last result.
  FP64 leg   : the same requests, same arrival times, through a one-slot server
               in plain FP64 (every vector FP64, full reorthogonalisation every
               step): the plain FP64 TRL one solve at a time.  It takes about
               570 s on one MI355X, so it is recorded by `--leg fp64` (eigenvalues,
               per-request solve times and latencies, source hash) into
               profiles/stream512_fp64_leg.json; the default run serves the stream
               live, checks every result against the recorded FP64 eigenvalues,
               and re-runs a seeded sample of the FP64 solves live (about a minute
               of solve time) to show the record still holds.  Without a record
               (or for another stream) the FP64 leg runs live in full.
Latencies of both legs are arrival-to-result.  Beside them: the H·v kernel's
timing and rooflines on the stream's largest basis (HIP events on the launch
stream) and the oracle's CPU restatement of operate_H!/TRL timed on this host.

The stream's size is fixed by the workload (512 requests); --steps only sizes
the other workloads.  Other workloads: --workload served (one configuration
served in a burst), variants (64 interaction variants of Cr48), --mode hv
(H·v launches only, for profiler runs).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload stream|served|variants]
"""
import argparse
import json
import os
import platform
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nucleartoolkit.jl_amd"))

import numpy as np  # noqa: E402

METRIC = "converged eigensolves/s + p50/p99 latency vs plain FP64 TRL at |ΔE|≤1e-5 MeV"
HBM_PEAK_GBS = 8000.0
# LDS: 256 CUs x 128 bytes per clock x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4
TOL_MEV = 1e-5

CONFIGS = {
    # name: (space, zv, nv, seed offset, spe range, sigma, num_ev, lm, ls)
    "mg24": ("sd", 4, 4, 1, (-4.0, 0.0), 1.5, 5, 300, 20),
    "cr48": ("pf", 4, 4, 2, (-8.0, 0.0), 1.0, 10, 300, 20),
    "mn50": ("pf", 5, 5, 3, (-8.0, 0.0), 1.0, 5, 100, 20),
    "fe52": ("pf", 6, 6, 6, (-6.0, 0.0), 1.2, 5, 300, 20),
}


def percentile(x, q):
    return float(np.percentile(np.array(x), q)) if len(x) else None


def host_cpu():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "cores_visible": len(os.sched_getaffinity(0)), "cores_online": os.cpu_count()}


# ----------------------------------------------------------------------------- CPU
def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import sm_oracle as orc
    return orc


def _cpu_hv_seconds(orc, snt, nuc, reps, threads, warmups=3):
    orc.set_threads(threads)
    o = orc.OracleHamiltonian(snt, nuc)
    v = np.random.default_rng(0).standard_normal(o.mdim)
    for _ in range(warmups):
        o.operate_H(v)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.operate_H(v)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), o


def _mgs_seconds_per_element(n=2_000_000, k=16):
    rng = np.random.default_rng(1)
    V = rng.standard_normal((k, n))
    w = rng.standard_normal(n)
    for l in range(k - 1, -1, -1):           # warm-up
        w -= np.dot(V[l], w) * V[l]
    t0 = time.perf_counter()
    for l in range(k - 1, -1, -1):           # ReORTH (shellmodel_main.jl:1005-1012)
        w -= np.dot(V[l], w) * V[l]
    return (time.perf_counter() - t0) / (k * n)


def cpu_baseline_stream(reqs, fp64_hv, work, budget_s=25.0):
    """The oracle's restatement (C operate_H! with OpenMP over blocks, numpy MGS
    ReORTH) on this host.  Measured: (a) one full FP64 TRL solve of the C1 workload
    (Mg24, lowest 5) end to end; (b) the median operate_H! of the stream's bases up
    to 2e6 amplitudes, 3 warm-ups + >= 3 timed calls each, on all visible cores and
    for the largest of them on 1 core; (c) numpy MGS per basis element.  The stream
    rate extrapolates (b) and (c) to every request: seconds per FMA term from the
    measured bases (their median) times the request's terms per H·v and the FP64
    leg's H·v count, plus its ReORTH elements."""
    orc = _oracle()
    from smgpu import synth
    cores = len(os.sched_getaffinity(0))
    out = {"kind": "port", "cores": cores, "cpu": host_cpu(), "unit": "solves/s"}
    t_start = time.perf_counter()
    # (a) full C1 solve
    tmp = tempfile.mkdtemp(prefix="smgpu_cpu_")
    snt, nuc = synth.config_snt(os.path.join(tmp, "mg24.snt"), "sd", 4, 4, synth.BASE_SEED + 1, (-4.0, 0.0), 1.5)
    orc.set_threads(cores)
    o = orc.OracleHamiltonian(snt, nuc)
    t0 = time.perf_counter()
    en, elit, nhv, nres, _, _ = orc.trl(o, 5)
    out["c1_full_solve"] = {"nucleus": nuc, "mdim": int(o.mdim), "seconds": time.perf_counter() - t0,
                            "iterations": int(elit), "n_hv": int(nhv), "eigenvalues_MeV": [float(x) for x in en]}
    del o
    # (b) H·v on the stream's bases up to 2e6 amplitudes, largest first, within the budget
    seen, per_term = {}, []
    cand = sorted({(r.space, r.zv, r.nv): r for r in reqs}.values(),
                  key=lambda r: -work[(r.space, r.zv, r.nv)]["mdim"])
    cand = [r for r in cand if work[(r.space, r.zv, r.nv)]["mdim"] <= 2_000_000]
    measured = []
    for r in cand:
        if time.perf_counter() - t_start > budget_s:
            break
        key = (r.space, r.zv, r.nv)
        t_hv, oh = _cpu_hv_seconds(orc, r.snt, r.nucleus, 3, cores)
        seen[f"{r.space}{r.zv}p{r.nv}n"] = {"mdim": int(oh.mdim), "ms": t_hv * 1e3}
        per_term.append(t_hv / work[key]["terms"])
        measured.append((r, t_hv))
        del oh
    # one core: the largest measured basis up to 5e5 amplitudes (1 warm-up, 3 timed)
    one = next(((r, t) for r, t in measured if work[(r.space, r.zv, r.nv)]["mdim"] <= 500_000), None)
    if one is not None:
        t1, _ = _cpu_hv_seconds(orc, one[0].snt, one[0].nucleus, 3, 1, warmups=1)
        out["one_core"] = {"basis": f"{one[0].space}{one[0].zv}p{one[0].nv}n",
                           "mdim": work[(one[0].space, one[0].zv, one[0].nv)]["mdim"],
                           "ms_per_hv_1core": t1 * 1e3, "ms_per_hv_all_cores": one[1] * 1e3,
                           "speedup_all_cores": t1 / one[1]}
        orc.set_threads(cores)
    sec_per_term = float(np.median(per_term))
    sec_per_elem = _mgs_seconds_per_element()
    total = 0.0
    for r, (nhv, lm) in zip(reqs, fp64_hv):
        w = work[(r.space, r.zv, r.nv)]
        nhv = max(1, nhv)
        basis_avg = min(nhv, lm) / 2.0
        total += nhv * (w["terms"] * sec_per_term + basis_avg * w["mdim"] * sec_per_elem)
    out.update({"value": len(reqs) / total, "hv_bases": seen, "seconds_per_term": sec_per_term,
                "mgs_seconds_per_element": sec_per_elem, "stream_seconds_estimate": total,
                "sample": ("measured: one full FP64 TRL solve of Mg24 (C1) on the oracle, and the oracle's "
                           f"operate_H! on {len(seen)} of the stream's bases (<= 2e6 amplitudes; median of 3 after "
                           "3 warm-ups) plus numpy MGS; extrapolated: every request of the stream at its FP64 "
                           "leg's H·v count (seconds per FMA term x terms + ReORTH elements x seconds per element)"),
                "measure_s": time.perf_counter() - t_start})
    return out


# ------------------------------------------------------------------- H·v roofline
def hv_roofline(ham, reps, lds_model=True):
    """Dominant kernel (k_hv, FP32, one vector): HIP-event timing on its launch stream,
    algorithmic HBM bytes (every table array once + vector in + vector out) and the
    LDS bytes of the kernel's access pattern."""
    t32 = ham.time_operate_H("mixed", reps)
    t64 = ham.time_operate_H("fp64", max(3, reps // 2))
    tb32 = ham.table_bytes("mixed")
    B32 = sum(tb32.values())
    B64 = sum(ham.table_bytes("fp64").values())
    terms = ham.work_terms()
    W = terms["pn"] + terms["pp"] + terms["nn"] + terms["diag"]
    out = {"ms_fp32": t32, "ms_fp64": t64, "hv_per_s_fp32": 1e3 / t32, "hv_per_s_fp64": 1e3 / t64,
           "fma_terms": W, "pn_terms_padded": terms["pn_padded"], "terms_per_s_fp32": W / (t32 * 1e-3),
           "alg_bytes_fp32": B32, "alg_bytes_fp64": B64, "gbs_fp32": B32 / (t32 * 1e-3) / 1e9,
           "gbs_fp64": B64 / (t64 * 1e-3) / 1e9}
    lds = ham.lds_bytes("mixed")
    out["lds_bytes_fp32"] = lds
    out["lds_gbs_fp32"] = sum(lds.values()) / (t32 * 1e-3) / 1e9
    return out


def calibration_copies(ham, mib=256, reps=3):
    """Device-to-device copies of a known size on the handle's device: the byte
    counters of a profiler run (FETCH_SIZE / WRITE_SIZE) are scaled by what they read
    for these copies (each reads and writes `mib` MiB)."""
    import ctypes
    from smgpu._lib import check, lib
    L = lib()
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    check(L.smgpu_malloc(ham._h, n, ctypes.byref(a)))
    check(L.smgpu_malloc(ham._h, n, ctypes.byref(b)))
    try:
        for _ in range(reps):
            check(L.smgpu_memcpy(ham._h, b, a, n, 2))
        check(L.smgpu_synchronize(ham._h))
    finally:
        check(L.smgpu_free(ham._h, a))
        check(L.smgpu_free(ham._h, b))
    return {"bytes_read": n, "bytes_written": n, "copies": reps}


def source_hash():
    """sha256 (16 hex) of the library sources: the kernels and solver a record was taken with."""
    import hashlib
    d = os.path.join(ROOT, "nucleartoolkit.jl_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".cpp", ".h")):
            h.update(name.encode())
            h.update(open(os.path.join(d, name), "rb").read())
    return h.hexdigest()[:16]


def profiled_traffic(config):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, only when it was
    taken of the current kernel source (its hash) and this workload."""
    import hashlib
    path = os.path.join(ROOT, "profiles", f"hv_{config}_pmc.json")
    if not os.path.exists(path):
        return None, "no pmc summary"
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None, "unreadable pmc summary"
    src = os.path.join(ROOT, "nucleartoolkit.jl_amd", "csrc", "sm_hv.hip")
    sha = hashlib.sha256(open(src, "rb").read()).hexdigest()[:16]
    if d.get("sm_hv_sha16") != sha:
        return None, f"pmc summary of another kernel source ({d.get('sm_hv_sha16')} vs {sha})"
    return d.get("hbm_bytes_per_launch"), f"profiles/hv_{config}_pmc.json ({d.get('kernel', '')})"


def _line(value, unit, steps, warmup, t_all, config, extra, dtype="fp32 H·v, fp64 recurrence"):
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": t_all * 1e3 / max(1, steps), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded isospin interactions written as .snt)",
           "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------------ stream
def _serve(reqs, server, precision):
    """Submit every request at its arrival time; returns (results, wall s, first submit)."""
    t0 = time.perf_counter()
    ids = []
    for r in reqs:
        dt = t0 + r.arrival_s - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(r.snt, r.nucleus, r.num_ev, precision=precision, seed=r.seed % (1 << 31)))
    # progress in completion order (requests finish out of submission order)
    last = 0
    for rid in ids:
        while not server.wait(rid, 10000.0):
            done = server.stats()["completed"]
            print(f"# {precision}: {done}/{len(ids)} done at {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
        done = server.stats()["completed"]
        if done - last >= 64:
            last = done
            print(f"# {precision}: {done}/{len(ids)} done at {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
    t_all = time.perf_counter() - t0
    res = [server.result(rid, raise_on_error=False) for rid in ids]
    return res, t_all


def run_stream(args):
    import smgpu
    from smgpu import synth
    from smgpu.server import Server
    tmp = tempfile.mkdtemp(prefix="smgpu_stream_")
    lo, hi = (int(x) for x in args.particles.split(","))
    reqs = synth.serving_stream(tmp, args.requests, rate_per_s=args.rate, particles=(lo, hi))
    # per-basis sizes and H·v terms (host tables only)
    work = {}
    for r in reqs:
        key = (r.space, r.zv, r.nv)
        if key not in work:
            p = smgpu.Problem(r.snt, r.nucleus)
            work[key] = {"mdim": p.mdim, "nucleus": r.nucleus, "problem": p}
    print(f"# stream: {len(reqs)} requests over {len(work)} bases, last arrival {reqs[-1].arrival_s:.2f} s",
          file=sys.stderr, flush=True)
    # warm-up requests (untimed): one small basis per space
    warm = []
    for i in range(max(1, args.warmup)):
        space = ("sd", "pf")[i % 2]
        path, nuc = synth.config_snt(os.path.join(tmp, f"warm_{i}.snt"), space, 2, 2, 999 + i, (-6.0, 0.0), 1.2)
        warm.append((path, nuc))

    def warmup(srv, precision):
        ids = [srv.submit(p, n, 3, precision=precision) for p, n in warm]
        for rid in ids:
            srv.wait(rid)
            srv.result(rid)

    if args.leg == "fp64":
        seq = Server(max_inflight=1, host_threads=args.host_threads, max_batch=1)
        warmup(seq, "fp64")
        sres, t_seq = _serve(reqs, seq, "fp64")
        seq.close()
        rec = {"what": "sequential FP64 leg of the C4 serving stream: the same requests and arrival times as the "
                       "served leg, one slot, plain FP64 TRL", "source_sha16": source_hash(),
               "requests": len(reqs), "rate_per_s": args.rate, "particles": [lo, hi], "wall_s": t_seq,
               "host": platform.node(), "date": time.strftime("%Y-%m-%d %H:%M:%S"),
               "per_request": [{"i": i, "nucleus": r.nucleus, "k": r.num_ev, "state": info["state"],
                                "ev": [float("%.12g" % x) for x in ev], "solve_ms": round(info["solve_ms"], 3),
                                "latency_ms": round(info["latency_ms"], 3), "n_hv": info["n_hv"], "lm": info["lm"]}
                               for i, (r, (ev, info)) in enumerate(zip(reqs, sres))]}
        with open(args.record, "w") as f:
            json.dump(rec, f, separators=(",", ":"))
        done = sum(1 for _, i in sres if i["state"] == "done")
        print(json.dumps({"leg": "fp64", "record": args.record, "wall_s": t_seq, "done": done,
                          "solves_per_s": done / t_seq, "source_sha16": rec["source_sha16"]}), flush=True)
        return
    srv = Server(max_inflight=args.inflight, host_threads=args.host_threads, max_batch=args.batch,
                 large_mdim=args.large_mdim, reserve_cus=args.reserve_cus)
    warmup(srv, "mixed")
    res, t_all = _serve(reqs, srv, "mixed")
    sst = srv.stats()
    srv.close()
    print(f"# served leg: {t_all:.1f} s", file=sys.stderr, flush=True)
    if args.timeline:
        with open(args.timeline, "w") as f:
            json.dump([dict(i=k, nucleus=r.nucleus, space=r.space, zv=r.zv, nv=r.nv, k=r.num_ev,
                            arrival_s=r.arrival_s, ev=[float(x) for x in ev], **info)
                       for k, (r, (ev, info)) in enumerate(zip(reqs, res))], f)
    # the sequential FP64 leg: the recorded full run of this stream (its eigenvalues check
    # every served result), and a live re-run of a sample of its requests
    rec = json.load(open(args.record)) if os.path.exists(args.record) else None
    if rec is not None and (rec["requests"], rec["rate_per_s"], rec["particles"]) != (len(reqs), args.rate, [lo, hi]):
        rec = None
    if rec is None:
        seq = Server(max_inflight=1, host_threads=args.host_threads, max_batch=1)
        warmup(seq, "fp64")
        sres, t_seq = _serve(reqs, seq, "fp64")
        seq.close()
        fp64_src = {"source": "measured in this run"}
        ev64 = [ev for ev, _ in sres]
        st64 = [i["state"] for _, i in sres]
        lat64 = [i["latency_ms"] for _, i in sres if i["state"] == "done"]
        nhv64 = [(i["n_hv"], i["lm"]) for _, i in sres]
    else:
        t_seq = rec["wall_s"]
        pr = rec["per_request"]
        ev64 = [np.array(q["ev"]) for q in pr]
        st64 = [q["state"] for q in pr]
        lat64 = [q["latency_ms"] for q in pr if q["state"] == "done"]
        nhv64 = [(q["n_hv"], q["lm"]) for q in pr]
        # live spot check: a seeded sample of the recorded requests, solved again one at a
        # time in FP64, within about a minute of recorded solve time
        rng = np.random.default_rng(7)
        sample, budget = [], 0.0
        for i in rng.permutation(len(pr)):
            if budget + pr[i]["solve_ms"] * 1e-3 > args.spot_s or len(sample) >= 24:
                continue
            sample.append(int(i))
            budget += pr[i]["solve_ms"] * 1e-3
        seq = Server(max_inflight=1, host_threads=args.host_threads, max_batch=1)
        warmup(seq, "fp64")
        warmup(seq, "mixed")
        live = []

        def solo(r, **kw):
            rid = seq.submit(r.snt, r.nucleus, r.num_ev, seed=r.seed % (1 << 31), **kw)
            seq.wait(rid)
            return seq.result(rid, raise_on_error=False)

        for i in sorted(sample):
            r = reqs[i]
            ev, info = solo(r, precision="fp64")
            # the same request one at a time as FP64 with partial reorthogonalisation and as
            # mixed precision: the speed-up of the served mixed form split into its parts
            evp, infp = solo(r, precision="fp64", reorth="partial")
            evm, infm = solo(r, precision="mixed")
            dev = (lambda a: float(np.abs(a - ev64[i]).max()) if len(a) == len(ev64[i]) else None)
            live.append({"i": i, "nucleus": r.nucleus, "solve_ms": info["solve_ms"],
                         "recorded_ms": pr[i]["solve_ms"], "max_abs_dev_MeV": dev(ev),
                         "fp64_partial_ms": infp["solve_ms"], "fp64_partial_dev_MeV": dev(evp),
                         "mixed_ms": infm["solve_ms"], "mixed_dev_MeV": dev(evm)})
        seq.close()
        ratio = sum(q["solve_ms"] for q in live) / max(1e-9, sum(q["recorded_ms"] for q in live))
        t_full = sum(q["solve_ms"] for q in live)
        split = {"fp64_full_s": t_full * 1e-3,
                 "fp64_partial_s": sum(q["fp64_partial_ms"] for q in live) * 1e-3,
                 "mixed_one_at_a_time_s": sum(q["mixed_ms"] for q in live) * 1e-3,
                 "note": "the sampled requests solved one at a time: full FP64 TRL (the baseline), FP64 with "
                         "partial reorthogonalisation (the algorithm change alone), mixed precision with partial "
                         "reorthogonalisation (algorithm and precision); the rest of the served speed-up is "
                         "concurrency and batching"}
        fp64_src = {"source": f"recorded sequential FP64 run ({os.path.relpath(args.record, ROOT)}, {rec['date']}, "
                              f"same requests and arrival times)",
                    "source_sha16_recorded": rec["source_sha16"], "source_sha16_now": source_hash(),
                    "speedup_split": split,
                    "spot_check": {"requests": len(live), "live_over_recorded_time": ratio,
                                   "max_abs_dev_MeV": max((q["max_abs_dev_MeV"] or 0.0) for q in live) if live
                                   else None, "solves": live}}
    dE, ok, bad = [], 0, []
    for i, (a, ia) in enumerate(res):
        if ia["state"] == "done" and st64[i] == "done" and len(a) == len(ev64[i]) > 0:
            d = float(np.abs(a - ev64[i]).max())
            dE.append(d)
            if d <= TOL_MEV:
                ok += 1
            else:
                bad.append({"request": i, "nucleus": reqs[i].nucleus, "k": reqs[i].num_ev, "dE_MeV": d})
        else:
            bad.append({"request": i, "nucleus": reqs[i].nucleus, "state": [ia["state"], st64[i]],
                        "error": ia.get("error", "")})
    value = ok / t_all
    fp64_rate = sum(1 for x in st64 if x == "done") / t_seq
    lat = [i["latency_ms"] for _, i in res if i["state"] == "done"]
    mdims = [i["mdim"] for _, i in res]
    # kernel rooflines on the stream's largest basis (the pf 6p6n requests dominate its H·v time)
    big = max(work.values(), key=lambda w: w["mdim"])
    ham = smgpu.Hamiltonian(big["problem"])
    hv = hv_roofline(ham, args.hv_reps)
    del ham
    cpu = None
    if not args.no_cpu:
        # FMA terms per H·v of every basis (device tables, one basis at a time) for the CPU model
        for w in work.values():
            h = smgpu.Hamiltonian(w["problem"])
            t = h.work_terms()
            w["terms"] = t["pn"] + t["pp"] + t["nn"] + t["diag"]
            del h
        cpu = cpu_baseline_stream(reqs, nhv64, work)
    c5 = None
    if not args.no_c5:
        # the interaction-variant sweep (64 Cr48 variants, every one checked against FP64)
        print("# c5: 64 Cr48 interaction variants", file=sys.stderr, flush=True)
        c5 = variants_leg(64, 2, 2, args.host_threads)[0]
    traffic, traffic_src = profiled_traffic("fe52")
    roof = {"bound": "hbm", "achieved": hv["gbs_fp32"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": hv["gbs_fp32"] / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "kernel": f"k_hv FP32 one vector on {big['nucleus']} (mdim {big['mdim']})",
            "lds": {"achieved": hv["lds_gbs_fp32"], "peak": LDS_PEAK_GBS, "unit": "GB/s",
                    "frac": hv["lds_gbs_fp32"] / LDS_PEAK_GBS}}
    # per basis: requests, and each leg's summed solve time, H·v count and smallest lm
    fp64_ms = [q["solve_ms"] for q in rec["per_request"]] if rec is not None else [i["solve_ms"] for _, i in sres]
    counts = {}
    for r, (_, ia), t64, (nh64, lm64) in zip(reqs, res, fp64_ms, nhv64):
        k = f"{r.space} {r.zv}p{r.nv}n"
        c = counts.setdefault(k, {"mdim": work[(r.space, r.zv, r.nv)]["mdim"], "requests": 0, "mixed_solve_s": 0.0,
                                  "mixed_n_hv": 0, "fp64_solve_s": 0.0, "fp64_n_hv": 0, "fp64_lm": 300})
        c["requests"] += 1
        c["mixed_solve_s"] += ia["solve_ms"] * 1e-3
        c["mixed_n_hv"] += ia["n_hv"]
        c["fp64_solve_s"] += t64 * 1e-3
        c["fp64_n_hv"] += nh64
        c["fp64_lm"] = min(c["fp64_lm"], lm64)
    counts = dict(sorted(counts.items(), key=lambda kv: -kv[1]["mdim"]))
    _line(value, "solves/s", len(reqs), args.warmup, t_all,
          {"workload": (f"C4 serving stream: {len(reqs)} requests, Poisson {args.rate}/s, sd/pf, valence {lo}..{hi} "
                        f"per species, 2M = A%2, k 1..10, per-request seeded interactions (SPE U[-6,0], TBME "
                        f"N(0,1.2)); mdim {min(mdims)}..{max(mdims)}; served mixed precision in {args.inflight} "
                        f"slots (batches of up to {args.batch}) vs the plain FP64 TRL one solve at a time"),
           "requests": len(reqs), "rate_per_s": args.rate, "particles": [lo, hi], "inflight": args.inflight,
           "batch": args.batch, "large_mdim": args.large_mdim, "reserve_cus": args.reserve_cus,
           "lm": 300, "ls": 20, "tol": 1e-8, "steps_note": "one step = one request; the "
           "stream fixes its own size (--steps is not used by this workload)"},
          {"latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99), "note": "arrival to result"},
           "fp64": {"solves_per_s": fp64_rate, "wall_s": t_seq, "latency_ms_p50": percentile(lat64, 50),
                    "latency_ms_p99": percentile(lat64, 99), "note": "same requests and arrival times, one slot, "
                    "plain FP64 (full reorthogonalisation every step)", **fp64_src},
           "speedup_vs_fp64": value / fp64_rate, "within_tol": ok, "failed_or_off": bad[:8], "n_off": len(bad),
           "max_abs_dE_vs_fp64_MeV": max(dE) if dE else None,
           "served": {"wall_s": t_all, "peak_running": sst["peak_running"], "hbm_peak_bytes": sst["hbm_peak"],
                      "batched": sst["batched"], "index_misses": sst["index_misses"],
                      "index_hits": sst["index_hits"],
                      "lm_bounded": sum(1 for _, i in res if i["lm"] < 300),
                      "fallbacks": sum(1 for _, i in res if i.get("n_fallback", 0))},
           "bases": counts, "hv": hv, "roofline": roof, "cpu_baseline": cpu, "c5_variants": c5})


# ---------------------------------------------------------------- served burst
def run_served(args):
    import smgpu
    from smgpu import synth
    from smgpu.server import Server
    space, zv, nv, soff, spe, sig, nev, lm, ls = CONFIGS[args.config]
    tmp = tempfile.mkdtemp(prefix="smgpu_bench_")
    snt, nuc = synth.config_snt(os.path.join(tmp, f"{args.config}.snt"), space, zv, nv,
                                synth.BASE_SEED + soff, spe, sig)
    prob = smgpu.Problem(snt, nuc)
    ham = smgpu.Hamiltonian(prob, 0)
    if args.mode == "hv":
        out = {"mode": "hv", "config": args.config, "mdim": prob.mdim, **hv_roofline(ham, args.hv_reps)}
        out["copy_calibration"] = calibration_copies(ham)
        print(json.dumps(out))
        return
    # plain FP64 path, one solve at a time (tables resident; solve only): >= 3 timed solves
    ham.trl(nev, lm=lm, ls=ls, precision="fp64")
    lat64, ev64, st64 = [], None, None
    for i in range(max(3, args.fp64_steps)):
        t0 = time.perf_counter()
        ev64, _, st64 = ham.trl(nev, lm=lm, ls=ls, precision="fp64", seed=3000 + i)
        lat64.append((time.perf_counter() - t0) * 1e3)
    fp64_rate = 1e3 / float(np.mean(lat64))
    # FP64 with partial reorthogonalisation, one at a time (separates the algorithm change)
    t0 = time.perf_counter()
    evp, _, stp = ham.trl(nev, lm=lm, ls=ls, precision="fp64", reorth="partial")
    fp64_partial_ms = (time.perf_counter() - t0) * 1e3
    ev1, _, st1 = ham.trl(nev, lm=lm, ls=ls, precision="mixed")
    t0 = time.perf_counter()
    ev1, _, st1 = ham.trl(nev, lm=lm, ls=ls, precision="mixed")
    mixed_single_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    evf, _, stf = ham.trl(nev, lm=lm, ls=ls, precision="mixed", reorth="full")
    mixed_full_ms = (time.perf_counter() - t0) * 1e3
    srv = Server(max_inflight=args.inflight, host_threads=args.host_threads, max_batch=args.batch)
    warm = [srv.submit(snt, nuc, nev, lm=lm, ls=ls, precision="mixed", seed=1000 + i)
            for i in range(max(args.warmup, args.inflight * args.batch))]
    for rid in warm:
        srv.wait(rid)
        srv.result(rid)
    t_all0 = time.perf_counter()
    ids = [srv.submit(snt, nuc, nev, lm=lm, ls=ls, precision="mixed", seed=2000 + i) for i in range(args.steps)]
    for rid in ids:
        srv.wait(rid)
    t_all = time.perf_counter() - t_all0
    res = [srv.result(rid) for rid in ids]
    sstats = srv.stats()
    srv.close()
    dE = [float(np.abs(ev - ev64).max()) for ev, _ in res]
    n_ok = sum(1 for d in dE if d <= TOL_MEV)
    value = n_ok / t_all
    lat = [info["latency_ms"] for _, info in res]
    hv = hv_roofline(ham, args.hv_reps)
    traffic, traffic_src = profiled_traffic(args.config)
    roof = {"bound": "hbm", "achieved": hv["gbs_fp32"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": hv["gbs_fp32"] / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "lds": {"achieved": hv["lds_gbs_fp32"], "peak": LDS_PEAK_GBS, "unit": "GB/s",
                    "frac": hv["lds_gbs_fp32"] / LDS_PEAK_GBS}}
    _line(value, "solves/s", args.steps, args.warmup, t_all,
          {"workload": f"{nuc} {space} {zv}p{nv}n 2M={prob.info.twoM} mdim {prob.mdim}, lowest {nev}, tol 1e-8, "
                       f"lm {lm}, ls {ls}; {args.steps} requests submitted at once, {args.inflight} slots, batches "
                       f"of up to {args.batch}", "nucleus": nuc, "mdim": prob.mdim, "num_ev": nev},
          {"latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99), "note": "submit to result; "
                          "all requests submitted at t=0"},
           "fp64": {"solves_per_s": fp64_rate, "solves_timed": len(lat64), "latency_ms": lat64,
                    "iterations": st64["iterations"], "n_hv": st64["n_hv"], "n_reorth": st64["n_reorth"]},
           "fp64_partial_reorth": {"solves_per_s": 1e3 / fp64_partial_ms, "n_reorth": stp["n_reorth"],
                                   "max_abs_dE_vs_fp64_MeV": float(np.abs(evp - ev64).max())},
           "speedup_vs_fp64": value / fp64_rate,
           "mixed_single": {"solves_per_s": 1e3 / mixed_single_ms, "n_reorth": st1["n_reorth"],
                            "max_abs_dE_vs_fp64_MeV": float(np.abs(ev1 - ev64).max())},
           "mixed_single_full_reorth": {"solves_per_s": 1e3 / mixed_full_ms, "n_reorth": stf["n_reorth"],
                                        "max_abs_dE_vs_fp64_MeV": float(np.abs(evf - ev64).max())},
           "served": {"within_tol": n_ok, "requests": len(res), "max_abs_dE_vs_fp64_MeV": max(dE),
                      "peak_running": sstats["peak_running"]},
           "hv": hv, "roofline": roof})


def variants_leg(n_variants, inflight, batch, host_threads):
    """64 interaction variants of Cr48 (N(0, 0.05) on every TBME), lowest 3, all
    submitted at once to the serving loop; every variant's eigenvalues are checked
    against its plain FP64 solve (one at a time, coefficients swapped on one handle).
    Returns (summary dict, served wall s, per-request results, nucleus, nev)."""
    import smgpu
    from smgpu import synth
    from smgpu.server import Server
    tmp = tempfile.mkdtemp(prefix="smgpu_variants_")
    paths, nuc = synth.variant_snts(tmp, n_variants)
    nev = 3
    srv = Server(max_inflight=inflight, host_threads=host_threads, max_batch=batch)
    warm = [srv.submit(paths[i % len(paths)], nuc, nev, seed=900 + i) for i in range(inflight)]
    for rid in warm:
        srv.wait(rid)
        srv.result(rid)
    t0 = time.perf_counter()
    ids = [srv.submit(p, nuc, nev, seed=1000 + i) for i, p in enumerate(paths)]
    for rid in ids:
        srv.wait(rid)
    t_all = time.perf_counter() - t0
    res = [srv.result(rid) for rid in ids]
    sst = srv.stats()
    srv.close()
    prob = smgpu.Problem(paths[0], nuc)
    ham = smgpu.Hamiltonian(prob)
    dE, t64 = [], 0.0
    for i in range(len(paths)):
        prob.set_interaction(paths[i])
        ham.update_interaction(prob)
        t1 = time.perf_counter()
        ev64, _, _ = ham.trl(nev, precision="fp64", seed=1000 + i)
        t64 += time.perf_counter() - t1
        dE.append(float(np.abs(res[i][0] - ev64).max()))
    fp64_rate = len(paths) / t64
    ok = sum(1 for d in dE if d <= TOL_MEV)
    lat = [info["latency_ms"] for _, info in res]
    summary = {"solves_per_s": ok / t_all, "fp64_solves_per_s": fp64_rate, "speedup_vs_fp64": ok / t_all / fp64_rate,
               "variants": len(paths), "variants_checked": len(paths), "within_tol": ok,
               "max_abs_dE_vs_fp64_MeV": max(dE), "latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
               "index_uploads": sst["index_uploads"], "batched": sst["batched"], "inflight": inflight, "batch": batch}
    return summary, t_all, res, nuc, nev


def run_variants(args):
    c5, t_all, res, nuc, nev = variants_leg(args.variants, args.inflight, args.batch, args.host_threads)
    _line(c5["solves_per_s"], "solves/s", c5["variants"], args.inflight, t_all,
          {"workload": f"{c5['variants']} interaction variants of {nuc} (pf 4p4n, mdim {res[0][1]['mdim']}), lowest "
                       f"{nev} each, one shared index, {args.inflight} slots", "nucleus": nuc, "variants": c5["variants"]},
          {"latency_ms": c5["latency_ms"],
           "fp64": {"solves_per_s": c5["fp64_solves_per_s"], "variants_checked": c5["variants_checked"]},
           "speedup_vs_fp64": c5["speedup_vs_fp64"], "max_abs_dE_vs_fp64_MeV": c5["max_abs_dE_vs_fp64_MeV"],
           "within_tol": c5["within_tol"], "index_uploads": c5["index_uploads"], "batched": c5["batched"]})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="requests of the served burst (--workload served)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="stream", choices=["stream", "served", "variants"])
    ap.add_argument("--config", default="cr48", choices=sorted(CONFIGS), help="--workload served / --mode hv")
    ap.add_argument("--mode", default="solve", choices=["solve", "hv"])
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c5", action="store_true", help="stream: skip the 64-variant Cr48 sweep summary")
    ap.add_argument("--hv-reps", type=int, default=20)
    ap.add_argument("--inflight", type=int, default=None)
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--host-threads", type=int, default=4)
    ap.add_argument("--variants", type=int, default=64)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--particles", default="2,6", help="valence range per species")
    ap.add_argument("--leg", default="served", choices=["served", "fp64"],
                    help="stream: fp64 runs only the sequential FP64 leg and writes --record")
    ap.add_argument("--record", default=os.path.join(ROOT, "profiles", "stream512_fp64_leg.json"),
                    help="stream: the recorded sequential FP64 leg (read by the served leg, written by --leg fp64)")
    ap.add_argument("--large-mdim", type=int, default=0,
                    help="stream: served leg runs requests of at least this dimension one at a time, the rest "
                         "smallest first (0: arrival order)")
    ap.add_argument("--reserve-cus", type=int, default=0,
                    help="stream, with --large-mdim: the large requests run in slot 0 on all CUs but this many")
    ap.add_argument("--timeline", default=None,
                    help="stream: write every served request's record (slot, times, counts) to this JSON file")
    ap.add_argument("--spot-s", type=float, default=40.0,
                    help="stream: recorded FP64 solve time re-run live as a spot check")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); independent solves shard one per GPU "
              "and are measured per GPU", file=sys.stderr)
        sys.exit(2)
    if args.inflight is None:
        args.inflight = 4 if args.workload == "stream" else 2
    if args.batch is None:
        # stream: 4 slots of lockstep batches of two (measured 2.36 solves/s against 2.22 with
        # single solves, 2.30 with batches of three, 2.34 with 3 slots of two)
        args.batch = 2
    if args.mode == "hv" or args.workload == "served":
        run_served(args)
    elif args.workload == "variants":
        run_variants(args)
    else:
        run_stream(args)


if __name__ == "__main__":
    main()
