"""Serving at the benchmark's scales: six resident undirected Kronecker graphs
(scales 18 to 28), queries arriving as a Poisson process with the 70/20/10
BFS / personalised PageRank / WCC mix, served by continuous batching under an
HBM budget.  Every result is checked against the oracle: on the scale 18-22
graphs against the oracle run on the host (BFS depths and WCC labels bit for
bit, personalised PageRank within 1e-6 L1 of the FP64 oracle over all
vertices); on scales 24-28 against the oracle's fixtures (depth and label
CRC32s from the fixture roots, personalised PageRank against the FP64 sample
at 24 and 26).  At scale 28, where no FP64 personalised PageRank was run,
its queries are checked against the same query run alone and its total mass.
"""
import json
import os
import time
import zlib

import numpy as np
import pytest

import gstore_amd as ga
import oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ORACLE_SCALES = (18, 20, 22)
FIXTURE_SCALES = tuple(s for s in (24, 26, 28) if os.path.exists(os.path.join(GOLDEN, f"scale_kron{s}.json")))
THREADS = min(16, len(os.sched_getaffinity(0)))


def _say(*a):
    print("[gpu-serve-scale]", *a, flush=True)


def crc(a):
    return zlib.crc32(np.ascontiguousarray(a).view(np.uint8))


def pct(xs, q):
    xs = sorted(xs)
    return round(xs[min(len(xs) - 1, int(q * len(xs)))], 2) if xs else None


@pytest.fixture(scope="module")
def fleet():
    out = {}
    for s in ORACLE_SCALES + FIXTURE_SCALES:
        t0 = time.time()
        g = ga.Graph.kron(s)
        rec = {"g": g, "n": g.n}
        if s in ORACLE_SCALES:
            start, grid, deg = oracle.convert(oracle.kron(s, 0, 16 << s), 1 << s)
            # the device's conversion is the oracle's (the grid up to in-tile order)
            gs, gg, gd = g.arrays()
            assert np.array_equal(gs, start) and np.array_equal(gd, deg)
            assert oracle.grid_checksum(gs, gg) == oracle.grid_checksum(start, grid)
            rec.update(start=start, grid=grid, deg=deg, nz=np.flatnonzero(deg))
        else:
            with open(os.path.join(GOLDEN, f"scale_kron{s}.json")) as f:
                rec["fx"] = json.load(f)
            p = os.path.join(GOLDEN, f"scale_kron{s}_samples.npz")
            rec["smp"] = np.load(p) if os.path.exists(p) else None
        out[s] = rec
        _say(f"Kron-{s} resident ({time.time() - t0:.1f}s)")
    yield out
    for rec in out.values():
        rec["g"].close()


def _plan(fleet, nq, rng):
    scales = sorted(fleet)
    plan = []
    for _ in range(nq):
        s = scales[int(rng.integers(0, len(scales)))]
        rec = fleet[s]
        u = rng.random()
        if u < 0.7:
            kind = ga.Q_BFS
            arg = int(rng.choice(rec["nz"])) if "nz" in rec else int(rng.choice([b["root"] for b in rec["fx"]["bfs"]]))
        elif u < 0.9:
            kind = ga.Q_PPR
            if "nz" in rec:
                arg = int(rng.choice(rec["nz"]))
            elif "ppr64" in rec["fx"]:
                arg = rec["fx"]["ppr64"]["source"]
            else:
                arg = int(rec["fx"]["bfs"][0]["root"])
        else:
            kind, arg = ga.Q_WCC, 0
        plan.append((s, kind, arg))
    return plan


def test_poisson_mix_over_six_graphs(fleet):
    rng = np.random.default_rng(2718)
    nq, rate = 240, 60.0
    plan = _plan(fleet, nq, rng)
    gaps = rng.exponential(1.0 / rate, nq)
    outs = {}
    with ga.Server(hbm_budget=96 << 30, workers=4) as srv:
        gid = {s: srv.add_graph(fleet[s]["g"]) for s in sorted(fleet)}
        for s in sorted(fleet):
            _say(f"Kron-{s} BFS batching rule", srv.bfs_policy(gid[s]))
        t0 = time.perf_counter()
        due = t0
        for (s, kind, arg), gap in zip(plan, gaps):
            due += gap
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            n = fleet[s]["n"]
            out = np.empty(n, np.uint8 if kind == ga.Q_BFS else np.uint32 if kind == ga.Q_WCC else np.float32)
            outs[srv.submit(gid[s], kind, arg, out)] = (s, kind, arg, out)
        srv.drain()
        dt = time.perf_counter() - t0
        res = srv.poll(nq + 10)
        mem = srv.memory()
        assert srv.last_error() == ""
    assert len(res) == nq and all(r["status"] == 0 for r in res)
    lat = {}
    for r in res:
        lat.setdefault(r["kind"], []).append(r["t_done_ms"] - r["t_submit_ms"])
    names = {ga.Q_BFS: "bfs", ga.Q_PPR: "ppr", ga.Q_WCC: "wcc"}
    _say(f"{nq} queries offered at {rate}/s served in {dt:.2f}s; workspace peak accounting {mem}; latency ms",
         {names[k]: {"n": len(v), "p50": pct(v, 0.5), "p99": pct(v, 0.99)} for k, v in lat.items()},
         "BFS batch sizes", sorted({r["batch_size"] for r in res if r["kind"] == ga.Q_BFS}))

    memo = {}
    checked = {"oracle": 0, "fixture": 0, "alone": 0}
    for r in res:
        s, kind, arg, out = outs[r["id"]]
        rec = fleet[s]
        if "nz" in rec:   # the oracle on the host
            key = (s, kind, arg)
            if key not in memo:
                if kind == ga.Q_BFS:
                    memo[key] = oracle.bfs(rec["n"], rec["start"], rec["grid"], arg, nthreads=THREADS)[0]
                elif kind == ga.Q_WCC:
                    memo[key] = oracle.wcc(rec["n"], rec["start"], rec["grid"], nthreads=THREADS)[0]
                else:
                    memo[key] = oracle.ppr(rec["n"], rec["start"], rec["grid"], rec["deg"], arg, 20, nthreads=THREADS)
            want = memo[key]
            if kind == ga.Q_PPR:
                l1 = float(np.abs(out.astype(np.float64) - want).sum())
                assert l1 <= 1e-6, (s, arg, l1)
            else:
                assert np.array_equal(out, want), (s, kind, arg)
            checked["oracle"] += 1
        elif kind == ga.Q_BFS:
            fx = next(b for b in rec["fx"]["bfs"] if b["root"] == arg)
            assert crc(out) == fx["depth_crc32"], (s, arg)
            assert r["value"] == fx["reached"], (s, arg, r["batch_size"], r["value"], fx["reached"])
            checked["fixture"] += 1
        elif kind == ga.Q_WCC:
            assert crc(out) == rec["fx"]["wcc"]["label_crc32"] and r["value"] == rec["fx"]["wcc"]["ncomp"]
            checked["fixture"] += 1
        elif rec["smp"] is not None and "ppr64" in rec["fx"]:
            idx = rec["smp"]["idx"]
            err = np.abs(out[idx].astype(np.float64) - rec["smp"]["ppr64"])
            assert err.max() <= 1e-6 and err.sum() <= 1e-6, (s, float(err.max()), float(err.sum()))
            assert abs(float(out.astype(np.float64).sum()) - rec["fx"]["ppr64"]["sum"]) <= 1e-6
            checked["fixture"] += 1
        else:
            key = (s, kind, arg)
            if key not in memo:
                memo[key] = rec["g"].ppr(arg, 20)
            assert np.max(np.abs(out - memo[key])) <= 1e-6 and float(out.astype(np.float64).sum()) <= 1.0 + 1e-5
            checked["alone"] += 1
    _say("checked", checked)
