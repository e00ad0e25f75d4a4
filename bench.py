#!/usr/bin/env python3
"""Benchmark: the G-Store tile-grid sweep on one MI355X.

Workload (one "step" = one job of each algorithm on a resident graph):
  undirected Graph500 Kronecker graph, scale S (default 28), edge factor 16,
  seeds (1, 2), generated and converted to the tile grid on the device
  (untimed, like the reference's load), then
    PageRank      20 iterations, float degree factor (-j 21 -a 20)
    BFS           from 16 seeded random roots with degree > 0 (-j 20)
    WCC           (-j 23)
    k-core        k = 8 (-j 22 -a 8)
value = aggregate GTEPS over the step, using per algorithm: PageRank 2E per
iteration; BFS Graph500 M_reach per root (stored edges in the reached
component); WCC 2 x edges swept; k-core 2 x edges in swept tiles.

Prints one JSON line (stdout).  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_bytes(name, n, E, nz, nchunk, calls, scanned):
    """Algorithmic HBM bytes per launch of a kernel class (averaged over calls)."""
    if name == "pr_scatter":      # edges + tile metadata + prior read + acc read-modify-write
        return 4 * E + 12 * nz + 4 * (nchunk + 1) + 12 * n
    if name == "pr_finalize":     # acc read + factor read + prior write + acc clear
        return 16 * n
    if name == "wcc_hook":        # edges + metadata + parent read/write
        return 4 * E + 12 * nz + 8 * n
    if name == "wcc_check":
        return 4 * E + 12 * nz + 4 * n
    if name == "bfs_expand":      # active edges + frontier/visited/next bitmaps
        return (4 * scanned) / max(calls, 1) + 3 * n / 8
    if name == "kcore_scan":      # active edges + active bitmap + degree read-modify-write
        return (4 * scanned) / max(calls, 1) + n / 8 + 8 * n
    return None


def cpu_baseline(scale, threads):
    import oracle
    t0 = time.time()
    n = 1 << scale
    edges = oracle.kron(scale, 0, 16 << scale)
    start, grid, degree = oracle.convert(edges, n)
    del edges
    log(f"cpu baseline graph s{scale} ready in {time.time() - t0:.1f}s, E={grid.size}")
    iters = 5
    t = time.perf_counter()
    oracle.pagerank(n, start, grid, degree, iters, mode=1, nthreads=threads)
    dt = time.perf_counter() - t
    return {"value": round(2 * grid.size * iters / dt / 1e9, 4), "unit": "GTEPS", "cores": threads,
            "kind": "port",
            "sample": f"PageRank {iters} iterations (float degree factor) on Kron-{scale} "
                      f"(E={grid.size}), C/OpenMP restatement with CAS float atomics like pr.cpp:60-80, "
                      f"{dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=28)
    ap.add_argument("--pr-iters", type=int, default=20)
    ap.add_argument("--bfs-roots", type=int, default=16)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--cpu-scale", type=int, default=22)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import gstore_amd as ga

    log(f"generating Kronecker scale {args.scale} on the device")
    t0 = time.time()
    g = ga.Graph.kron(args.scale)
    gen_s = time.time() - t0
    n, E = g.n, g.E
    nchunk = (E + 4095) // 4096
    log(f"graph ready in {gen_s:.1f}s: N={n} E={E} nonempty tiles={g.nonempty_tiles}")
    deg = g.degree()
    rng = np.random.default_rng(20240601)
    roots = [int(r) for r in rng.choice(np.flatnonzero(deg), args.bfs_roots, replace=False)]

    # Graph500 M_reach per root (untimed pre-pass; BFS output is deterministic)
    m_reach = {}
    deg64 = deg.astype(np.int64)
    for r in roots:
        d = g.bfs(r).depth
        m_reach[r] = int(deg64[d > 0].sum() // 2)
    log("BFS M_reach per root computed")

    g.set_kernel_timing(True)
    kern = {}
    per = {"pagerank": [0.0, 0.0], "bfs": [0.0, 0.0, 0.0], "wcc": [0.0, 0.0], "kcore": [0.0, 0.0]}

    def absorb(st):
        for k, v in st["kernels"].items():
            a = kern.setdefault(k, [0.0, 0, 0])
            a[0] += v["ms"]
            a[1] += v["calls"]

    def step(record):
        g.pagerank(args.pr_iters, ga.PR_INTENDED, copy_out=False)
        st = g.stats()
        if record:
            absorb(st)
            per["pagerank"][0] += st["ms"]
            per["pagerank"][1] += 2 * E * args.pr_iters
        for r in roots:
            g.bfs(r, copy_out=False)
            st = g.stats()
            if record:
                absorb(st)
                kern.setdefault("bfs_expand", [0.0, 0, 0])[2] += st["edges_scanned"]
                per["bfs"][0] += st["ms"]
                per["bfs"][1] += m_reach[r]
                per["bfs"][2] += 2 * st["edges_scanned"]
        g.wcc(copy_out=False)
        st = g.stats()
        if record:
            absorb(st)
            per["wcc"][0] += st["ms"]
            per["wcc"][1] += 2 * st["edges_scanned"]
        g.kcore(args.k, copy_out=False)
        st = g.stats()
        if record:
            absorb(st)
            kern.setdefault("kcore_scan", [0.0, 0, 0])[2] += st["edges_scanned"]
            per["kcore"][0] += st["ms"]
            per["kcore"][1] += 2 * st["edges_scanned"]

    for i in range(args.warmup):
        step(False)
        log(f"warmup step {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t = time.perf_counter()
    for i in range(args.steps):
        step(True)
        log(f"timed step {i + 1}/{args.steps} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    ms_per_step = elapsed * 1e3 / args.steps
    traversed = per["pagerank"][1] + per["bfs"][1] + per["wcc"][1] + per["kcore"][1]
    value = traversed / elapsed / 1e9

    per_algo = {
        "pagerank": {"ms": per["pagerank"][0] / args.steps, "iterations": args.pr_iters,
                     "gteps": per["pagerank"][1] / (per["pagerank"][0] / 1e3) / 1e9},
        "bfs": {"ms_per_root": per["bfs"][0] / args.steps / len(roots), "roots": len(roots),
                "gteps_graph500": per["bfs"][1] / (per["bfs"][0] / 1e3) / 1e9,
                "gteps_scanned": per["bfs"][2] / (per["bfs"][0] / 1e3) / 1e9},
        "wcc": {"ms": per["wcc"][0] / args.steps, "gteps": per["wcc"][1] / (per["wcc"][0] / 1e3) / 1e9},
        "kcore": {"ms": per["kcore"][0] / args.steps, "k": args.k,
                  "gteps": per["kcore"][1] / (per["kcore"][0] / 1e3) / 1e9},
    }
    for a in per_algo.values():
        for k2 in list(a):
            if isinstance(a[k2], float):
                a[k2] = round(a[k2], 4)

    kernels = {}
    for name, (ms, calls, scanned) in kern.items():
        avg_ms = ms / max(calls, 1)
        b = kernel_bytes(name, n, E, g.nonempty_tiles, nchunk, calls, scanned)
        kernels[name] = {"total_ms": round(ms / args.steps, 3), "launches_per_step": calls / args.steps,
                         "avg_ms": round(avg_ms, 4),
                         "alg_gbs": round(b / (avg_ms / 1e3) / 1e9, 1) if b else None}
    dom = max(kern, key=lambda k: kern[k][0])
    d_ms, d_calls, d_scanned = kern[dom]
    d_avg = d_ms / d_calls
    d_bytes = kernel_bytes(dom, n, E, g.nonempty_tiles, nchunk, d_calls, d_scanned)
    achieved = d_bytes / (d_avg / 1e3) / 1e9
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(d_bytes), "avg_launch_ms": round(d_avg, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(args.cpu_scale, threads)

    out = {
        "metric": "GTEPS + % HBM peak per algo at Kron-28; batched BFS queries/s, p50/p99 lat",
        "value": round(value, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+u32", "data": "synthetic (Graph500 Kronecker, seeds 1,2, generated on device)",
        "config": {"workload": f"Kron-{args.scale} ef16 undirected, resident; PR {args.pr_iters} it + "
                               f"BFS {len(roots)} roots + WCC + k-core k={args.k}",
                   "scale": args.scale, "vertices": n, "stored_edges": E, "nonempty_tiles": g.nonempty_tiles,
                   "gen_convert_s": round(gen_s, 2), "parallelism": "single GPU"},
        "per_algo": per_algo, "kernels": kernels, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
