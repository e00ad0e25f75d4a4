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
value = aggregate GTEPS over the step.  Traversed edges per job, independent
of the GPU method: PageRank 2E per iteration; BFS Graph500 M_reach per root
(stored edges with a reached endpoint); WCC 2E per run (every edge once);
k-core 2 x the stored edges the reference's tile sweeps scan (tiles whose row
or column holds a removed vertex, counted by the tile method in an untimed
pre-pass).

Per algorithm the line also gives: ref_formula_gbs, the fixed algorithmic
bytes of the reference's tile sweeps per unit of time (a rate, not a fraction:
methods that skip work the tile sweep does exceed the HBM peak on it); and,
where a rocprofv3 --pmc record for this workload is kept in
profiles/traffic.json and its kernel time is within 30% of this run's,
traffic_bytes and traffic_frac = measured HBM bytes / time / 8 TB/s.  The
roofline entry is one PageRank iteration.  A `serving` entry measures batched
BFS serving on the same graph, and `cpu_baseline` times the oracle's threaded
forms of all four jobs on the host.

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


def alg_bytes_pr_iter(n, E, T):
    """PageRank, one iteration: tiles + offsets + 24 B per vertex (prior read,
    accumulator read-modify-write and read, degree factor read, new prior
    write).  Fixed formula, independent of the GPU method."""
    return 4 * E + 8 * (T + 1) + 24 * n


def alg_bytes_bfs(n, e_scan, t_scan, levels, reached):
    """BFS, one root: per level 4 B per scanned edge + 8 B per scanned tile
    (reference tile-skip rule) + three N-bit bitmaps, plus 1 B per reached vertex."""
    return 4 * e_scan + 8 * t_scan + levels * 3 * n / 8 + reached


def alg_bytes_wcc(n, E, T, passes, jumps):
    return passes * (4 * E + 8 * (T + 1) + 8 * n) + jumps * 8 * n


def alg_bytes_kcore(n, e_scan, t_scan, rounds):
    return 4 * e_scan + 8 * t_scan + rounds * (4 * n + n / 8)


def measured_traffic(workload, key, live_ms):
    """HBM-side bytes of one unit of work (a PageRank iteration, one BFS root,
    one WCC or k-core run) measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy calibration) and kept in
    profiles/traffic.json with the kernel time of that unit in an unprofiled
    run.  Returned only when this run's time is within 30% of it (the same
    kernels doing the same work)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload, {}).get(key)
    except (OSError, ValueError):
        return None
    ref_ms = rec and rec.get("ms", rec.get("iteration_ms"))
    if not rec or not ref_ms or abs(live_ms - ref_ms) > 0.3 * ref_ms:
        return None
    return {"bytes": int(rec["fetch_bytes"] + rec["write_bytes"]), "source": rec["source"], "rec": rec}


# kernel classes of one PageRank iteration, per method
PR_ITER_KERNELS = ("pb_gather", "pb_accum", "pb_reduce", "pr_pull", "pr_scatter", "pr_finalize")


def host_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when set (the GPU boxes give each GPU a share of
    the host and set it)."""
    visible = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(visible, cap) if cap > 0 else visible), visible


def cpu_baseline(scale, threads, visible):
    """The oracle's threaded forms (tile sweeps with atomics, the reference's
    OpenMP scheme) of the four jobs on Kron-<scale>, with the GPU line's
    traversed-edge conventions: PageRank 5 iterations, BFS from 2 of the
    bench's roots, WCC, k-core k = 8."""
    import oracle
    t0 = time.time()
    n = 1 << scale
    edges = oracle.kron(scale, 0, 16 << scale)
    start, grid, degree = oracle.convert(edges, n)
    del edges
    E = grid.size
    log(f"cpu baseline graph s{scale} ready in {time.time() - t0:.1f}s, E={E}")
    roots = [int(r) for r in np.random.default_rng(20240601).choice(np.flatnonzero(degree), 2, replace=False)]
    deg64 = degree.astype(np.int64)
    per = {}
    iters = 5
    t = time.perf_counter()
    oracle.pagerank(n, start, grid, degree, iters, mode=1, nthreads=threads)
    per["pagerank"] = (time.perf_counter() - t, 2 * E * iters)
    t = time.perf_counter()
    m = 0
    for r in roots:
        d, _ = oracle.bfs(n, start, grid, r, nthreads=threads)
        m += int(deg64[d > 0].sum() // 2)
    per["bfs"] = (time.perf_counter() - t, m)
    t = time.perf_counter()
    oracle.wcc(n, start, grid, nthreads=threads)
    per["wcc"] = (time.perf_counter() - t, 2 * E)
    scanned = []
    t = time.perf_counter()
    oracle.kcore(n, start, grid, degree, 8, nthreads=threads, scanned=scanned)
    per["kcore"] = (time.perf_counter() - t, 2 * scanned[0])
    tot_s = sum(v[0] for v in per.values())
    tot_e = sum(v[1] for v in per.values())
    return {"value": round(tot_e / tot_s / 1e9, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "host_cpus_visible": visible,
            "per_algo": {k: {"s": round(v[0], 2), "gteps": round(v[1] / v[0] / 1e9, 4)} for k, v in per.items()},
            "sample": f"Kron-{scale} (E={E}): PageRank {iters} iterations (float degree factor), BFS from "
                      f"{len(roots)} roots, WCC, k-core k=8; the oracle's C/OpenMP threaded forms (tile sweeps "
                      f"with atomics like pr.cpp:60-80, kcore.cpp:69-103), {tot_s:.1f} s on {threads} threads"}


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    return round(xs[min(len(xs) - 1, int(q * len(xs)))], 3)


def bfs_saturation(ga, g, roots, workers, tile_ms_per_root=None, do_ms_per_root=None):
    """BFS throughput through the server on one graph: the same roots
    (submitted together) one root at a time on the single-root path
    (max_batch 1, one worker), in forced 64-lane batches, and under the
    server's own measured batching rule; beside them the direct single-root
    BFS per launch, tile sweep (the reference's method) and
    direction-optimizing (device time per root; measured here unless given)."""
    K = len(roots)
    sat = {"queries": K}
    for label, max_batch, nw, force in (("server_single_root", 1, 1, False),
                                        ("server_batched_64", 64, workers, True),
                                        ("server_auto", 64, workers, False)):
        with ga.Server(workers=nw) as srv:
            gid = srv.add_graph(g)
            srv.configure(max_batch=max_batch)
            if label == "server_auto":
                sat["policy"] = {k: round(v, 3) if isinstance(v, float) else v for k, v in srv.bfs_policy(gid).items()}
            if force:
                srv.set_batch_min(gid, 1)
            srv.submit(gid, ga.Q_BFS, int(roots[0]))     # warm the workspace
            srv.drain()
            srv.poll()
            t = time.perf_counter()
            for r in roots:
                srv.submit(gid, ga.Q_BFS, int(r))
            srv.drain()
            dt = time.perf_counter() - t
            res = srv.poll(K + 10)
        lat = [r["t_done_ms"] - r["t_submit_ms"] for r in res]
        sat[label] = {"qps": round(K / dt, 2), "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                      "mean_batch": round(float(np.mean([r["batch_size"] for r in res])), 2)}
    nd = min(K, 16)
    for label, mode, given in (("tile_single_root", 0, tile_ms_per_root), ("do_single_root", 1, do_ms_per_root)):
        if given is None:
            g.set_bfs_mode(mode)
            g.bfs(int(roots[0]), copy_out=False)
            ms = 0.0
            for r in roots[:nd]:
                g.bfs(int(r), copy_out=False)
                ms += g.stats()["ms"]
            given = ms / nd
        sat[label] = {"qps": round(1e3 / given, 2), "ms_per_root": round(given, 3)}
    b = sat["server_batched_64"]["qps"]
    best_single = max(sat["do_single_root"]["qps"], sat["server_single_root"]["qps"])
    sat["batched_vs_best_single_root"] = round(b / best_single, 2)
    sat["batched_vs_tile_single_root"] = round(b / sat["tile_single_root"]["qps"], 2)
    sat["server_single_vs_do_single"] = round(sat["server_single_root"]["qps"] / sat["do_single_root"]["qps"], 3)
    return sat


def bfs_poisson(ga, g, pool, rate, seconds, workers, rng):
    """BFS queries (uniform roots from `pool`) arriving as a Poisson process at
    `rate` per second for `seconds`, served with the server's own batching
    rule: achieved q/s and p50/p99 latency (submit to done)."""
    n_q = max(10, int(rate * seconds))
    gaps = rng.exponential(1.0 / rate, n_q)
    roots = rng.choice(pool, n_q, replace=True)
    with ga.Server(workers=workers) as srv:
        gid = srv.add_graph(g)
        srv.submit(gid, ga.Q_BFS, int(roots[0]))
        srv.drain()
        srv.poll()
        t0 = time.perf_counter()
        due = t0
        for r, gap in zip(roots, gaps):
            due += gap
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(gid, ga.Q_BFS, int(r))
        srv.drain()
        dt = time.perf_counter() - t0
        res = srv.poll(n_q + 10)
    lat = [r["t_done_ms"] - r["t_submit_ms"] for r in res]
    return {"offered_qps": rate, "queries": n_q, "achieved_qps": round(n_q / dt, 2), "p50_ms": pct(lat, 0.5),
            "p99_ms": pct(lat, 0.99), "mean_batch": round(float(np.mean([r["batch_size"] for r in res])), 2)}


def serve_mode(args):
    """Batched serving (Kronecker scales --serve-scales resident together): BFS
    saturation throughput per graph (batched, single root through the server,
    direct single roots), then Poisson arrivals of a 70/20/10 BFS/PPR/WCC mix
    at swept rates, p50/p99 latency per kind."""
    import gstore_amd as ga
    scales = [int(x) for x in args.serve_scales.split(",")]
    graphs = []
    for s in scales:
        t0 = time.time()
        graphs.append(ga.Graph.kron(s))
        log(f"serve: Kron-{s} resident ({time.time() - t0:.1f}s)")
    degs = [g.degree() for g in graphs]
    nz = [np.flatnonzero(d) for d in degs]
    rng = np.random.default_rng(4242)
    out = {"metric": "batched BFS queries/s, p50/p99 latency (serving)", "unit": "queries/s", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic (Graph500 Kronecker, seeds 1,2)",
           "config": {"workload": "serving", "scales": scales, "workers": args.workers}}
    sat_all = []
    for gi, g in enumerate(graphs):
        sat = {"graph_scale": scales[gi], **bfs_saturation(ga, g, rng.choice(nz[gi], args.sat_queries, replace=True),
                                                           args.workers)}
        g.set_bfs_mode(0)
        sat_all.append(sat)
        log(f"serve: Kron-{scales[gi]} batched {sat['server_batched_64']['qps']:.1f} q/s; "
            f"x{sat['batched_vs_best_single_root']} vs the best single root, "
            f"x{sat['batched_vs_tile_single_root']} vs the tile single root; policy {sat['policy']}")
    out["bfs_saturation"] = sat_all
    big = sat_all[-1]

    # 2. Poisson mix
    rates = [float(x) for x in args.serve_rates.split(",")]
    mix = []
    for lam in rates:
        n_q = max(10, int(lam * args.serve_seconds))
        gaps = rng.exponential(1.0 / lam, n_q)
        plan = []
        for i in range(n_q):
            gi = int(rng.integers(0, len(graphs)))
            u = rng.random()
            kind = ga.Q_BFS if u < 0.7 else ga.Q_PPR if u < 0.9 else ga.Q_WCC
            arg = int(rng.choice(nz[gi])) if kind == ga.Q_PPR else int(rng.integers(0, graphs[gi].n))
            plan.append((gi, kind, arg))
        with ga.Server(workers=args.workers) as srv:
            gids = [srv.add_graph(g) for g in graphs]
            t0 = time.perf_counter()
            due = t0
            for (gi, kind, arg), gap in zip(plan, gaps):
                due += gap
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                srv.submit(gids[gi], kind, arg)
            srv.drain()
            dt = time.perf_counter() - t0
            res = srv.poll(n_q + 10)
        by = {}
        for r in res:
            by.setdefault(r["kind"], []).append(r["t_done_ms"] - r["t_submit_ms"])
        names = {ga.Q_BFS: "bfs", ga.Q_PPR: "ppr", ga.Q_WCC: "wcc"}
        allat = [x for v in by.values() for x in v]
        mix.append({"offered_qps": lam, "queries": n_q, "achieved_qps": round(n_q / dt, 2),
                    "p50_ms": pct(allat, 0.5), "p99_ms": pct(allat, 0.99),
                    "per_kind": {names[k]: {"n": len(v), "p50_ms": pct(v, 0.5), "p99_ms": pct(v, 0.99)}
                                 for k, v in by.items()},
                    "bfs_mean_batch": round(float(np.mean([r["batch_size"] for r in res if r["kind"] == 0])), 2)})
        log(f"serve: rate {lam}/s -> {n_q / dt:.1f} q/s, p99 {mix[-1]['p99_ms']} ms")
    out["poisson_mix"] = mix
    out["value"] = big["server_batched_64"]["qps"]
    out["batched_vs_best_single_root"] = big["batched_vs_best_single_root"]
    out["batched_vs_tile_single_root"] = big["batched_vs_tile_single_root"]
    print(json.dumps(out), flush=True)
    for g in graphs:
        g.close()


def pinned_h2d_peak(gb=1):
    """Host-to-device copy rate of a pinned buffer (hipMemcpyAsync through
    torch), the roof of the semi-external stream: best of 5 copies of `gb` GiB."""
    import torch
    h = torch.empty(gb << 28, dtype=torch.float32, pin_memory=True)
    d = torch.empty(gb << 28, dtype=torch.float32, device="cuda")
    best = None
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        d.copy_(h, non_blocking=True)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del h, d
    torch.cuda.empty_cache()
    if hasattr(torch._C, "_host_emptyCache"):   # give the pinned block back before the engine pins the grid
        torch._C._host_emptyCache()
    return round((gb << 30) / (best / 1e3) / 1e9, 2)


def stream_mode(args):
    """Semi-external: Kron-S grid in pinned host memory (or, --external, left
    in a .grid file on disk) with the device tile budget capped; PageRank
    iterations and BFS roots streamed.  Reports the achieved host-to-device
    rate beside the measured pinned-copy peak, the compute/transfer overlap
    and, for the external mode, the file read rate."""
    # torch (the copy-peak probe) initialises the device before the engine does
    peak = pinned_h2d_peak()
    log(f"stream: pinned host-to-device copy peak {peak} GB/s")
    import gstore_amd as ga
    t0 = time.time()
    g = ga.Graph.kron(args.scale)
    log(f"stream: Kron-{args.scale} generated ({time.time() - t0:.1f}s), E={g.E}")
    t0 = time.time()
    ext_dir = None
    if args.external:
        ext_dir = args.external_dir or os.path.join(os.environ.get("TMPDIR", "/tmp"), f"gstore_ext_{os.getpid()}")
        os.makedirs(ext_dir, exist_ok=True)
        pre = os.path.join(ext_dir, f"kron{args.scale}")
        g.save(pre)
        n = g.n
        g.close()
        log(f"stream: tile files written to {ext_dir} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        g = ga.Graph.open_external(pre, n, args.budget_gb << 30, segment_bytes=args.segment_mb << 20)
        log(f"stream: opened from disk, resident edges {g.resident_edges} of {g.E} ({time.time() - t0:.1f}s), "
            f"io {g.io_stats()}")
    else:
        g.set_device_budget(args.budget_gb << 30, args.segment_mb << 20)
        log(f"stream: moved to pinned host, resident edges {g.resident_edges} of {g.E} ({time.time() - t0:.1f}s)")
    g.set_kernel_timing(True)
    res = {}
    io0 = g.io_stats()
    g.pagerank(args.stream_pr_iters, ga.PR_INTENDED, copy_out=False)
    st = g.stats()
    io1 = g.io_stats()
    k = st["kernels"]
    kern_ms = sum(v["ms"] for n_, v in k.items() if n_ != "h2d_copy")
    h2d_ms = k.get("h2d_copy", {}).get("ms", 0.0)
    exposed = max(0.0, st["ms"] - kern_ms)
    res["pagerank"] = {"iterations": args.stream_pr_iters, "ms": round(st["ms"], 1),
                       "ms_per_iter": round(st["ms"] / args.stream_pr_iters, 1),
                       "h2d_gb": round(st["h2d_bytes"] / 1e9, 2),
                       "h2d_gbs_wall": round(st["h2d_bytes"] / (st["ms"] / 1e3) / 1e9, 2),
                       "h2d_gbs_copy": round(st["h2d_bytes"] / (h2d_ms / 1e3) / 1e9, 2) if h2d_ms else None,
                       "kernel_ms": round(kern_ms, 1), "copy_ms": round(h2d_ms, 1),
                       "overlap_frac": round(1 - exposed / st["ms"], 4),
                       "copy_hidden_frac": round(1 - exposed / h2d_ms, 4) if h2d_ms else None,
                       "gteps": round(2 * g.E * args.stream_pr_iters / (st["ms"] / 1e3) / 1e9, 3)}
    res["pagerank"]["h2d_frac_of_pinned_peak"] = round(res["pagerank"]["h2d_gbs_wall"] / peak, 4)
    if args.external:
        db = io1["disk_bytes"] - io0["disk_bytes"]
        res["pagerank"].update(disk_gb=round(db / 1e9, 2), disk_gbs_wall=round(db / (st["ms"] / 1e3) / 1e9, 2),
                               read_wait_ms=round(io1["wait_ms"] - io0["wait_ms"], 1),
                               read_mode=["O_DIRECT into pinned", "O_DIRECT via bounce buffer",
                                          "buffered"][io1["read_mode"]])
    log(f"stream: pagerank {res['pagerank']}")
    deg = g.degree()
    roots = [int(r) for r in np.random.default_rng(5).choice(np.flatnonzero(deg), args.stream_bfs_roots, replace=False)]
    bfs = []
    for r in roots:
        io0 = g.io_stats()
        g.bfs(r, copy_out=False)
        st = g.stats()
        io1 = g.io_stats()
        e = {"root": r, "levels": st["iterations"], "ms": round(st["ms"], 1), "h2d_gb": round(st["h2d_bytes"] / 1e9, 2)}
        if args.external:
            e["disk_gb"] = round((io1["disk_bytes"] - io0["disk_bytes"]) / 1e9, 2)
        bfs.append(e)
        log(f"stream: bfs {e}")
    res["bfs"] = bfs
    mode = "external (grid read from disk every pass)" if args.external else "semi-external (grid in pinned host memory)"
    out = {"metric": "semi-external PageRank GTEPS, H2D GB/s and overlap", "value": res["pagerank"]["gteps"],
           "unit": "GTEPS", "n_gpus": 1, "higher_is_better": True,
           "data": "synthetic (Graph500 Kronecker, seeds 1,2)",
           "config": {"workload": f"Kron-{args.scale} {mode}, device tile budget {args.budget_gb} GB",
                      "scale": args.scale, "stored_edges": g.E, "resident_edges": g.resident_edges,
                      "segment_edges": g.segment_edges},
           "pinned_h2d_peak_gbs": peak, **res}
    print(json.dumps(out), flush=True)
    g.close()
    if ext_dir and not args.external_dir:
        for f in os.listdir(ext_dir):
            os.remove(os.path.join(ext_dir, f))
        os.rmdir(ext_dir)


def jobs_mode(args):
    """The remaining jobs on Kron-S resident: -j 30 asynchronous BFS from root
    1, -j 4 colour propagation from pivot 1 (device time, sweeps); and -c 4
    text to binary on the text of Kron-<text-scale>'s tuples: device parse
    rate (text bytes per second, the host-to-device copy of the text
    included) beside the oracle's serial parse of the same text on the host."""
    import gstore_amd as ga
    import oracle
    out = {"metric": "remaining jobs: -j 30 / -j 4 device ms, -c 4 text GB/s", "n_gpus": 1,
           "data": "synthetic (Graph500 Kronecker, seeds 1,2)", "config": {"workload": f"Kron-{args.scale} resident"}}
    with ga.Graph.kron(args.scale) as g:
        deg = g.degree().astype(np.int64)
        for label, fn in (("bfs_async", lambda: g.bfs_async(1)), ("traverse", lambda: g.traverse(1, 0))):
            fn()
            r = fn()
            st = g.stats()
            rec = {"ms": round(st["ms"], 2), "sweeps": st["iterations"]}
            if label == "bfs_async":
                d = r[0]
                m = int(deg[d > 0].sum() // 2)
                rec.update(reached=int(np.count_nonzero(d)), gteps_m_reach=round(m / (st["ms"] / 1e3) / 1e9, 2))
            else:
                rec.update(colored=r[2])
            out[label] = rec
            log(f"jobs: {label} {rec}")
    ts = args.text_scale
    e = oracle.kron(ts, 0, 16 << ts)
    text = b"% Kronecker tuples\n" + "\n".join(f"{a} {b}" for a, b in e).encode() + b"\n"
    del e
    tdir = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"gstore_txt_{os.getpid()}")
    os.makedirs(tdir, exist_ok=True)
    path = os.path.join(tdir, "edges.txt")
    with open(path, "wb") as f:
        f.write(text)
    ga.text_to_bin(text[:1 << 20])       # warm the device
    t = time.perf_counter()
    r = ga.text_file_to_bin(path)
    dev_s = time.perf_counter() - t
    t = time.perf_counter()
    want = oracle.text_to_bin(text)
    cpu_s = time.perf_counter() - t
    got = np.fromfile(path + ".edge", np.uint32).reshape(-1, 2)
    assert np.array_equal(got, want[0]), "text conversion differs from the oracle"
    out["text_to_bin"] = {"text_bytes": len(text), "pairs": r["pairs"],
                          "device_s": round(dev_s, 3), "device_gbs": round(len(text) / dev_s / 1e9, 2),
                          "host_serial_s": round(cpu_s, 3), "host_serial_gbs": round(len(text) / cpu_s / 1e9, 3),
                          "note": "device time includes reading the file and writing the .edge file"}
    log(f"jobs: text_to_bin {out['text_to_bin']}")
    for f_ in os.listdir(tdir):
        os.remove(os.path.join(tdir, f_))
    os.rmdir(tdir)
    print(json.dumps(out), flush=True)


def copy_mode(args):
    """Reference copy for counter calibration: 5 device-to-device copies of 1 GiB.
    Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE each dispatch should read and
    write 1 GiB; the ratio to what the counters report scales kernel traffic."""
    import torch
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / 5
    print(json.dumps({"metric": "device copy", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                      "ms": round(dt * 1e3, 3), "gbs": round(2 * (1 << 30) / dt / 1e9, 1)}), flush=True)


def _pmc_dispatches(path):
    """rocprofv3 --pmc counter CSV -> [(kernel name, value)] in dispatch order,
    the counter's values summed per dispatch (one counter per run)."""
    import csv
    per = {}
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        cols = {c.lower(): c for c in rd.fieldnames}
        kd, kn, kv = cols["dispatch_id"], cols["kernel_name"], cols["counter_value"]
        for row in rd:
            d = int(row[kd])
            name, v = per.get(d, (row[kn], 0.0))
            per[d] = (name, v + float(row[kv]))
    return [per[d] for d in sorted(per)]


def _short(name):
    base = name.split("(")[0].split("<")[0].split("::")[-1].strip()
    return base.replace("void ", "")


def traffic_mode(args):
    """HBM traffic per algorithm from two rocprofv3 --pmc runs (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch) of one bench invocation, scaled by the
    copy calibration (FETCH_SIZE reads half of a known copy's bytes: factor 2;
    WRITE_SIZE: 1).  The bench's step order splits the dispatches: PageRank
    iterations (the blocked kernels, per iteration), then BFS from the first
    k_do_root (direction-optimizing; one level per k_do_begin), then WCC from
    its first k_aff_/k_wcc_init kernel, then k-core from its first peel."""
    fetch, write = _pmc_dispatches(args.pmc_fetch), _pmc_dispatches(args.pmc_write)
    if [n for n, _ in fetch] != [n for n, _ in write]:
        sys.exit("the two runs dispatched different kernel sequences")
    ff, wf = args.fetch_factor, args.write_factor
    rows = [(_short(n), fv * 1024 * ff, wv * 1024 * wf) for (n, fv), (_, wv) in zip(fetch, write)]
    names = [r[0] for r in rows]
    out = {"workload": f"kron{args.scale}", "fetch_factor": ff, "write_factor": wf, "dispatches": len(rows)}

    def total(sel):
        f = sum(rows[i][1] for i in sel)
        w = sum(rows[i][2] for i in sel)
        by = {}
        for i in sel:
            b = by.setdefault(rows[i][0], [0, 0.0, 0.0])
            b[0] += 1
            b[1] += rows[i][1]
            b[2] += rows[i][2]
        top = sorted(by.items(), key=lambda kv: -(kv[1][1] + kv[1][2]))[:8]
        return {"fetch_bytes": int(f), "write_bytes": int(w),
                "kernels": {k: {"calls": v[0], "fetch_bytes": int(v[1]), "write_bytes": int(v[2])} for k, v in top}}

    pb = [i for i, n in enumerate(names) if n in ("k_pb_gather", "k_pb_accum", "k_pb_reduce")]
    iters = names.count("k_pb_gather")
    if iters:
        t = total(pb)
        out["pagerank"] = {"iterations": iters, "fetch_bytes": t["fetch_bytes"] // iters,
                           "write_bytes": t["write_bytes"] // iters, "kernels": t["kernels"]}

    def first(pred, lo=0):
        return next((i for i in range(lo, len(names)) if pred(names[i])), None)

    b0 = first(lambda n: n == "k_do_root")
    w0 = first(lambda n: n.startswith("k_aff_") or n == "k_wcc_init", (b0 or 0))
    k0 = first(lambda n: n in ("k_kc_peel_list", "k_kc_peel"), (w0 or b0 or 0))
    if b0 is not None:
        b1 = w0 if w0 is not None else (k0 if k0 is not None else len(rows))
        t = total(range(b0, b1))
        levels, cur = [], None
        for i in range(b0, b1):
            if names[i] == "k_do_begin":
                cur = [i]
                levels.append(cur)
            elif cur is not None:
                cur.append(i)
                if names[i] == "k_do_end":   # a level closes with k_do_end; what follows the last
                    cur = None               # one (unpermuting the depths) belongs to no level
        lv = []
        for li, sel in enumerate(levels):
            tl = total(sel)
            lv.append({"level": li + 1, "fetch_bytes": tl["fetch_bytes"], "write_bytes": tl["write_bytes"],
                       "top_kernel": next(iter(tl["kernels"]), None)})
        out["bfs"] = {"fetch_bytes": t["fetch_bytes"], "write_bytes": t["write_bytes"], "kernels": t["kernels"],
                      "levels": lv}
    if w0 is not None:
        t = total(range(w0, k0 if k0 is not None else len(rows)))
        out["wcc"] = t
    if k0 is not None:
        out["kcore"] = total(range(k0, len(rows)))
    print(json.dumps(out, indent=1), flush=True)


def level_method_bytes(n_act_words, lv):
    """Bytes one direction-optimizing BFS level touches by the engine's own
    counters: 4 per list entry examined (gs_level_stat.edges), 16 per frontier
    vertex on a top-down level (queue entry, degree, two offsets), the three
    bitmaps over the vertices with edges read and written once (12 bytes per
    word) and one depth byte per vertex reached.  An estimate of the
    algorithm's own bytes (random 4-byte reads priced at 4 bytes), beside the
    counter-measured traffic."""
    top = lv["direction"] == "top-down"
    return 4 * lv["edges"] + (16 * lv["frontier"] if top else 0) + 12 * n_act_words + lv["reached"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["main", "serve", "stream", "copy", "traffic", "jobs"], default="main")
    ap.add_argument("--text-scale", type=int, default=20, help="jobs mode: Kronecker scale of the -c 4 text")
    ap.add_argument("--serve-scales", default="18,20,22,24,26,28")
    ap.add_argument("--serve-rates", default="5,20,80")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--sat-queries", type=int, default=256)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--budget-gb", type=int, default=16)
    ap.add_argument("--segment-mb", type=int, default=512)
    ap.add_argument("--stream-pr-iters", type=int, default=10)
    ap.add_argument("--stream-bfs-roots", type=int, default=4)
    ap.add_argument("--external", action="store_true", help="stream mode: keep the grid in a file on disk")
    ap.add_argument("--external-dir", default=None, help="stream mode --external: where the tile files go")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=28)
    ap.add_argument("--pr-iters", type=int, default=20)
    ap.add_argument("--bfs-roots", type=int, default=16)
    ap.add_argument("--bfs-run-roots", type=int, default=0,
                    help="run the step's BFS from the first K of the drawn roots only (profiling runs); 0 = all")
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--cpu-scale", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=300.0, help="offered BFS queries/s of the serving entry")
    ap.add_argument("--algos", default="pr,bfs,wcc,kcore",
                    help="subset of the step's jobs (profiling runs); the default is the full step")
    ap.add_argument("--bfs-mode", type=int, default=-1, help="-1: autotune; 0: tile sweep; 1: direction-optimizing")
    ap.add_argument("--tune", type=int, default=1, help="run the sweep autotuner before timing")
    ap.add_argument("--pr-method", choices=["blocked", "pull", "scatter"], default="blocked",
                    help="PageRank: two streamed passes over the blocked index copy, pull over the "
                         "neighbour-list index, or the tile scatter with atomics")
    ap.add_argument("--pmc-fetch", help="traffic mode: counter CSV of the FETCH_SIZE run")
    ap.add_argument("--pmc-write", help="traffic mode: counter CSV of the WRITE_SIZE run")
    ap.add_argument("--fetch-factor", type=float, default=2.0)
    ap.add_argument("--write-factor", type=float, default=1.0)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path "
              "(tiles and index blocks would shard over a 2-D checkerboard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.mode == "serve":
        return serve_mode(args)
    if args.mode == "stream":
        return stream_mode(args)
    if args.mode == "copy":
        return copy_mode(args)
    if args.mode == "traffic":
        return traffic_mode(args)
    if args.mode == "jobs":
        return jobs_mode(args)

    import torch
    import gstore_amd as ga

    log(f"generating Kronecker scale {args.scale} on the device")
    t0 = time.time()
    g = ga.Graph.kron(args.scale)
    gen_s = time.time() - t0
    n, E = g.n, g.E
    log(f"graph ready in {gen_s:.1f}s: N={n} E={E} nonempty tiles={g.nonempty_tiles}")
    deg = g.degree()
    rng = np.random.default_rng(20240601)
    roots = [int(r) for r in rng.choice(np.flatnonzero(deg), args.bfs_roots, replace=False)]
    run_roots = roots[:args.bfs_run_roots] if args.bfs_run_roots > 0 else roots
    algos = set(args.algos.split(","))
    workload = f"kron{args.scale}"

    # Untimed pre-pass with the tile BFS (its scans follow the reference's
    # tile-skip rule): Graph500 M_reach and the scanned edges / tiles of each
    # root, for the fixed formula.  BFS output is deterministic, so these hold
    # for every BFS mode.  Its device time is the reference method's time per root.
    m_reach, ref_scan = {}, {}
    deg64 = deg.astype(np.int64)
    g.set_bfs_mode(0)
    ref_levels = None
    tile_ms = []
    for r in run_roots:
        d = g.bfs(r).depth
        st = g.stats()
        tile_ms.append(st["ms"])
        m_reach[r] = int(deg64[d > 0].sum() // 2)
        ref_scan[r] = (st["edges_scanned"], st["tiles_scanned"], st["iterations"], int(np.count_nonzero(d)))
        if ref_levels is None:
            ref_levels = g.level_stats()
    log("BFS M_reach and reference scan counts per root computed")
    ref_kc = None
    if "kcore" in algos:
        # the reference's tile scans of k-core (same rounds and degrees as the default method)
        g.set_kcore_method(0)
        g.kcore(args.k, copy_out=False)
        st = g.stats()
        ref_kc = (st["edges_scanned"], st["tiles_scanned"], st["iterations"])
        g.set_kcore_method(1)
        log(f"k-core reference scans: {ref_kc}")

    tune = None
    if args.tune:
        tune = g.tune()
        log(f"autotune: chunk {tune['chunk']} xcd_order {tune['xcd_order']} ({tune['best_ms']:.2f} ms of tile sweeps), "
            f"BFS alpha {tune['do_alpha']} beta {tune['do_beta']} ({tune['do_ms']:.2f} ms for two roots)")
    bfs_mode = args.bfs_mode
    if bfs_mode < 0 and "bfs" in algos:
        best = None
        for mode in (0, 1):
            g.set_bfs_mode(mode)
            g.bfs(roots[0], copy_out=False)
            g.bfs(roots[1], copy_out=False)
            ms = g.stats()["ms"]
            log(f"bfs mode {mode}: {ms:.2f} ms (root {roots[1]})")
            if best is None or ms < best[1]:
                best = (mode, ms)
        bfs_mode = best[0]
    g.set_bfs_mode(max(bfs_mode, 0))
    g.set_pagerank_method({"scatter": 0, "pull": 1, "blocked": 2}[args.pr_method])
    g.set_kernel_timing(True)
    kern = {}
    T = g.T
    # per algorithm: device ms, traversed edges (GTEPS convention), fixed reference-formula bytes
    per = {k: [0.0, 0.0, 0.0] for k in ("pagerank", "bfs", "wcc", "kcore")}
    bfs_scanned = [0.0]
    first_root_levels = []

    def absorb(st):
        for k, v in st["kernels"].items():
            a = kern.setdefault(k, [0.0, 0])
            a[0] += v["ms"]
            a[1] += v["calls"]

    def step(record):
        if "pr" in algos:
            g.pagerank(args.pr_iters, ga.PR_INTENDED, copy_out=False)
            st = g.stats()
            if record:
                absorb(st)
                per["pagerank"][0] += st["ms"]
                per["pagerank"][1] += 2 * E * args.pr_iters
                per["pagerank"][2] += args.pr_iters * alg_bytes_pr_iter(n, E, T)
        for r in (run_roots if "bfs" in algos else []):
            g.bfs(r, copy_out=False)
            st = g.stats()
            if record and r == roots[0] and not first_root_levels:
                first_root_levels.extend(g.level_stats())
            if record:
                absorb(st)
                e_scan, t_scan, levels, reached = ref_scan[r]
                per["bfs"][0] += st["ms"]
                per["bfs"][1] += m_reach[r]
                per["bfs"][2] += alg_bytes_bfs(n, e_scan, t_scan, levels, reached)
                bfs_scanned[0] += 2 * e_scan
        if "wcc" in algos:
            g.wcc(copy_out=False)
            st = g.stats()
            if record:
                absorb(st)
                passes = st["iterations"]
                jumps = st["kernels"].get("wcc_compress", {}).get("calls", 0)
                per["wcc"][0] += st["ms"]
                per["wcc"][1] += 2 * E
                per["wcc"][2] += alg_bytes_wcc(n, E, T, passes, jumps)
        if "kcore" not in algos:
            return
        g.kcore(args.k, copy_out=False)
        st = g.stats()
        if record:
            absorb(st)
            per["kcore"][0] += st["ms"]
            per["kcore"][1] += 2 * ref_kc[0]
            per["kcore"][2] += alg_bytes_kcore(n, *ref_kc)

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
    value = sum(v[1] for v in per.values()) / elapsed / 1e9

    def rate(x, ms):
        return x / (ms / 1e3) / 1e9 if ms > 0 else None

    idx = g.index_info()
    act_words = (idx["vertices"] + 31) // 32
    names = {"pagerank": "pr", "bfs": "bfs", "wcc": "wcc", "kcore": "kcore"}
    trf_keys = {"pagerank": args.pr_method, "bfs": "bfs_do" if bfs_mode == 1 else "bfs_tile",
                "wcc": "wcc_index", "kcore": "kcore_index"}
    per_algo = {}
    for algo, (ms, teps, ab) in per.items():
        if names[algo] not in algos:
            continue
        unit_ms = ms / args.steps
        d = {"ms": unit_ms, "gteps": rate(teps, ms), "ref_formula_gbs": rate(ab, ms)}
        if algo == "pagerank":
            unit_ms /= args.pr_iters
            d.update(iterations=args.pr_iters, method=args.pr_method,
                     gteps_convention="2E per iteration")
        if algo == "bfs":
            d.update(ms_per_root=ms / args.steps / len(run_roots), roots=len(run_roots),
                     mode=["tile sweep", "direction-optimizing"][max(bfs_mode, 0)],
                     gteps_convention="Graph500 M_reach per root", gteps_ref_scanned=rate(bfs_scanned[0], ms))
            unit_ms = sum(x["ms"] for x in first_root_levels)   # the profiled unit is the first root
            trf = measured_traffic(workload, trf_keys[algo], unit_ms)
            lv = []
            rec_lv = trf["rec"].get("levels", []) if trf else []
            for i, x in enumerate(first_root_levels):
                ref = (ref_levels or [])[i] if i < len(ref_levels or []) else None
                e = {"level": x["level"], "direction": x["direction"], "frontier": x["frontier"],
                     "edges_examined": x["edges"], "reached": x["reached"], "ms": round(x["ms"], 4),
                     "method_bytes_est": int(level_method_bytes(act_words, x))}
                if ref:
                    rb = 4 * ref["edges"] + 8 * ref["tiles"] + 3 * n / 8 + ref["reached"]
                    e["ref_formula_gbs"] = round(rb / (x["ms"] / 1e3) / 1e9, 1) if x["ms"] > 0 else None
                if i < len(rec_lv) and x["ms"] > 0:
                    lb = rec_lv[i]["fetch_bytes"] + rec_lv[i]["write_bytes"]
                    e["traffic_bytes"] = int(lb)
                    e["traffic_frac"] = round(lb / (x["ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                lv.append(e)
            d["levels_first_root"] = lv
            d["first_root_ms"] = round(unit_ms, 4)
            d["first_root_method_bytes_est"] = int(sum(e["method_bytes_est"] for e in lv))
        else:
            trf = measured_traffic(workload, trf_keys[algo], unit_ms)
        if algo == "wcc":
            d["gteps_convention"] = "2E per run"
        if algo == "kcore":
            d.update(k=args.k, gteps_convention="2 x stored edges scanned by the reference's tile sweeps")
        if trf:
            d["traffic_bytes"] = trf["bytes"]
            d["traffic_frac"] = trf["bytes"] / (unit_ms / 1e3) / 1e9 / HBM_PEAK_GBS
            d["traffic_unit"] = {"pagerank": "iteration", "bfs": "first root", "wcc": "run", "kcore": "run"}[algo]
            d["traffic_source"] = trf["source"]
        per_algo[algo] = {k2: round(v, 4) if isinstance(v, float) else v for k2, v in d.items()}

    chunk = tune["chunk"] if tune else 4096
    kernels = {name: {"total_ms": round(ms / args.steps, 3), "launches_per_step": calls / args.steps,
                      "avg_ms": round(ms / max(calls, 1), 4)} for name, (ms, calls) in kern.items()}
    # roofline: one PageRank iteration (all its kernels) against the fixed formula
    roofline = None
    if "pr" in algos and per["pagerank"][0] > 0:
        iters = args.pr_iters * args.steps
        it_kern = {k: kern[k][0] / iters for k in PR_ITER_KERNELS if k in kern}
        it_ms = sum(it_kern.values())
        b = alg_bytes_pr_iter(n, E, T)
        achieved = b / (it_ms / 1e3) / 1e9
        roofline = {"kernel": "pagerank iteration (" + " + ".join(it_kern) + ")", "bound": "hbm",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "bytes_per_launch": int(b), "avg_launch_ms": round(it_ms, 4),
                    "kernel_ms": {k: round(v, 4) for k, v in it_kern.items()}}
        trf = measured_traffic(workload, args.pr_method, it_ms)
        if trf:
            roofline["traffic"] = trf["bytes"]
            roofline["traffic_frac"] = round(trf["bytes"] / (it_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
            roofline["traffic_source"] = trf["source"]

    serving = None
    if not args.no_serving and "bfs" in algos:
        g.set_kernel_timing(False)
        srng = np.random.default_rng(777)
        pool = np.flatnonzero(deg)
        log("serving: BFS saturation and Poisson arrivals on the same graph")
        serving = bfs_saturation(ga, g, srng.choice(pool, 256, replace=True), args.workers,
                                 tile_ms_per_root=float(np.mean(tile_ms)),
                                 do_ms_per_root=per_algo["bfs"]["ms_per_root"] if bfs_mode == 1 else None)
        serving["poisson"] = bfs_poisson(ga, g, pool, args.serve_rate, 3.0, args.workers, srng)
        serving["graph_scale"] = args.scale
        log(f"serving: {serving}")

    cpu = None
    if not args.no_cpu_baseline:
        threads, visible = host_threads()
        log(f"cpu baseline on {threads} threads ({visible} CPUs visible)")
        cpu = cpu_baseline(args.cpu_scale, threads, visible)

    out = {
        "metric": "GTEPS + % HBM peak per algo at Kron-28; batched BFS queries/s, p50/p99 lat",
        "value": round(value, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+u32", "data": "synthetic (Graph500 Kronecker, seeds 1,2, generated on device)",
        "config": {"workload": f"Kron-{args.scale} ef16 undirected, resident; PR {args.pr_iters} it + "
                               f"BFS {len(run_roots)} roots + WCC + k-core k={args.k}",
                   "scale": args.scale, "vertices": n, "stored_edges": E, "nonempty_tiles": g.nonempty_tiles,
                   "gen_convert_s": round(gen_s, 2), "parallelism": "single GPU",
                   "index": {"entries": idx["entries"], "build_ms": round(idx["build_ms"], 1),
                             "device_gb": round(idx["device_bytes"] / 1e9, 1)},
                   "tuned": {"chunk": chunk, "xcd_order": tune["xcd_order"] if tune else 1,
                             "bfs_alpha": tune["do_alpha"] if tune else 14, "bfs_beta": tune["do_beta"] if tune else 24}},
        "per_algo": per_algo, "kernels": kernels, "roofline": roofline, "serving": serving, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
