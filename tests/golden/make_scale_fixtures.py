"""Writes the scale fixtures: the oracle's results on the Kronecker graphs the
benchmark configurations name, reduced to counts, CRC32s of whole result
arrays and FP64 values on a vertex sample, so that a GPU test can compare the
engine at full size without the oracle re-running there.

    python tests/golden/make_scale_fixtures.py 25     # ~2 min on 8 cores
    python tests/golden/make_scale_fixtures.py 28     # ~1 h on 8 cores, 25 GB of RAM
    python tests/golden/make_scale_fixtures.py 28 fp64  # adds the FP64 samples to it

Graph: undirected Graph500 Kronecker, edge factor 16, seeds (1, 2), generated
and converted in chunks by the oracle (or_kron_convert_*).  Results:
  * files: CRC32 of .start and .degree bytes, order-free checksum of .grid
  * BFS from the 16 roots the benchmark draws (numpy default_rng(20240601),
    16 distinct vertices with degree > 0) and from vertex 1: exact per-level
    counts and CRC32 of the u8 depth array (bfs.cpp encoding)
  * WCC: component count and CRC32 of the min-id labels
  * k-core k = 8 (and k = 4): counts, rounds, CRC32 of the final degrees
  * PageRank compat, 20 iterations: CRC32 of the FP32 output (bit-exact at any
    summation order: only degree-1 vertices are non-zero and each has one term)
  * scale <= 26, or with the argument fp64 (which adds them to an existing
    fixture and recomputes nothing else): FP64 intended PageRank (20
    iterations) and FP64 personalised PageRank (20 iterations, one seeded
    source) on a vertex sample (the 1024 highest-degree vertices and 65536
    uniform ones), with the FP64 sums over all vertices, in
    scale_kronS_samples.npz
"""
import json
import os
import sys
import time
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

THREADS = min(8, os.cpu_count() or 1)


def crc(a: np.ndarray) -> int:
    return zlib.crc32(np.ascontiguousarray(a).view(np.uint8))


def bench_roots(degree: np.ndarray, k: int = 16):
    rng = np.random.default_rng(20240601)
    return [int(r) for r in rng.choice(np.flatnonzero(degree), k, replace=False)]


def sample_vertices(degree: np.ndarray):
    n = degree.size
    top = np.argsort(degree, kind="stable")[::-1][:1024]
    uni = np.random.default_rng(99).choice(n, 65536, replace=False)
    return np.unique(np.concatenate([top, uni])).astype(np.uint32)


def log(*a):
    print(f"[{time.strftime('%H:%M:%S')}]", *a, flush=True)


def fp64_samples(scale, n, start, grid, degree, fx, t0):
    idx = sample_vertices(degree)
    pr64 = oracle.pagerank64(n, start, grid, degree, 20, nthreads=THREADS)
    log(f"s{scale}: FP64 PageRank ({time.time() - t0:.0f}s)")
    src = int(np.random.default_rng(7).choice(np.flatnonzero(degree)))
    ppr64 = oracle.ppr(n, start, grid, degree, src, 20, nthreads=THREADS)
    np.savez_compressed(os.path.join(HERE, f"scale_kron{scale}_samples.npz"), idx=idx,
                        pr64=pr64[idx], ppr64=ppr64[idx])
    fx["pagerank_intended64"] = {"iterations": 20, "sum": float(pr64.sum()), "sample": int(idx.size)}
    fx["ppr64"] = {"iterations": 20, "source": src, "sum": float(ppr64.sum()), "sample": int(idx.size)}
    log(f"s{scale}: FP64 PageRank and PPR from {src} ({time.time() - t0:.0f}s)")


def main(scale: int, only_fp64: bool = False):
    t0 = time.time()
    n = 1 << scale
    start, grid, degree = oracle.kron_convert(scale)
    log(f"s{scale}: generated and converted, E={grid.size} ({time.time() - t0:.0f}s)")
    if only_fp64:
        path = os.path.join(HERE, f"scale_kron{scale}.json")
        with open(path) as f:
            fx = json.load(f)
        if crc(start) != fx["start_crc32"] or crc(degree) != fx["degree_crc32"]:
            raise SystemExit(f"s{scale}: graph differs from the fixture's")
        fp64_samples(scale, n, start, grid, degree, fx, t0)
        fx["oracle_seconds_fp64"] = round(time.time() - t0, 1)
        with open(path, "w") as f:
            json.dump(fx, f, indent=1)
        log(f"s{scale}: added the FP64 samples to scale_kron{scale}.json")
        return
    T = start.size - 1
    fx = {"scale": scale, "edge_factor": 16, "seeds": [1, 2], "vertices": n, "stored_edges": int(grid.size),
          "tile_slots": int(T), "nonempty_tiles": int(np.count_nonzero(np.diff(start))),
          "start_crc32": crc(start), "degree_crc32": crc(degree),
          "grid_checksum": oracle.grid_checksum(start, grid)}
    roots = bench_roots(degree) + [1]

    def one_bfs(r):
        d, lc = oracle.bfs(n, start, grid, r)
        return {"root": r, "level_counts": lc, "depth_crc32": crc(d), "reached": int(np.count_nonzero(d))}

    with ThreadPoolExecutor(THREADS) as ex:
        fx["bfs"] = list(ex.map(one_bfs, roots))
    log(f"s{scale}: BFS from {len(roots)} roots ({time.time() - t0:.0f}s)")
    lab, nc = oracle.wcc(n, start, grid)
    fx["wcc"] = {"ncomp": nc, "label_crc32": crc(lab)}
    del lab
    log(f"s{scale}: WCC count {nc} ({time.time() - t0:.0f}s)")
    fx["kcore"] = []
    for k in (8, 4):
        kd, ne0, pos, rounds = oracle.kcore(n, start, grid, degree, k)
        fx["kcore"].append({"k": k, "count_ne0": ne0, "count_pos": pos, "rounds": rounds, "deg_crc32": crc(kd)})
        log(f"s{scale}: k-core k={k}: {ne0} ({time.time() - t0:.0f}s)")
    del kd
    pr = oracle.pagerank(n, start, grid, degree, 20, mode=0, nthreads=THREADS)
    fx["pagerank_compat"] = {"iterations": 20, "crc32": crc(pr), "nonzero": int(np.count_nonzero(pr))}
    del pr
    log(f"s{scale}: PageRank compat ({time.time() - t0:.0f}s)")
    if scale <= 26:
        fp64_samples(scale, n, start, grid, degree, fx, t0)
    fx["oracle_seconds"] = round(time.time() - t0, 1)
    with open(os.path.join(HERE, f"scale_kron{scale}.json"), "w") as f:
        json.dump(fx, f, indent=1)
    log(f"s{scale}: wrote scale_kron{scale}.json")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 25, len(sys.argv) > 2 and sys.argv[2] == "fp64")
