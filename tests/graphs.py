"""Edge lists used by the tests (u32 tuples, as the -c 1 converter reads them)."""
import numpy as np

import oracle


def path(n=100):
    v = np.arange(n - 1, dtype=np.uint32)
    return n, np.stack([v, v + 1], 1)


def star_dups_loops():
    # hub 7 with leaves, a duplicated edge, self-loops (dropped by the converter)
    e = [(7, i) for i in range(8, 40)] + [(7, 8), (8, 7), (3, 3), (7, 7), (50, 51), (52, 52)]
    return 64, np.array(e, np.uint32)


def tile_crossing():
    # ids on both sides of the 2^16 tile boundaries, a long path across tiles
    n = 200_000
    rng = np.random.default_rng(7)
    a = rng.integers(0, n, 3000, dtype=np.uint64)
    b = rng.integers(0, n, 3000, dtype=np.uint64)
    extra = np.array([(65535, 65536), (65536, 131071), (131072, 0), (199_999, 65535), (1, 2), (2, 1)], np.uint64)
    e = np.concatenate([np.stack([a, b], 1), extra]).astype(np.uint32)
    return n, e


def two_groups():
    # N > 2^24: two physical groups, edges in both diagonal groups and the off-diagonal one
    g = 1 << 24
    n = g + 70_000
    rng = np.random.default_rng(11)
    lo = rng.integers(0, 5000, 2000)
    hi = g + rng.integers(0, 70_000, 2000)
    e = np.concatenate([
        np.stack([lo, hi], 1), np.stack([hi[::-1], lo], 1), np.stack([lo, lo[::-1]], 1),
        np.stack([hi, hi[::-1]], 1),
        np.array([(g - 1, g), (g, g + 65536), (0, n - 1), (g - 65536, g - 1)])]).astype(np.uint32)
    return n, e


def components():
    # isolated K2s, a triangle, a chain of 5, isolated vertices between them
    e = [(0, 1), (10, 11), (20, 21), (21, 22), (22, 20), (30, 31), (31, 32), (32, 33), (33, 34), (40, 41), (41, 40)]
    return 64, np.array(e, np.uint32)


def kron(scale, ef=16):
    return 1 << scale, oracle.kron(scale, 0, ef << scale)


HAND = {"path": path, "star": star_dups_loops, "tiles": tile_crossing, "components": components}


def directed_pr_check(got, n, start, grid, degree, iters):
    """Intended PageRank on a directed grid against the FP64 oracle, relative
    to max(1, rank).  Out-degree normalisation leaves ranks above 1 there,
    where the FP32 summation order alone moves a rank by a few ulps: on
    Kron-14 the FP32 oracle itself sits 9.9e-7 from the FP64 result.  The
    engine is held to twice the FP32 oracle's own distance from FP64 on the
    same graph, and never more loosely than 1e-6.  Returns (error, bound)."""
    w64 = oracle.pagerank64(n, start, grid, degree, iters, directed=True)
    f32 = oracle.pagerank(n, start, grid, degree, iters, 1, directed=True)
    scale = np.maximum(1.0, np.abs(w64))
    own = float(np.max(np.abs(f32.astype(np.float64) - w64) / scale))
    err = float(np.max(np.abs(got.astype(np.float64) - w64) / scale))
    bound = max(1e-6, 2 * own)
    assert err <= bound, (err, bound, own)
    return err, bound
