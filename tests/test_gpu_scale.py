"""The benchmark graphs at full size against the oracle's fixtures.

tests/golden/scale_kronS.json (written by tests/golden/make_scale_fixtures.py
from the oracle) holds, for the undirected Kronecker graph of scale S: CRC32s
of the .start/.degree arrays and an order-free checksum of the grid; BFS level
counts and depth CRC32s from the benchmark's 16 roots and vertex 1; the WCC
count and label CRC32; k-core (k = 8, 4) counts, rounds and degree CRC32; the
PageRank compat CRC32 after 20 iterations; and, up to scale 26, FP64 PageRank
(intended form) and personalised PageRank values on a vertex sample.  The
graph is generated and converted on the device, every result is computed by
the HIP engine and compared bit for bit (integer results, compat PageRank) or
within the stated FP32 tolerance (1e-6 max-abs, 1e-4 relative, L1 reported).
Scale 25 is the plain GPU drop-in configuration, 28 the at-scale one; the
streamed (semi-external) path runs the same checks with a device budget below
the grid."""
import json
import os
import zlib

import numpy as np
import pytest

import gstore_amd as ga
import oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SCALES = [s for s in (24, 25, 26, 28) if os.path.exists(os.path.join(GOLDEN, f"scale_kron{s}.json"))]


def crc(a):
    return zlib.crc32(np.ascontiguousarray(a).view(np.uint8))


def fixture_of(s):
    with open(os.path.join(GOLDEN, f"scale_kron{s}.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module", params=SCALES, ids=[f"kron{s}" for s in SCALES])
def big(request):
    s = request.param
    print(f"scale test: generating Kron-{s} on the device", flush=True)
    g = ga.Graph.kron(s)
    yield s, g, fixture_of(s)
    g.close()


def test_files_match_oracle_conversion(big):
    s, g, fx = big
    assert (g.n, g.E, g.T, g.nonempty_tiles) == (fx["vertices"], fx["stored_edges"], fx["tile_slots"],
                                                 fx["nonempty_tiles"])
    start, grid, degree = g.arrays()
    assert crc(start) == fx["start_crc32"]
    assert crc(degree) == fx["degree_crc32"]
    assert oracle.grid_checksum(start, grid) == fx["grid_checksum"]


@pytest.mark.parametrize("mode", [1, 0], ids=["direction_optimizing", "tile"])
def test_bfs_roots(big, mode):
    s, g, fx = big
    g.set_bfs_mode(mode)
    for rec in fx["bfs"]:
        r = g.bfs(rec["root"])
        assert r.level_counts == rec["level_counts"], rec["root"]
        assert crc(r.depth) == rec["depth_crc32"], rec["root"]
    g.set_bfs_mode(0)


def test_bfs_batch_lanes(big):
    s, g, fx = big
    roots = [rec["root"] for rec in fx["bfs"]]
    depth, counts = g.bfs_batch(roots)
    for i, rec in enumerate(fx["bfs"]):
        assert crc(depth[i]) == rec["depth_crc32"], rec["root"]
        lc = rec["level_counts"]
        assert counts[i][:len(lc)] == lc and not any(counts[i][len(lc):])


@pytest.mark.parametrize("method", [1, 0], ids=["index", "tile"])
def test_wcc(big, method):
    s, g, fx = big
    g.set_wcc_method(method)
    lab, nc = g.wcc()
    g.set_wcc_method(1)
    assert nc == fx["wcc"]["ncomp"]
    assert crc(lab) == fx["wcc"]["label_crc32"]


@pytest.mark.parametrize("method", [1, 0], ids=["index", "tile"])
def test_kcore(big, method):
    s, g, fx = big
    g.set_kcore_method(method)
    for rec in fx["kcore"]:
        deg, ne0, pos, rounds = g.kcore(rec["k"])
        assert (ne0, pos, rounds) == (rec["count_ne0"], rec["count_pos"], rec["rounds"]), rec["k"]
        assert crc(deg) == rec["deg_crc32"], rec["k"]
    g.set_kcore_method(1)


@pytest.mark.parametrize("method", [2, 1], ids=["blocked", "pull"])
def test_pagerank_compat_bit_exact(big, method):
    s, g, fx = big
    g.set_pagerank_method(method)
    r = g.pagerank(20, ga.PR_COMPAT)
    g.set_pagerank_method(2)
    assert crc(r) == fx["pagerank_compat"]["crc32"]
    assert np.count_nonzero(r) == fx["pagerank_compat"]["nonzero"]


def _samples(s):
    p = os.path.join(GOLDEN, f"scale_kron{s}_samples.npz")
    return np.load(p) if os.path.exists(p) else None


def _fp32_vs_fp64(got, want, what):
    err = np.abs(got.astype(np.float64) - want)
    rel = err / np.maximum(np.abs(want), 1e-30)
    floor = want > 1e-3 * np.max(want)   # relative bound on the values that carry the vector
    stats = {"max_abs": float(err.max()), "l1_sample": float(err.sum()),
             "max_rel": float(rel[floor].max()), "n": int(want.size)}
    print(f"{what}: {stats}", flush=True)
    assert stats["max_abs"] <= 1e-6, stats
    assert stats["max_rel"] <= 1e-4, stats
    return stats


def test_pagerank_intended_vs_fp64(big):
    s, g, fx = big
    smp = _samples(s)
    if smp is None or "pagerank_intended64" not in fx:
        pytest.skip(f"no FP64 sample at Kron-{s} (the FP64 oracle is run up to scale 26)")
    idx = smp["idx"]
    r = g.pagerank(20, ga.PR_INTENDED)
    _fp32_vs_fp64(r[idx], smp["pr64"], f"Kron-{s} PageRank intended, sample")
    # the sum over all vertices follows the FP64 sum to FP32 accumulation error
    assert abs(float(r.astype(np.float64).sum()) - fx["pagerank_intended64"]["sum"]) <= 1e-6 * g.n


def test_ppr_vs_fp64(big):
    s, g, fx = big
    smp = _samples(s)
    if smp is None or "ppr64" not in fx:
        pytest.skip(f"no FP64 sample at Kron-{s}")
    idx = smp["idx"]
    r = g.ppr(fx["ppr64"]["source"], 20)
    st = _fp32_vs_fp64(r[idx], smp["ppr64"], f"Kron-{s} personalised PageRank, sample")
    assert st["l1_sample"] <= 1e-6
    assert abs(float(r.astype(np.float64).sum()) - fx["ppr64"]["sum"]) <= 1e-6


def test_streamed_matches_oracle(big):
    """Semi-external mode (grid in pinned host memory, device budget half the
    grid) gives the fixtures' results bit for bit."""
    s, g, fx = big
    with ga.Graph.kron(s) as h:
        h.set_device_budget(max(256 << 20, 2 * h.E), 64 << 20)
        assert not h.resident
        for rec in fx["bfs"][:4] + fx["bfs"][-1:]:
            r = h.bfs(rec["root"])
            assert r.level_counts == rec["level_counts"] and crc(r.depth) == rec["depth_crc32"], rec["root"]
            assert h.stats()["h2d_bytes"] > 0 or rec["reached"] == 1
        lab, nc = h.wcc()
        assert nc == fx["wcc"]["ncomp"] and crc(lab) == fx["wcc"]["label_crc32"]
        rec = fx["kcore"][0]
        deg, ne0, pos, rounds = h.kcore(rec["k"])
        assert (ne0, pos, rounds) == (rec["count_ne0"], rec["count_pos"], rec["rounds"])
        assert crc(deg) == rec["deg_crc32"]
        r = h.pagerank(20, ga.PR_COMPAT)
        assert crc(r) == fx["pagerank_compat"]["crc32"]


def test_config1_kron20_against_full_oracle():
    """The plain CPU reference configuration (Kron-20, edge factor 16): PageRank
    5 iterations, BFS from root 1, WCC, k-core k = 4 -- every output array
    against the oracle run on the host over the same tuples.  Compat PageRank
    and all integer outputs bit for bit; intended PageRank within 1e-6
    max-abs and 1e-4 relative of FP64, its L1 reported."""
    s, threads = 20, min(16, len(os.sched_getaffinity(0)))
    n = 1 << s
    start, grid, deg = oracle.convert(oracle.kron(s, 0, 16 << s), n)
    with ga.Graph.kron(s) as g:
        gs_, gg, gd = g.arrays()
        assert np.array_equal(gs_, start) and np.array_equal(gd, deg)
        assert oracle.grid_checksum(gs_, gg) == oracle.grid_checksum(start, grid)
        assert np.array_equal(g.pagerank(5, ga.PR_COMPAT), oracle.pagerank(n, start, grid, deg, 5, 0, nthreads=threads))
        got = g.pagerank(5, ga.PR_INTENDED)
        w64 = oracle.pagerank64(n, start, grid, deg, 5, nthreads=threads)
        d = np.abs(got.astype(np.float64) - w64)
        nz = w64 > 0
        print(f"Kron-20 intended PageRank vs FP64: max-abs {d.max():.3g}, max-rel {np.max(d[nz] / w64[nz]):.3g}, "
              f"L1 {d.sum():.3g}", flush=True)
        assert d.max() <= 1e-6 and np.max(d[nz] / w64[nz]) <= 1e-4
        want_d, want_c = oracle.bfs(n, start, grid, 1, nthreads=threads)
        for mode in (0, 1):
            g.set_bfs_mode(mode)
            r = g.bfs(1)
            assert np.array_equal(r.depth, want_d) and r.level_counts == want_c
        lab, nc = g.wcc()
        olab, onc = oracle.wcc(n, start, grid, nthreads=threads)
        assert np.array_equal(lab, olab) and nc == onc
        kd, ne0, pos, rounds = g.kcore(4)
        okd, one0, opos, orounds = oracle.kcore(n, start, grid, deg, 4)
        assert np.array_equal(kd, okd) and (ne0, pos, rounds) == (one0, opos, orounds)
