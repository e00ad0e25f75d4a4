"""Parity of the HIP engine against the CPU oracle, through the C ABI.

Integer outputs (BFS depth, level counts, WCC labels and count, k-core degrees
and counts, converted tile files, generated tuples) must be bit-exact.
PageRank compat mode is bit-exact (every non-zero value has a single-term
sum); intended mode accumulates FP32 in a different order than the oracle's
sequential sweep, so it is held to 1e-6 max-abs per vertex, the tolerance
the engine is specified to, and its L1 error against an FP64 recurrence is
reported.
"""
import os

import numpy as np
import pytest

import gstore_amd as ga
import oracle
from tests import graphs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PR_TOL = 1e-6


def _say(*a):
    print("[gpu-parity]", *a, flush=True)


def _graph_cases():
    out = []
    for name, fn in graphs.HAND.items():
        n, e = fn()
        out.append((name, n, e))
    out.append(("two_groups",) + graphs.two_groups())
    for s in (10, 14, 16):
        n, e = graphs.kron(s)
        out.append((f"kron{s}", n, e))
    return out


CASES = _graph_cases()
IDS = [c[0] for c in CASES]


@pytest.fixture(scope="module")
def loaded():
    """(name, n, arrays, Graph) for every case, the same arrays on both sides."""
    res = {}
    for name, n, e in CASES:
        _say("load", name, "N", n, "tuples", len(e))
        arrays = oracle.convert(e, n)
        res[name] = (n, arrays, ga.Graph.from_arrays(n, *arrays))
    yield res
    for _, _, g in res.values():
        g.close()


# ---- generator and converter ---------------------------------------------------

def test_generator_golden_vectors():
    vec = np.load(os.path.join(GOLDEN, "kron_ref_vectors.npz"))
    for key in vec.files:
        s, e0 = key[1:].split("_e")
        want = vec[key]
        _say("kron", key)
        assert np.array_equal(ga.kron_edges(int(s), int(e0), int(e0) + len(want)), want), key


def test_generator_matches_oracle_bulk():
    for s, e0, cnt in ((16, 0, 1 << 20), (22, 12345, 300000), (28, 4 << 28, 100000)):
        _say("kron bulk", s, e0, cnt)
        assert np.array_equal(ga.kron_edges(s, e0, e0 + cnt), oracle.kron(s, e0, e0 + cnt))


def _same_tiles(start, grid_a, grid_b):
    """Grids equal up to the order of edges inside each tile (the converter's
    in-tile order depends on atomic timing, as the reference's does)."""
    tid = np.repeat(np.arange(start.size - 1), np.diff(start.astype(np.int64)))
    ka = np.sort(tid.astype(np.uint64) << np.uint64(32) | grid_a.astype(np.uint64))
    kb = np.sort(tid.astype(np.uint64) << np.uint64(32) | grid_b.astype(np.uint64))
    return np.array_equal(ka, kb)


@pytest.mark.parametrize("directed", [False, True])
def test_convert_files_match_oracle(tmp_path, directed):
    for name, n, e in CASES[:6]:
        _say("convert", name, "directed", directed)
        pre = str(tmp_path / f"{name}_{int(directed)}")
        ga.convert(e, n, pre, directed)
        o_start, o_grid, o_deg = oracle.convert(e, n, directed)
        start = np.fromfile(pre + ".start", np.uint64)
        grid = np.fromfile(pre + ".grid", np.uint32)
        deg = np.fromfile(pre + ".degree", np.int32)
        assert np.array_equal(start, o_start) and np.array_equal(deg, o_deg)
        assert _same_tiles(start, grid, o_grid)
        with ga.Graph.open(pre, n, directed) as g:
            s2, g2, d2 = g.arrays()
            assert np.array_equal(s2, start) and np.array_equal(g2, grid) and np.array_equal(d2, deg)


def test_open_kron_matches_oracle_conversion():
    for s in (12, 17):
        _say("open_kron", s)
        with ga.Graph.kron(s) as g:
            start, grid, deg = g.arrays()
            o_start, o_grid, o_deg = oracle.convert(oracle.kron(s, 0, 16 << s), 1 << s)
            assert np.array_equal(start, o_start) and np.array_equal(deg, o_deg)
            assert _same_tiles(start, grid, o_grid)


# ---- algorithms ------------------------------------------------------------------

@pytest.mark.parametrize("name", IDS)
@pytest.mark.parametrize("iters", [1, 5, 20])
def test_pagerank_compat_bit_exact(loaded, name, iters):
    n, (start, grid, deg), g = loaded[name]
    _say("pr compat", name, iters)
    assert np.array_equal(g.pagerank(iters, ga.PR_COMPAT), oracle.pagerank(n, start, grid, deg, iters, 0))


@pytest.mark.parametrize("name", IDS)
def test_pagerank_intended_within_tolerance(loaded, name):
    n, (start, grid, deg), g = loaded[name]
    _say("pr intended", name)
    got = g.pagerank(20, ga.PR_INTENDED)
    want = oracle.pagerank(n, start, grid, deg, 20, 1)
    err = float(np.max(np.abs(got.astype(np.float64) - want)))
    _say("  max-abs vs oracle", err)
    assert err <= PR_TOL
    # against the FP64 recurrence: absolute (the spec), relative on every
    # non-zero value (ranks start at 1/N, so a 1e-6 absolute bound alone would
    # let a small rank be off by tens of percent) and L1, reported
    w64 = oracle.pagerank64(n, start, grid, deg, 20)
    d = np.abs(got.astype(np.float64) - w64)
    nz = w64 > 0
    rel = float(np.max(d[nz] / w64[nz])) if nz.any() else 0.0
    _say(f"  vs FP64: max-abs {d.max():.3g}, max-rel {rel:.3g}, L1 {d.sum():.3g}, L1/N {d.sum() / n:.3g}")
    assert d.max() <= PR_TOL
    assert rel <= 1e-4
    assert np.all(got[~nz] == 0)


@pytest.mark.parametrize("name", IDS)
def test_bfs_bit_exact(loaded, name):
    n, (start, grid, deg), g = loaded[name]
    roots = sorted({0, 1, n - 1, int(np.argmax(deg)), int(np.nonzero(deg)[0][len(np.nonzero(deg)[0]) // 2])})
    for root in roots:
        _say("bfs", name, root)
        res = g.bfs(root)
        d, lc = oracle.bfs(n, start, grid, root)
        assert np.array_equal(res.depth, d), (name, root)
        assert res.level_counts == lc


@pytest.mark.parametrize("name", IDS)
def test_wcc_bit_exact(loaded, name):
    n, (start, grid, deg), g = loaded[name]
    _say("wcc", name)
    lab, nc = g.wcc()
    olab, onc = oracle.wcc(n, start, grid)
    assert np.array_equal(lab, olab) and nc == onc


@pytest.mark.parametrize("n", [2000, 300000])
def test_wcc_long_first_entry_chains(n):
    """Paths: all inner vertices share a degree, so the first entries chain
    through the whole path and the first link round cuts its walks (the
    compress after it then runs); a second path with shuffled ids and a
    Kronecker part keep the labels and the count against the oracle."""
    _say("wcc long chains", n)
    rng = np.random.default_rng(n)
    m1, p1 = graphs.path(n)
    perm = rng.permutation(n).astype(np.uint32)
    p2 = perm[graphs.path(n)[1]] + np.uint32(n)
    m3, k3 = graphs.kron(10)
    e = np.concatenate([p1, p2, k3 + np.uint32(2 * n)])
    nn = 2 * n + m3
    start, grid, deg = oracle.convert(e, nn)
    with ga.Graph.from_arrays(nn, start, grid, deg) as g:
        g.set_wcc_method(1)
        lab, nc = g.wcc()
    olab, onc = oracle.wcc(nn, start, grid)
    assert np.array_equal(lab, olab) and nc == onc


@pytest.mark.parametrize("name", IDS)
@pytest.mark.parametrize("k", [2, 4, 8])
def test_kcore_bit_exact(loaded, name, k):
    n, (start, grid, deg), g = loaded[name]
    _say("kcore", name, k)
    kd, ne0, pos, rounds = g.kcore(k)
    okd, one0, opos, orounds = oracle.kcore(n, start, grid, deg, k)
    assert np.array_equal(kd, okd) and (ne0, pos, rounds) == (one0, opos, orounds)


def test_repeat_calls_identical(loaded):
    n, (start, grid, deg), g = loaded["kron14"]
    _say("repeat")
    a = g.bfs(3).depth.copy()
    g.pagerank(3, ga.PR_INTENDED)
    g.kcore(4)
    assert np.array_equal(g.bfs(3).depth, a)
    assert np.array_equal(g.wcc()[0], g.wcc()[0])


# ---- directed grids (gstored) ------------------------------------------------------

@pytest.mark.parametrize("name", ["tiles", "components", "kron14"])
def test_directed_parity(name):
    n, e = dict((c[0], (c[1], c[2])) for c in CASES)[name]
    start, grid, deg = oracle.convert(e, n, directed=True)
    _say("directed", name)
    with ga.Graph.from_arrays(n, start, grid, deg, directed=True) as g:
        # directed compat ranks sum several in-neighbours (FP32 order-dependent): tolerance, not bit-exact
        got = g.pagerank(5, ga.PR_COMPAT)
        want = oracle.pagerank(n, start, grid, deg, 5, 0, directed=True)
        assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= PR_TOL
        # out-degree normalisation leaves ranks above 1 on a directed grid: against FP64
        err, bound = graphs.directed_pr_check(g.pagerank(5, ga.PR_INTENDED), n, start, grid, deg, 5)
        _say(f"  intended vs FP64: {err:.3g} (bound {bound:.3g})")
        for root in (0, 1, int(np.argmax(deg))):
            assert np.array_equal(g.bfs(root).depth, oracle.bfs(n, start, grid, root, directed=True)[0])
        lab, nc = g.wcc()
        olab, onc = oracle.wcc(n, start, grid, directed=True)
        assert np.array_equal(lab, olab) and nc == onc
        with pytest.raises(ga.GStoreError) as ei:
            g.kcore(4)
        assert ei.value.status == 6


# ---- error paths -------------------------------------------------------------------

@pytest.mark.parametrize("nth", [1, 6, 8, 10])
def test_index_build_failure_releases_and_rebuilds(nth):
    """An allocation failure part way through the index build (the 1st, 6th,
    8th or last of its 10 allocations) returns GS_ERR_NOMEM, leaves no index
    behind, and the next call builds it again and answers correctly."""
    n, e = graphs.kron(14)
    start, grid, deg = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        base = g.index_info()["device_bytes"]
        g.debug_fail_alloc(nth)
        with pytest.raises(ga.GStoreError) as ei:
            g.set_bfs_mode(1)          # builds the index
        assert ei.value.status == 4
        info = g.index_info()
        assert info["entries"] == 0 and info["device_bytes"] <= base + 64 * n
        want, lc = oracle.bfs(n, start, grid, 0)
        got = g.bfs(0)
        assert np.array_equal(got.depth, want) and got.level_counts == lc
        assert g.index_info()["entries"] == 2 * g.E
        lab, nc = g.wcc()
        assert np.array_equal(lab, oracle.wcc(n, start, grid)[0])


def test_errors(tmp_path):
    _say("errors")
    n, e = graphs.path(300)
    start, grid, deg = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        with pytest.raises(ga.GStoreError) as ei:
            g.bfs(0)
        assert ei.value.status == 7          # depth would pass 255
        assert g.bfs(150).depth.max() == 151
        with pytest.raises(ga.GStoreError):
            g.bfs(n)                         # root out of range
        g.set_bfs_mode(1, 14, 24)            # the device-driven levels stop at the same place
        with pytest.raises(ga.GStoreError) as ei:
            g.bfs(0)
        assert ei.value.status == 7
        res = g.bfs(150)
        d, lc = oracle.bfs(n, start, grid, 150)
        assert np.array_equal(res.depth, d) and res.level_counts == lc
        g.set_bfs_mode(0, 14, 24)
    bad = deg.copy()
    bad[5] += 1
    with pytest.raises(ga.GStoreError) as ei:
        ga.Graph.from_arrays(n, start, grid, bad)
    assert ei.value.status == 3
    pre = str(tmp_path / "p")
    ga.convert(e, n, pre)
    with pytest.raises(ga.GStoreError) as ei:
        ga.Graph.open(pre, n + (1 << 24))    # wrong N: .start size differs
    assert ei.value.status == 3
    with open(pre + ".grid", "r+b") as f:
        f.truncate(8)
    with pytest.raises(ga.GStoreError) as ei:
        ga.Graph.open(pre, n)
    assert ei.value.status == 3
    with pytest.raises(ga.GStoreError) as ei:
        ga.Graph.open(str(tmp_path / "missing"), n)
    assert ei.value.status == 2
    with pytest.raises(ga.GStoreError) as ei:
        ga.convert(np.array([[0, 400]], np.uint32), n, pre)
    assert ei.value.status == 1


# ---- size-independent properties at a larger scale -----------------------------------

def test_properties_scale20():
    s = 20
    _say("properties s", s)
    with ga.Graph.kron(s) as g:
        start, grid, deg = g.arrays()
        from gstore_amd import tiles
        v0, v1 = tiles.decode(g.n, start, grid)
        v0 = v0.astype(np.int64)
        v1 = v1.astype(np.int64)
        root = int(np.argmax(deg))
        d = g.bfs(root).depth.astype(np.int64)
        both = (d[v0] > 0) & (d[v1] > 0)
        assert np.all(np.abs(d[v0] - d[v1])[both] <= 1)                  # BFS layering
        assert np.all((d[v0] > 0) == (d[v1] > 0))                          # no edge leaves the reached set
        reached = np.nonzero(d > 1)[0]
        has_parent = np.zeros(g.n, bool)
        has_parent[v1[d[v0] == d[v1] - 1]] = True
        has_parent[v0[d[v1] == d[v0] - 1]] = True
        assert np.all(has_parent[reached])
        lab, nc = g.wcc()
        assert np.all(lab[v0] == lab[v1])
        nz = deg > 0
        assert np.all(lab[nz] <= np.nonzero(nz)[0]) and np.all(lab[lab[nz]] == lab[nz])
        assert nc == int(np.sum(lab[nz] == np.nonzero(nz)[0]))
        assert np.all((d > 0)[nz] == (lab[nz] == lab[root]))              # BFS reach == root's component
        r = g.pagerank(5, ga.PR_COMPAT)
        assert np.all(r[deg != 1] == 0)
        kd, ne0, pos, _ = g.kcore(8)
        core = kd > 0
        cnt = np.bincount(v0[core[v0] & core[v1]], minlength=g.n) + np.bincount(v1[core[v0] & core[v1]], minlength=g.n)
        plain = core & (deg < 32768)   # the (int16) test leaves larger degrees to the quirk
        assert np.all(cnt[plain] >= 8) and pos == int(core.sum())          # core vertices keep >= k core edges


# ---- autotuner and sweep knobs --------------------------------------------------------

def test_knobs_do_not_change_results(loaded, tmp_path):
    n, (start, grid, deg), g = loaded["kron16"]
    want_pr = oracle.pagerank(n, start, grid, deg, 5, 0)
    want_bfs = oracle.bfs(n, start, grid, 7)[0]
    for chunk in (1024, 2048, 4096):
        for order in (0, 1):
            _say("knobs", chunk, order)
            g.set_chunk(chunk)
            g.set_sweep_order(bool(order))
            assert np.array_equal(g.pagerank(5, ga.PR_COMPAT), want_pr)
            assert np.array_equal(g.bfs(7).depth, want_bfs)
    cache = str(tmp_path / "tune.txt")
    r1 = g.tune(cache)
    _say("tune", r1)
    assert not r1["from_cache"] and len(r1["candidates"]) == 6
    r2 = g.tune(cache)
    assert r2["from_cache"] and (r2["chunk"], r2["xcd_order"]) == (r1["chunk"], r1["xcd_order"])
    assert np.array_equal(g.pagerank(5, ga.PR_COMPAT), want_pr)
    with pytest.raises(ga.GStoreError):
        g.set_chunk(1000)
    g.set_chunk(4096)
    g.set_sweep_order(True)


@pytest.mark.parametrize("name", [c for c in IDS if c != "two_groups"] + ["two_groups"])
def test_bfs_direction_optimizing_bit_exact(loaded, name):
    n, (start, grid, deg), g = loaded[name]
    roots = sorted({0, 1, n - 1, int(np.argmax(deg))})
    for alpha, beta in ((14, 24), (1, 1000000), (1000000, 1)):   # default, always bottom-up-ish, always top-down
        g.set_bfs_mode(1, alpha, beta)
        for root in roots:
            _say("do-bfs", name, root, alpha, beta)
            res = g.bfs(root)
            d, lc = oracle.bfs(n, start, grid, root)
            assert np.array_equal(res.depth, d), (name, root, alpha, beta)
            assert res.level_counts == lc
    g.set_bfs_mode(0, 14, 24)


@pytest.mark.parametrize("name", ["kron16", "two_groups", "star", "tiles"])
def test_pagerank_methods_same_results(loaded, name):
    """Tile scatter (plain atomics or heavy-vertex LDS table), the pull over
    neighbour lists and the blocked two-pass method change only the FP32
    summation order."""
    n, (start, grid, deg), g = loaded[name]
    want_c = oracle.pagerank(n, start, grid, deg, 10, 0)
    want_i = oracle.pagerank(n, start, grid, deg, 10, 1)
    src = int(np.argmax(deg))
    want_p = oracle.ppr(n, start, grid, deg, src, 20)
    for method, heavy in ((ga.PR_SCATTER, False), (ga.PR_SCATTER, True), (ga.PR_PULL, True),
                          (ga.PR_BLOCKED, True)):
        g.set_pagerank_method(method)
        g.set_pagerank_heavy(heavy)
        _say("pr method", name, "method", method, "heavy", heavy)
        assert np.array_equal(g.pagerank(10, ga.PR_COMPAT), want_c)
        assert np.max(np.abs(g.pagerank(10, ga.PR_INTENDED) - want_i)) <= PR_TOL
        assert np.array_equal(g.pagerank(0, ga.PR_COMPAT), np.full(n, np.float32(1.0 / n)))
        p = g.ppr(src, 20)
        assert p.min() >= 0 and float(np.abs(p - want_p).sum()) <= 1e-6
    g.set_pagerank_method(ga.PR_BLOCKED)
    g.set_pagerank_heavy(True)


@pytest.mark.parametrize("s", [16, 18])
def test_pagerank_blocked_split_paths(s):
    """Blocked PageRank with tiny work limits, so that gather units split runs
    and cut at the run limit, and destination blocks split into partial sums
    (Kron-18 also has several 2^15-vertex blocks): same results as the oracle."""
    n, e = graphs.kron(s)
    start, grid, deg = oracle.convert(e, n)
    want_c = oracle.pagerank(n, start, grid, deg, 5, 0)
    want_i = oracle.pagerank(n, start, grid, deg, 5, 1)
    src = int(np.argmax(deg))
    want_p = oracle.ppr(n, start, grid, deg, src, 20)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        for lim in ((1 << 20, 2048, 1 << 20, 64, 500), (97, 3, 5000, 0, 1000), (1000, 1, 1 << 12, 1 << 30, 0),
                    (4096, 1024, 1 << 30, 8, 1000), (160, 2, 3000, 1 << 30, 1000), (1 << 20, 1024, 1 << 20, 8192, 500)):
            _say("blocked limits", s, lim)
            g.set_pagerank_blocking(*lim)
            assert np.array_equal(g.pagerank(5, ga.PR_COMPAT), want_c)
            assert np.max(np.abs(g.pagerank(5, ga.PR_INTENDED) - want_i)) <= PR_TOL
            p = g.ppr(src, 20)
            assert p.min() >= 0 and float(np.abs(p - want_p).sum()) <= 1e-6
        with pytest.raises(ga.GStoreError):
            g.set_pagerank_blocking(0, 1, 1, 64)
        with pytest.raises(ga.GStoreError):
            g.set_pagerank_blocking((1 << 20) + 1, 1, 1, 64)
        with pytest.raises(ga.GStoreError):
            g.set_pagerank_blocking(1, 2049, 1, 64)
        with pytest.raises(ga.GStoreError):
            g.set_pagerank_blocking(1, 1, 1, 64, 1001)
        with pytest.raises(ga.GStoreError):
            g.set_pagerank_method(3)


def _scan_counts(n, start, grid, depth):
    """Per BFS level, edges and non-empty tiles in tile rows / columns holding a
    frontier vertex (the reference's skip rule, bfs.cpp:290-312), summed."""
    from gstore_amd import tiles
    v0, v1 = tiles.decode(n, start, grid)
    tr, tc = (v0 >> 16).astype(np.int64), (v1 >> 16).astype(np.int64)
    side = int(max(tr.max(initial=0), tc.max(initial=0))) + 1
    tile, cnt = np.unique(tr * side + tc, return_counts=True)
    r, c = tile // side, tile % side
    e_scan = t_scan = 0
    for lev in range(1, int(depth.max()) + 1):
        act = np.zeros(side, bool)
        f = np.nonzero(depth == lev)[0] >> 16
        act[f[f < side]] = True
        m = act[r] | act[c]
        e_scan += int(cnt[m].sum())
        t_scan += int(m.sum())
    return e_scan, t_scan


@pytest.mark.parametrize("name", ["kron14", "tiles", "two_groups"])
def test_bfs_scan_counts_follow_skip_rule(loaded, name):
    """The tile BFS reports the edges and tiles its levels scan, which the bench
    prices with the fixed byte formula: they follow the reference's skip rule."""
    n, (start, grid, deg), g = loaded[name]
    g.set_bfs_mode(0)
    root = int(np.argmax(deg))
    d = g.bfs(root).depth
    st = g.stats()
    assert (st["edges_scanned"], st["tiles_scanned"]) == _scan_counts(n, start, grid, d)
    assert st["iterations"] == int(d.max())


@pytest.mark.parametrize("name", IDS)
def test_wcc_methods_identical(loaded, name):
    """Sampled linking over the neighbour-list index and hook passes over the
    tiles give the same canonical labels and count as the oracle."""
    n, (start, grid, deg), g = loaded[name]
    want_l, want_c = oracle.wcc(n, start, grid)
    for method in (0, 1):
        g.set_wcc_method(method)
        lab, nc = g.wcc()
        _say("wcc method", name, method, nc)
        assert nc == want_c and np.array_equal(lab, want_l)
    g.set_wcc_method(1)
    with pytest.raises(ga.GStoreError):
        g.set_wcc_method(2)


@pytest.mark.parametrize("name", [c for c in IDS if c != "two_groups"] + ["two_groups"])
@pytest.mark.parametrize("k", [2, 8, 40])
def test_kcore_methods_identical(loaded, name, k):
    """Peeled lists pushed over the index and tile rescans give the oracle's
    final degrees, counts and rounds."""
    n, (start, grid, deg), g = loaded[name]
    want = oracle.kcore(n, start, grid, deg, k)
    for method in (0, 1):
        g.set_kcore_method(method)
        got = g.kcore(k)
        _say("kcore method", name, k, method, got[1:])
        assert np.array_equal(got[0], want[0]) and tuple(got[1:]) == tuple(want[1:])
    g.set_kcore_method(1)
    with pytest.raises(ga.GStoreError):
        g.set_kcore_method(2)


@pytest.mark.parametrize("name", ["kron14", "tiles"])
def test_bfs_level_stats(loaded, name):
    """Per-level records: one per level, reached = the level counts, the first
    frontier is the root; tile levels add up to the scan totals."""
    n, (start, grid, deg), g = loaded[name]
    root = int(np.argmax(deg))
    want = oracle.bfs(n, start, grid, root)[1]
    for mode in (0, 1):
        g.set_bfs_mode(mode)
        res = g.bfs(root)
        st = g.stats()
        lv = g.level_stats()
        _say("levels", name, mode, [(x["direction"], x["frontier"], x["edges"]) for x in lv])
        assert len(lv) == st["iterations"] == len(want)
        assert [x["reached"] for x in lv] == want == res.level_counts
        assert lv[0]["frontier"] == 1 and all(x["ms"] >= 0 for x in lv)
        assert [x["frontier"] for x in lv[1:]] == want[:-1]
        if mode == 0:
            assert sum(x["edges"] for x in lv) == st["edges_scanned"]
            assert sum(x["tiles"] for x in lv) == st["tiles_scanned"]
            assert all(x["direction"] == "tile" for x in lv)
        else:
            assert all(x["direction"] in ("top-down", "bottom-up") for x in lv)
    g.set_bfs_mode(0)
