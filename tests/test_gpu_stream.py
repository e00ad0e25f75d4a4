"""Semi-external mode: the grid in pinned host memory, a resident prefix in
HBM and the rest streamed through two or three segment buffers.  Every algorithm must
give the resident result (integer outputs bit for bit, PageRank compat bit for
bit, intended within FP32 tolerance)."""
import numpy as np
import pytest

import gstore_amd as ga
import oracle
from tests import graphs

pytestmark = pytest.mark.gpu
CH = 4096 * 4   # bytes per chunk


def _say(*a):
    print("[gpu-stream]", *a, flush=True)


@pytest.mark.parametrize("res_chunks", [0, 5, 12])
def test_streaming_equals_resident(res_chunks):
    """Two segment buffers while the budget holds fewer than four segments,
    three from there on (res_chunks 12: the copy engine runs two segments
    ahead of the sweep)."""
    n, e = graphs.kron(15)
    start, grid, deg = oracle.convert(e, n)
    seg = 3 * CH
    budget = 2 * seg + res_chunks * CH + CH
    nbuf = 3 if budget >= 4 * seg + CH else 2
    with ga.Graph.from_arrays(n, start, grid, deg) as ref, ga.Graph.from_arrays(n, start, grid, deg) as g:
        g.set_device_budget(budget, seg)
        assert not g.resident and g.resident_edges == min(g.E, (budget - nbuf * seg) // CH * 4096)
        _say("streaming, resident edges", g.resident_edges, "of", g.E)
        assert np.array_equal(g.pagerank(5, ga.PR_COMPAT), ref.pagerank(5, ga.PR_COMPAT))
        assert np.max(np.abs(g.pagerank(10, ga.PR_INTENDED) - ref.pagerank(10, ga.PR_INTENDED))) <= 1e-6
        st = g.stats()
        assert st["h2d_bytes"] >= 10 * 4 * (g.E - g.resident_edges)
        for root in (0, int(np.argmax(deg))):
            assert np.array_equal(g.bfs(root).depth, ref.bfs(root).depth)
        assert np.array_equal(g.bfs_batch([1, 2, 3])[0], ref.bfs_batch([1, 2, 3])[0])
        assert np.array_equal(g.wcc()[0], ref.wcc()[0])
        a, b = g.kcore(8), ref.kcore(8)
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]
        assert np.max(np.abs(g.ppr(5) - ref.ppr(5))) <= 1e-6
        s2, g2, d2 = g.arrays()
        assert np.array_equal(s2, start) and np.array_equal(g2, grid)


@pytest.mark.parametrize("directed", [False, True], ids=["undirected", "directed"])
@pytest.mark.parametrize("method", [2, 0], ids=["binned", "atomic"])
def test_streamed_pagerank_methods_match_oracle(directed, method):
    """Semi-external PageRank by the binned scatter (default: contributions
    binned per tile row and column, summed in LDS, no global atomics) and by
    the tile scatter with atomics (method 0), against the oracle: compat bit
    for bit (undirected), intended within FP32 tolerance, personalised
    PageRank within 1e-6 L1 of the FP64 oracle.  Segments of 3 chunks and a
    resident prefix of 7 chunks, so the binned scatter cuts the prefix into
    several pieces and tiles span pieces and chunks."""
    n, e = graphs.kron(16)
    start, grid, deg = oracle.convert(e, n, directed=directed)
    seg = 3 * CH
    with ga.Graph.from_arrays(n, start, grid, deg, directed=directed) as g:
        g.set_device_budget(2 * seg + 7 * CH, seg)
        g.set_pagerank_method(method)
        assert not g.resident
        _say("streamed pagerank", "directed" if directed else "undirected", "method", method)
        got = g.pagerank(5, ga.PR_COMPAT)
        want = oracle.pagerank(n, start, grid, deg, 5, 0, directed=directed)
        if directed:   # several in-neighbours' ranks: FP32 order-dependent
            assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-6
        else:
            assert np.array_equal(got, want)
        got = g.pagerank(10, ga.PR_INTENDED)
        if directed:   # ranks above 1: against FP64, as close as FP32 allows
            graphs.directed_pr_check(got, n, start, grid, deg, 10)
        else:
            want = oracle.pagerank(n, start, grid, deg, 10, 1)
            assert np.max(np.abs(got - want)) <= 1e-6
        src = int(np.argmax(deg))
        p = g.ppr(src, 20)
        assert float(np.abs(p - oracle.ppr(n, start, grid, deg, src, 20, directed=directed)).sum()) <= 1e-6
        assert g.stats()["h2d_bytes"] > 0


@pytest.mark.parametrize("method", [2, 0], ids=["binned", "atomic"])
def test_directed_resident_pagerank_methods_match_oracle(method):
    """Resident directed grids (gstored) also take the binned scatter by
    default; both methods against the oracle, on a graph of four tile rows,
    so every bin gathers row and column contributions of several tiles.  The
    hubs here sum ~10^4 FP32 in-contributions in an order that differs from
    the oracle's, so the bound is 1e-5 relative to max(1, rank) (both methods
    stay within 3e-6 in practice)."""
    n, e = graphs.kron(18)
    start, grid, deg = oracle.convert(e, n, directed=True)
    with ga.Graph.from_arrays(n, start, grid, deg, directed=True) as g:
        g.set_pagerank_method(method)
        for mode, iters in ((ga.PR_COMPAT, 5), (ga.PR_INTENDED, 10)):
            got = g.pagerank(iters, mode)
            want = oracle.pagerank(n, start, grid, deg, iters, mode, directed=True)
            err = float(np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))))
            _say("directed resident method", method, "mode", mode, "max rel err", err)
            assert err <= 1e-5


def test_sparse_levels_read_host_grid_in_place():
    """Semi-external BFS levels whose frontier sits in at most 1/16 of the tile
    rows sweep the pinned host grid in place instead of copying segments: the
    same depths, level counts and scan counts as resident, and fewer bytes
    copied than one full stream per level."""
    n, e = graphs.kron(20)   # 16 tile rows: a level with one frontier row reads in place
    start, grid, deg = oracle.convert(e, n)
    seg = 64 * CH
    with ga.Graph.from_arrays(n, start, grid, deg) as ref, ga.Graph.from_arrays(n, start, grid, deg) as g:
        g.set_device_budget(2 * seg + CH, seg)
        assert not g.resident
        per_stream = 4 * (g.E - g.resident_edges)
        for root in (int(np.argmax(deg)), int(np.flatnonzero(deg)[-1]), 1):
            _say("direct sweeps, root", root)
            got, want = g.bfs(root), ref.bfs(root)
            assert np.array_equal(got.depth, want.depth) and got.level_counts == want.level_counts
            st, rs = g.stats(), ref.stats()
            assert st["edges_scanned"] == rs["edges_scanned"] and st["tiles_scanned"] == rs["tiles_scanned"]
            assert st["h2d_bytes"] < len(got.level_counts) * per_stream


def test_open_with_budget(tmp_path):
    n, e = graphs.kron(14)
    pre = str(tmp_path / "k14")
    ga.convert(e, n, pre)
    grid = np.fromfile(pre + ".grid", np.uint32)
    with ga.Graph.open(pre, n, device_budget=4 * CH) as g, ga.Graph.open(pre, n) as ref:
        assert not g.resident and ref.resident
        assert np.array_equal(g.bfs(3).depth, ref.bfs(3).depth)
        assert np.array_equal(g.pagerank(3, ga.PR_COMPAT), ref.pagerank(3, ga.PR_COMPAT))
    with pytest.raises(ga.GStoreError):
        ga.Graph.open(pre, n, device_budget=CH)          # below two segments
    assert grid.size > 4 * 4096


def test_streaming_graph_not_served():
    n, e = graphs.kron(12)
    arrays = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, *arrays) as g, ga.Server(workers=1) as srv:
        g.set_device_budget(2 * CH + CH, CH)
        with pytest.raises(ga.GStoreError):
            srv.add_graph(g)


def test_host_grid_view():
    """gs_get_host_grid: the pinned host grid of a semi-external graph, read in
    place (equal to the grid the tile arrays return); refused for resident
    graphs."""
    n, e = graphs.kron(13)
    start, grid, deg = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        with pytest.raises(ga.GStoreError):
            g.host_grid()
        g.set_device_budget(2 * 3 * CH + CH, 3 * CH)
        assert not g.resident
        view = g.host_grid()
        assert view.size == g.E and not view.flags.writeable
        assert np.array_equal(view, grid)
        s2, g2, d2 = g.arrays(with_grid=False)
        assert g2 is None and np.array_equal(s2, start) and np.array_equal(d2, deg)

