"""gs_bfs_noalloc (the BFS entry point that allocates nothing) against gs_bfs,
called through the C ABI: two calls in a row, the second on a smaller graph,
and the size limit of its static work array."""
import ctypes

import numpy as np
import pytest

import gstore_amd as ga
import oracle

pytestmark = pytest.mark.gpu


def _noalloc(g, root, max_levels=256):
    lib = ga.load_library()
    fn = lib.gs_bfs_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_int)]
    depth = np.full(g.n, 0xAB, np.uint8)
    counts = np.zeros(max_levels, np.uint64)
    nlev = ctypes.c_int(0)
    rc = fn(g._h, root, depth.ctypes.data, counts.ctypes.data, max_levels, ctypes.byref(nlev))
    return rc, depth, [int(c) for c in counts[:nlev.value]]


def _check_against_gs_bfs(g, root):
    g.set_bfs_mode(0)
    want = g.bfs(root)
    rc, depth, counts = _noalloc(g, root)
    assert rc == 0, ga.GStoreError(rc, "")
    assert np.array_equal(depth, want.depth)
    assert counts == want.level_counts
    return depth, counts


def test_noalloc_bfs_two_calls_second_smaller():
    print("noalloc: Kron-16 root = max degree, then Kron-12 root 1", flush=True)
    with ga.Graph.kron(16) as g:
        deg = g.degree()
        root = int(np.argmax(deg))
        d1, c1 = _check_against_gs_bfs(g, root)
        start, grid, degree = g.arrays()
        od, oc = oracle.bfs(g.n, start, grid, root)
        assert np.array_equal(d1, od) and c1 == oc
    with ga.Graph.kron(12) as g:
        # the scratch array still holds the Kron-16 run: the smaller call must clear what it uses
        d2, c2 = _check_against_gs_bfs(g, 1)
        start, grid, degree = g.arrays()
        od, oc = oracle.bfs(g.n, start, grid, 1)
        assert np.array_equal(d2, od) and c2 == oc


def test_noalloc_bfs_tile_crossing_and_directed():
    from tests import graphs
    n, e = graphs.tile_crossing()
    for directed in (False, True):
        start, grid, degree = oracle.convert(e, n, directed)
        with ga.Graph.from_arrays(n, start, grid, degree, directed=directed) as g:
            for root in (0, 65535, 131072):
                _check_against_gs_bfs(g, root)


def test_noalloc_bfs_rejects_graphs_beyond_its_array():
    n = 1 << 25          # 1.375 N bytes of work arrays > 32 MiB
    start, grid, degree = oracle.convert(np.array([[0, 1], [1, n - 1]], np.uint32), n)
    with ga.Graph.from_arrays(n, start, grid, degree) as g:
        rc, _, _ = _noalloc(g, 0)
        assert rc == 1   # GS_ERR_ARG
        buf = ctypes.create_string_buffer(512)
        ga.load_library().gs_last_error(buf, 512)
        assert b"static array" in buf.value
        # the allocating entry point still serves this graph
        assert g.bfs(0).level_counts == [1, 1, 0]
