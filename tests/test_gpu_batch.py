"""Batched (multi-source, 64 bit-lane) BFS: every lane must equal the
single-root BFS of the oracle bit for bit, whatever else shares the batch."""
import numpy as np
import pytest

import gstore_amd as ga
import oracle
from tests import graphs

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-batch]", *a, flush=True)


@pytest.fixture(scope="module")
def kron14():
    n, e = graphs.kron(14)
    arrays = oracle.convert(e, n)
    g = ga.Graph.from_arrays(n, *arrays)
    yield n, arrays, g
    g.close()


def _check_batch(g, n, start, grid, roots):
    depth, counts = g.bfs_batch(roots)
    for i, r in enumerate(roots):
        d, lc = oracle.bfs(n, start, grid, int(r))
        assert np.array_equal(depth[i], d), (i, r)
        assert counts[i][:len(lc)] == lc and all(c == 0 for c in counts[i][len(lc):])


def test_batch_64_roots_equal_single(kron14):
    n, (start, grid, deg), g = kron14
    rng = np.random.default_rng(1)
    roots = rng.choice(np.flatnonzero(deg), 64, replace=False)
    _say("64 roots")
    _check_batch(g, n, start, grid, roots)


@pytest.mark.parametrize("nr", [1, 7, 33])
def test_batch_partial_and_duplicate_roots(kron14, nr):
    n, (start, grid, deg), g = kron14
    rng = np.random.default_rng(nr)
    roots = rng.integers(0, n, nr)          # isolated roots and repeats allowed
    if nr > 1:
        roots[-1] = roots[0]
    _say("roots", nr)
    _check_batch(g, n, start, grid, roots)


def test_batch_lane_independent_of_mix(kron14):
    n, (start, grid, deg), g = kron14
    a, _ = g.bfs_batch([5])
    b, _ = g.bfs_batch(list(range(60, 123)) + [5])
    assert np.array_equal(a[0], b[-1])


@pytest.mark.parametrize("name", ["path", "tiles", "components"])
def test_batch_hand_graphs(name):
    n, e = graphs.HAND[name]()
    start, grid, deg = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        _say(name)
        _check_batch(g, n, start, grid, [0, 1, n - 1, int(np.argmax(deg))])


def test_batch_directed():
    n, e = graphs.kron(12)
    start, grid, deg = oracle.convert(e, n, directed=True)
    with ga.Graph.from_arrays(n, start, grid, deg, directed=True) as g:
        roots = [0, 3, 9, int(np.argmax(deg))]
        depth, _ = g.bfs_batch(roots)
        for i, r in enumerate(roots):
            assert np.array_equal(depth[i], oracle.bfs(n, start, grid, r, directed=True)[0])


def test_batch_errors(kron14):
    n, _, g = kron14
    with pytest.raises(ga.GStoreError):
        g.bfs_batch(list(range(65)))
    with pytest.raises(ga.GStoreError):
        g.bfs_batch([n])


@pytest.mark.parametrize("nr", [1, 64])
def test_batch_methods_identical(kron14, nr):
    """Lanes over the neighbour-list index (host-driven and device-driven
    levels) and over the tile sweep give the same depths and level counts."""
    n, (start, grid, deg), g = kron14
    rng = np.random.default_rng(100 + nr)
    roots = rng.choice(np.flatnonzero(deg), nr, replace=False)
    got = []
    for method in (0, 1, 2):
        g.set_batch_bfs_method(method)
        got.append(g.bfs_batch(roots))
    g.set_batch_bfs_method(2)
    for k in (1, 2):
        assert np.array_equal(got[0][0], got[k][0]) and got[0][1] == got[k][1]
    with pytest.raises(ga.GStoreError):
        g.set_batch_bfs_method(3)


@pytest.mark.parametrize("s", [14, 18])
def test_batch_graph_replays(s):
    """The captured level sequence launched for one batch after another, with
    other work on the graph between them: every launch matches the host-driven
    levels and the oracle, including a root without edges, one root, 64 roots
    and batches deeper than one launch's four levels."""
    n, e = graphs.kron(s)
    start, grid, deg = oracle.convert(e, n)
    rng = np.random.default_rng(1000 + s)
    nz = np.flatnonzero(deg)
    iso = np.flatnonzero(deg == 0)
    batches = [rng.choice(nz, 64, replace=False), nz[-1:], rng.choice(nz, 7, replace=False),
               np.concatenate([iso[:1], rng.choice(nz, 20, replace=False)]) if iso.size else nz[:21],
               rng.choice(nz, 64, replace=False)]
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        for i, roots in enumerate(batches):
            _say("graph replay", s, i, len(roots))
            g.set_batch_bfs_method(1)
            want = g.bfs_batch(roots)
            g.bfs(int(roots[0]))          # eager work between the launches
            g.set_batch_bfs_method(2)
            got = g.bfs_batch(roots)
            assert np.array_equal(got[0], want[0]) and got[1] == want[1]
            d, lc = oracle.bfs(n, start, grid, int(roots[-1]))
            assert np.array_equal(got[0][-1], d) and got[1][-1][:len(lc)] == lc
        assert max(len(c) for c in got[1]) > 4 or s == 14


def test_batch_index_scale18():
    """A graph with several index blocks and both traversal directions."""
    n, e = graphs.kron(18)
    start, grid, deg = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        rng = np.random.default_rng(18)
        roots = rng.choice(np.flatnonzero(deg), 64, replace=False)
        depth, counts = g.bfs_batch(roots)
        for i in (0, 17, 63):
            d, lc = oracle.bfs(n, start, grid, int(roots[i]))
            assert np.array_equal(depth[i], d) and counts[i][:len(lc)] == lc
