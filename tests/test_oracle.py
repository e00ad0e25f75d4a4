"""Pins the CPU oracle (oracle/gs_oracle.c) against golden vectors, library
routines and closed forms.  CPU only."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

import oracle
from gstore_amd import tiles
from tests import graphs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- generator ------------------------------------------------------------

def test_kron_readme_known_answer():
    rows = [l.split() for l in open(os.path.join(GOLDEN, "graph500_readme_s21.txt")) if l.strip() and l[0] != "#"]
    want = np.array(rows, dtype=np.uint32)
    assert np.array_equal(oracle.kron(21, 0, len(want)), want)


def test_kron_reference_vectors():
    vec = np.load(os.path.join(GOLDEN, "kron_ref_vectors.npz"))
    assert len(vec.files) >= 8
    for key in vec.files:
        s, e0 = key[1:].split("_e")
        want = vec[key]
        got = oracle.kron(int(s), int(e0), int(e0) + len(want))
        assert np.array_equal(got, want), key


def test_kron_against_reference_build_live():
    L = oracle.ref_generator()
    if L is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(3)
    for s in (12, 19, 26):
        for e0 in rng.integers(0, 16 << s, 3):
            e0 = int(e0)
            assert np.array_equal(oracle.kron(s, e0, e0 + 500), oracle.ref_kron(L, s, e0, e0 + 500))


def test_kron_seed_changes_stream():
    a = oracle.kron(14, 0, 100, 1, 2)
    b = oracle.kron(14, 0, 100, 3, 4)
    assert not np.array_equal(a, b)
    assert a.max() < (1 << 14)


# ---- converter ------------------------------------------------------------

def _conv_check(n, e, directed=False):
    start, grid, degree = oracle.convert(e, n, directed)
    T = tiles.total_slots(tiles.tiles_per_side(n), directed)
    assert start.size == T + 1 and start[0] == 0 and np.all(np.diff(start.astype(np.int64)) >= 0)
    keep = e[:, 0] != e[:, 1]
    assert int(start[-1]) == int(keep.sum()) == grid.size
    a, b = e[keep, 0].astype(np.uint64), e[keep, 1].astype(np.uint64)
    if not directed:
        a, b = np.minimum(a, b), np.maximum(a, b)
    v0, v1 = tiles.decode(n, start, grid, directed)
    got = np.sort(v0 * (1 << 32) + v1)
    want = np.sort(a * (1 << 32) + b)
    assert np.array_equal(got, want)
    deg = np.bincount(e[keep, 0], minlength=n)
    if not directed:
        deg = deg + np.bincount(e[keep, 1], minlength=n)
    assert np.array_equal(degree, deg.astype(np.int32))
    # every edge sits in the slot of its (row, col) tile
    r, c = tiles.slot_coords(tiles.groups(n), directed)
    counts = np.diff(start.astype(np.int64))
    assert np.array_equal(np.repeat(r, counts), (v0 >> 16).astype(np.uint32))
    assert np.array_equal(np.repeat(c, counts), (v1 >> 16).astype(np.uint32))
    return start, grid, degree


@pytest.mark.parametrize("name", list(graphs.HAND))
def test_convert_invariants_hand(name):
    n, e = graphs.HAND[name]()
    _conv_check(n, e)
    _conv_check(n, e, directed=True)


def test_convert_invariants_two_groups():
    n, e = graphs.two_groups()
    _conv_check(n, e)


def test_convert_invariants_kron():
    n, e = graphs.kron(12)
    _conv_check(n, e)


def test_convert_rejects_out_of_range():
    with pytest.raises(ValueError):
        oracle.convert(np.array([[0, 70]], np.uint32), 64)


# ---- helpers: scipy views of a converted graph -------------------------------

def _adj(n, start, grid, directed=False):
    v0, v1 = tiles.decode(n, start, grid, directed)
    A = sp.coo_matrix((np.ones(v0.size), (v0.astype(np.int64), v1.astype(np.int64))), shape=(n, n)).tocsr()
    return A if directed else (A + A.T).tocsr()


def _cases():
    out = []
    for name in ("path", "star", "tiles", "components"):
        n, e = graphs.HAND[name]()
        out.append((name, n, e))
    n, e = graphs.kron(11)
    out.append(("kron11", n, e))
    return out


CASES = _cases()


# ---- PageRank -----------------------------------------------------------------

def _pr_compat_closed_form(n, degree, start, grid, iters):
    """Only degree-1 vertices are non-zero (the (int32)(1.0/deg) factor is 0 for
    deg >= 2).  A leaf u with neighbour w gets x1 = f32(0.15 + 0.85 f32(1/N));
    afterwards f32(0.15) if w is not a leaf, or x_{t+1} = f32(0.15 + 0.85 x_t)
    if {u, w} is an isolated edge."""
    v0, v1 = tiles.decode(n, start, grid)
    nb = np.zeros(n, np.int64)
    nb[v0.astype(np.int64)] = v1.astype(np.int64)
    nb[v1.astype(np.int64)] = v0.astype(np.int64)
    out = np.zeros(n, np.float32)
    x1 = np.float32(0.15 + 0.85 * float(np.float32(1.0 / n)))
    for u in np.nonzero(degree == 1)[0]:
        w = nb[u]
        x = x1
        for _ in range(iters - 1):
            x = np.float32(0.15 + 0.85 * float(x)) if degree[w] == 1 else np.float32(0.15)
        out[u] = x if iters >= 1 else np.float32(1.0 / n)
    return out


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("iters", [1, 2, 5])
def test_pagerank_compat_closed_form(name, n, e, iters):
    start, grid, degree = oracle.convert(e, n)
    got = oracle.pagerank(n, start, grid, degree, iters, mode=0)
    assert np.array_equal(got, _pr_compat_closed_form(n, degree, start, grid, iters))


def _pr_f64(A_T, n, degree, iters):
    fac = np.where(degree > 0, 1.0 / np.maximum(degree, 1), 0.0)
    x = np.full(n, 1.0 / n)
    for _ in range(iters):
        x = (0.15 + 0.85 * (A_T @ x)) * fac
    return x


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_pagerank_intended_vs_float64(name, n, e, directed):
    start, grid, degree = oracle.convert(e, n, directed)
    A = _adj(n, start, grid, directed)
    want = _pr_f64(A.T.tocsr(), n, degree, 10)
    got = oracle.pagerank(n, start, grid, degree, 10, mode=1, directed=directed)
    # FP32 accumulation against an FP64 recurrence; values here are O(1)
    assert np.max(np.abs(got.astype(np.float64) - want)) <= 2e-6 * max(1.0, np.max(np.abs(want)))
    # a mistake that swaps the scatter direction on a directed grid is visible:
    if directed and name == "kron11":
        wrong = _pr_f64(A.tocsr(), n, degree, 10)
        assert np.max(np.abs(wrong - want)) > 1e-3


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
def test_ppr_vs_scipy(name, n, e):
    start, grid, degree = oracle.convert(e, n)
    A = _adj(n, start, grid)
    inv = np.where(degree > 0, 1.0 / np.maximum(degree, 1), 0.0)
    for src in (0, int(np.argmax(degree))):
        r = np.zeros(n)
        r[src] = 1.0
        for _ in range(20):
            r = 0.85 * (A.T @ (r * inv))
            r[src] += 0.15
        got = oracle.ppr(n, start, grid, degree, src, 20)
        assert np.max(np.abs(got - r)) <= 1e-12
        assert got.sum() <= 1.0 + 1e-9


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_pagerank64_vs_float64(name, n, e, directed):
    start, grid, degree = oracle.convert(e, n, directed)
    A = _adj(n, start, grid, directed)
    want = _pr_f64(A.T.tocsr(), n, degree, 10)
    for nt in (1, 4):
        got = oracle.pagerank64(n, start, grid, degree, 10, nthreads=nt, directed=directed)
        assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
    # the FP32 oracle in intended mode sits within FP32 rounding of it
    f32 = oracle.pagerank(n, start, grid, degree, 10, mode=1, directed=directed)
    assert np.max(np.abs(f32 - want)) <= 2e-6 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_pr_step64_sample_is_one_step(name, n, e, directed):
    """Step k -> k+1 of the scipy FP64 recurrence at a vertex sample; the
    iterate of another step, or the grid read the wrong way, fails."""
    start, grid, degree = oracle.convert(e, n, directed)
    A = _adj(n, start, grid, directed)
    rng = np.random.default_rng(4)
    idx = np.unique(np.concatenate([np.argsort(degree)[::-1][:16], rng.choice(n, min(n, 500), replace=False)]))
    x4 = _pr_f64(A.T.tocsr(), n, degree, 4)
    want = _pr_f64(A.T.tocsr(), n, degree, 5)[idx]
    for nt in (1, 4):
        got = oracle.pr_step64_sample(n, start, grid, degree, x4, idx, directed=directed, nthreads=nt)
        assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
    if name == "kron11":
        x3 = _pr_f64(A.T.tocsr(), n, degree, 3)
        assert np.max(np.abs(oracle.pr_step64_sample(n, start, grid, degree, x3, idx, directed=directed) - want)) > 1e-6
        if directed:
            und = oracle.pr_step64_sample(n, start, grid, degree, x4, idx, directed=False)
            assert np.max(np.abs(und - want)) > 1e-6
    with pytest.raises(RuntimeError):   # the sample must be sorted and distinct
        oracle.pr_step64_sample(n, start, grid, degree, x4, idx[::-1].copy(), directed=directed)


@pytest.mark.parametrize("name,n,e", CASES + [("two_groups",) + graphs.two_groups()],
                         ids=[c[0] for c in CASES] + ["two_groups"])
@pytest.mark.parametrize("directed", [False, True])
def test_bfs_check_accepts_bfs_and_rejects_planted_mistakes(name, n, e, directed):
    start, grid, degree = oracle.convert(e, n, directed)
    A = _adj(n, start, grid, directed)
    rng = np.random.default_rng(6)
    roots = [0, int(np.argmax(degree)), n - 1]
    planted = 0
    for root in roots:
        d = _bfs_expected(A, root)
        assert oracle.bfs_check(n, start, grid, root, d, directed) == 0, (name, root)
        assert oracle.bfs_check(n, start, grid, root, d, directed, nthreads=4) == 0
        reached = np.flatnonzero(d > 1)
        if reached.size == 0:
            continue
        for v in rng.choice(reached, min(3, reached.size), replace=False):
            for delta in (1, -1):
                bad = d.copy()
                bad[v] = int(d[v]) + delta   # d[v] >= 2, so this stays a reached depth
                assert oracle.bfs_check(n, start, grid, root, bad, directed) > 0, (name, root, int(v), delta)
                planted += 1
            bad = d.copy()
            bad[v] = 0   # a reached vertex dropped
            assert oracle.bfs_check(n, start, grid, root, bad, directed) > 0
        unreached = np.flatnonzero((d == 0) & (degree > 0))
        if unreached.size:
            bad = d.copy()
            u = int(unreached[0])
            bad[u] = 2   # a vertex of another component marked reached
            assert oracle.bfs_check(n, start, grid, root, bad, directed) > 0
            bad[u] = 1   # or a second depth-1 vertex
            assert oracle.bfs_check(n, start, grid, root, bad, directed) > 0
    bad = _bfs_expected(A, 0)
    bad[0] = 2   # the root must have depth 1
    assert oracle.bfs_check(n, start, grid, 0, bad, directed) > 0
    assert planted > 0 or name in ("components",)


def test_ppr_threaded_vs_scipy():
    n, e = graphs.kron(12)
    start, grid, degree = oracle.convert(e, n)
    A = _adj(n, start, grid)
    inv = np.where(degree > 0, 1.0 / np.maximum(degree, 1), 0.0)
    src = int(np.argmax(degree))
    r = np.zeros(n)
    r[src] = 1.0
    for _ in range(20):
        r = 0.85 * (A.T @ (r * inv))
        r[src] += 0.15
    got = oracle.ppr(n, start, grid, degree, src, 20, nthreads=4)
    assert np.max(np.abs(got - r)) <= 1e-12


def test_kron_convert_chunked_equals_whole_list():
    for s, chunk in ((12, 1000), (17, 1 << 20)):
        want = oracle.convert(oracle.kron(s, 0, 16 << s), 1 << s)
        got = oracle.kron_convert(s, chunk=chunk)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


def test_grid_checksum_is_a_multiset_hash():
    n, e = graphs.tile_crossing()
    start, grid, degree = oracle.convert(e, n)
    c = oracle.grid_checksum(start, grid)
    # in-tile permutation leaves it unchanged
    rng = np.random.default_rng(1)
    g2 = grid.copy()
    full = np.flatnonzero(np.diff(start.astype(np.int64)) > 3)
    assert full.size >= 4
    for t in full:
        a, b = int(start[t]), int(start[t + 1])
        g2[a:b] = rng.permutation(g2[a:b])
    assert oracle.grid_checksum(start, g2) == c
    # an altered edge word, or an edge moved to the neighbouring tile, changes it
    g3 = grid.copy()
    g3[7] ^= 1
    assert oracle.grid_checksum(start, g3) != c
    t = int(np.flatnonzero(np.diff(start.astype(np.int64)) > 0)[3])
    s4 = start.copy()
    s4[t + 1] -= 1      # last edge of tile t now belongs to tile t+1
    assert oracle.grid_checksum(s4, grid) != c


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
def test_threaded_forms_equal_serial(name, n, e):
    for directed in (False, True):
        start, grid, degree = oracle.convert(e, n, directed)
        for root in (0, int(np.argmax(degree)), n - 1):
            a, b = oracle.bfs(n, start, grid, root, directed), oracle.bfs(n, start, grid, root, directed, nthreads=4)
            assert np.array_equal(a[0], b[0]) and a[1] == b[1]
        a, b = oracle.wcc(n, start, grid, directed), oracle.wcc(n, start, grid, directed, nthreads=4)
        assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    start, grid, degree = oracle.convert(e, n)
    for k in (2, 4, 8):
        a, b = oracle.kcore(n, start, grid, degree, k), oracle.kcore(n, start, grid, degree, k, nthreads=4)
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


def test_pagerank_parallel_matches_sequential_within_fp32():
    n, e = graphs.kron(12)
    start, grid, degree = oracle.convert(e, n)
    a = oracle.pagerank(n, start, grid, degree, 5, mode=1, nthreads=1)
    b = oracle.pagerank(n, start, grid, degree, 5, mode=1, nthreads=4)
    assert np.max(np.abs(a - b)) <= 1e-6


# ---- BFS -----------------------------------------------------------------------

def _bfs_expected(A, root):
    d = csgraph.shortest_path(A, unweighted=True, indices=root, directed=True)
    out = np.zeros(A.shape[0], np.uint8)
    ok = np.isfinite(d)
    out[ok] = (d[ok] + 1).astype(np.uint8)
    return out


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_bfs_vs_scipy(name, n, e, directed):
    start, grid, degree = oracle.convert(e, n, directed)
    A = _adj(n, start, grid, directed)
    roots = [0, 1, 7, int(np.argmax(degree)), n - 1]
    for root in roots:
        depth, counts = oracle.bfs(n, start, grid, root, directed)
        want = _bfs_expected(A, root)
        assert np.array_equal(depth, want), (name, root)
        hist = np.bincount(want, minlength=256)
        assert counts[:-1] == [int(hist[l]) for l in range(2, len(counts) + 1)]
        assert counts[-1] == 0


def test_bfs_depth_overflow_reported():
    n, e = graphs.path(300)
    start, grid, _ = oracle.convert(e, n)
    with pytest.raises(RuntimeError):
        oracle.bfs(n, start, grid, 0)
    depth, counts = oracle.bfs(n, start, grid, 150)  # eccentricity 150 < 254
    assert depth.max() == 151


# ---- WCC -----------------------------------------------------------------------

def _wcc_expected(A, degree):
    nc, lab = csgraph.connected_components(A, directed=False)
    n = A.shape[0]
    mins = np.full(nc, n, np.int64)
    np.minimum.at(mins, lab, np.arange(n))
    out = mins[lab].astype(np.uint32)
    out[degree == 0] = 0xFFFFFFFF
    return out, len(np.unique(lab[degree > 0]))


@pytest.mark.parametrize("name,n,e", CASES + [("two_groups",) + graphs.two_groups()],
                         ids=[c[0] for c in CASES] + ["two_groups"])
def test_wcc_vs_scipy(name, n, e):
    start, grid, degree = oracle.convert(e, n)
    A = _adj(n, start, grid)
    want, wn = _wcc_expected(A, degree)
    got, gn = oracle.wcc(n, start, grid)
    assert np.array_equal(got, want) and gn == wn


def test_wcc_directed_is_weak():
    n, e = graphs.components()
    s1, g1, _ = oracle.convert(e, n, directed=True)
    s2, g2, _ = oracle.convert(e, n)
    assert np.array_equal(oracle.wcc(n, s1, g1, directed=True)[0], oracle.wcc(n, s2, g2)[0])


# ---- k-core ----------------------------------------------------------------------

def _core_bruteforce(n, e, k):
    """Order-independent peeling, one vertex at a time, on the multigraph."""
    keep = e[:, 0] != e[:, 1]
    adj = [[] for _ in range(n)]
    deg = np.zeros(n, np.int64)
    for a, b in e[keep]:
        adj[a].append(b)
        adj[b].append(a)
        deg[a] += 1
        deg[b] += 1
    alive = deg > 0
    stack = [v for v in range(n) if alive[v] and deg[v] < k]
    gone = np.zeros(n, bool)
    while stack:
        v = stack.pop()
        if gone[v]:
            continue
        gone[v] = True
        for u in adj[v]:
            deg[u] -= 1
            if not gone[u] and deg[u] < k:
                stack.append(u)
    return alive & ~gone


def _kcore_final_degrees(n, e, degree0, k):
    """Synchronous rounds from the vertex side: R_r = {v : 0 < int16(d_r[v]) < k};
    removed vertices end at -#(incident edges whose other end left in the same
    or a later round), kept ones at deg0 - #(incident edges to removed ones)."""
    keep = e[:, 0] != e[:, 1]
    a, b = e[keep, 0].astype(np.int64), e[keep, 1].astype(np.int64)
    rnd = np.zeros(n, np.int64)     # 0 = never removed
    d = degree0.astype(np.int64).copy()
    r = 0
    while True:
        r += 1
        m = d.astype(np.int16)
        rem = (m > 0) & (m < k) & (rnd == 0)
        if not rem.any():
            break
        rnd[rem] = r
        d[rem] = 0
        d -= np.bincount(b[rem[a]], minlength=n) + np.bincount(a[rem[b]], minlength=n)
    ra, rb = rnd[a], rnd[b]
    inf = np.iinfo(np.int64).max
    ra2, rb2 = np.where(ra == 0, inf, ra), np.where(rb == 0, inf, rb)
    final = degree0.astype(np.int64).copy()
    removed = rnd > 0
    final[removed] = 0
    # removed v: loses 1 per edge to a removed u with rnd[u] >= rnd[v]
    final -= (np.bincount(a[(ra > 0) & (rb > 0) & (rb2 >= ra2)], minlength=n)
              + np.bincount(b[(rb > 0) & (ra > 0) & (ra2 >= rb2)], minlength=n))
    # kept v: loses 1 per edge to a removed u
    final -= np.bincount(a[(ra == 0) & (rb > 0)], minlength=n) + np.bincount(b[(rb == 0) & (ra > 0)], minlength=n)
    return final.astype(np.int32), r


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("k", [2, 3, 4, 8])
def test_kcore_vs_vertex_side_and_bruteforce(name, n, e, k):
    start, grid, degree = oracle.convert(e, n)
    deg, ne0, pos, rounds = oracle.kcore(n, start, grid, degree, k)
    want, r = _kcore_final_degrees(n, e, degree, k)
    assert np.array_equal(deg, want)
    assert rounds == r
    assert ne0 == int((want != 0).sum()) and pos == int((want > 0).sum())
    assert np.array_equal(deg > 0, _core_bruteforce(n, e, k))


def test_kcore_int16_quirk():
    # 65538 copies of edge (0, 1): degree 65538 reads as (int16) 2, so both go at k = 4
    e = np.concatenate([np.tile([[0, 1]], (65538, 1)), [[2, 3], [3, 4], [4, 2]]]).astype(np.uint32)
    n = 8
    start, grid, degree = oracle.convert(e, n)
    deg, ne0, pos, _ = oracle.kcore(n, start, grid, degree, 4)
    assert deg[0] == deg[1] == -65538
    assert ne0 == 5 and pos == 0   # 0, 1 negative; 2, 3, 4 removed at (2 < 4) and left at -2


# ---- -j 30 (bfs2), -j 4 (traverse), -c 4 (text to binary) ---------------------------

@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_bfs2_reaches_bfs_depths(name, n, e, directed):
    """The asynchronous sweeps' fixed point is the breadth-first depth."""
    start, grid, degree = oracle.convert(e, n, directed)
    A = _adj(n, start, grid, directed)
    for root in (0, 1, int(np.argmax(degree)), n - 1):
        depth, counts = oracle.bfs2(n, start, grid, root, directed)
        assert np.array_equal(depth, _bfs_expected(A, root)), (name, root)
        assert counts[-1] == 0 and all(c > 0 for c in counts[:-1])
        assert sum(counts) >= int(np.count_nonzero(depth)) - 1   # every reached vertex was set at least once


def test_bfs2_depth_overflow_reported():
    n, e = graphs.path(300)
    start, grid, _ = oracle.convert(e, n)
    with pytest.raises(RuntimeError):
        oracle.bfs2(n, start, grid, 0)


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_traverse_colours_the_oriented_reach(name, n, e, directed):
    """Colour spreads along stored edges v0 -> v1 only: on an undirected grid
    (v0 < v1) the coloured set is what increasing-id paths reach."""
    start, grid, degree = oracle.convert(e, n, directed)
    v0, v1 = tiles.decode(n, start, grid, directed)
    D = sp.coo_matrix((np.ones(v0.size), (v0.astype(np.int64), v1.astype(np.int64))), shape=(n, n)).tocsr()
    for pivot in (0, 1, int(np.argmax(degree)), n // 2):
        cid, counts = oracle.traverse(n, start, grid, pivot, color=7, directed=directed)
        reach = np.isfinite(csgraph.shortest_path(D, unweighted=True, indices=pivot, directed=True))
        assert np.array_equal(cid == 7, reach), (name, pivot)
        assert np.all(cid[~reach] == 0xFFFFFFFF)
        assert counts[-1] == 0 and sum(counts) == int(reach.sum()) - 1


def test_traverse_rejects_invalid_colour():
    n, e = graphs.path(10)
    start, grid, _ = oracle.convert(e, n)
    with pytest.raises(RuntimeError):
        oracle.traverse(n, start, grid, 0, color=0xFFFFFFFF)


@pytest.mark.parametrize("text,pairs,vmin,vmax,ntok", [
    (b"1 2\n3 4\n", [(0, 1), (2, 3)], 1, 4, 4),                     # ids shift down by the minimum
    (b"%%MatrixMarket\n% c\n0 5\n5 9\n", [(0, 5), (5, 9)], 0, 9, 4),  # '%' header lines skipped
    (b"3\t4\r\n5 6\r\n", [(0, 1), (2, 3)], 3, 6, 4),                # tabs, CRLF (\r stays in the token)
    (b" 5 6\n7\t8", [(0, 0), (1, 2)], 5, 8, 5),                    # leading blank: first number twice
    (b"2 3\n4", [(0, 1)], 2, 4, 3),                                 # odd token count: last one dropped
    (b"1000001 1000002\n1000003 1000004\n", [(2, 3), (4, 5)], 999999, 1000004, 4),  # v_min starts at 999999
    (b"7 -1\n", [(0, 0xFFFFFFF8)], 7, 0xFFFFFFFF, 2),               # atoi of -1 is 0xFFFFFFFF as u32, the maximum
    (b"", [], 999999, 0, 0),
])
def test_text_to_bin_known_answers(text, pairs, vmin, vmax, ntok):
    got, mn, mx, nt = oracle.text_to_bin(text)
    assert got.tolist() == [list(p) for p in pairs] and (mn, mx, nt) == (vmin, vmax, ntok)


def test_text_to_bin_round_trip():
    rng = np.random.default_rng(5)
    e = rng.integers(0, 1 << 20, (5000, 2)).astype(np.uint32)
    e[17] = (0, 3)                                   # the minimum id is 0: ids come back unchanged
    seps = [b" ", b"\t", b"  "]
    ends = [b"\n", b"\r\n", b" \n"]
    lines = [str(a).encode() + seps[i % 3] + str(b).encode() + ends[i % 3] for i, (a, b) in enumerate(e)]
    got, mn, mx, nt = oracle.text_to_bin(b"%header\n" + b"".join(lines))
    assert np.array_equal(got, e) and mn == 0 and mx == int(e.max()) and nt == 2 * len(e)
