"""Serving: continuous batching over several resident graphs.  Every query's
result in the mix must equal the oracle's (BFS depth and WCC labels bit for
bit, PageRank compat bit for bit, intended within 1e-6, personalised PageRank
within 1e-6 L1 of the FP64 oracle) and the same query run alone (FP32
tolerance where other workspaces' traffic changes the order of FP32 adds)."""
import numpy as np
import pytest

import gstore_amd as ga
import oracle
from tests import graphs

pytestmark = pytest.mark.gpu
PPR_L1 = 1e-6
PR_MAXABS = 1e-6


def _say(*a):
    print("[gpu-serve]", *a, flush=True)


@pytest.fixture(scope="module")
def fleet():
    out = []
    for name, (n, e) in {"kron12": graphs.kron(12), "kron15": graphs.kron(15), "tiles": graphs.tile_crossing()}.items():
        arrays = oracle.convert(e, n)
        out.append((name, n, arrays, ga.Graph.from_arrays(n, *arrays)))
    yield out
    for *_, g in out:
        g.close()


def test_ppr_single_vs_fp64(fleet):
    for name, n, (start, grid, deg), g in fleet:
        for src in (0, int(np.argmax(deg))):
            _say("ppr", name, src)
            got = g.ppr(src, 20)
            want = oracle.ppr(n, start, grid, deg, src, 20)
            l1 = float(np.abs(got.astype(np.float64) - want).sum())
            _say("  L1 vs fp64", l1)
            assert l1 <= PPR_L1


def _alone(g, kind, arg):
    if kind == ga.Q_BFS:
        return g.bfs(arg).depth
    if kind == ga.Q_PPR:
        return g.ppr(arg, 20)
    if kind == ga.Q_WCC:
        return g.wcc()[0]
    return g.pagerank(20, arg)


class _Oracle:
    """Oracle results per (graph, kind, arg), computed once."""

    def __init__(self, fleet):
        self.fleet, self.memo = fleet, {}

    def __call__(self, gi, kind, arg):
        key = (gi, kind, arg)
        if key not in self.memo:
            name, n, (start, grid, deg), g = self.fleet[gi]
            if kind == ga.Q_BFS:
                v = oracle.bfs(n, start, grid, arg)[0]
            elif kind == ga.Q_PPR:
                v = oracle.ppr(n, start, grid, deg, arg, 20)
            elif kind == ga.Q_WCC:
                v = oracle.wcc(n, start, grid)[0]
            else:
                v = oracle.pagerank(n, start, grid, deg, 20, arg)
            self.memo[key] = v
        return self.memo[key]


def _check_oracle(want, kind, arg, out):
    if kind in (ga.Q_BFS, ga.Q_WCC) or (kind == ga.Q_PAGERANK and arg == ga.PR_COMPAT):
        assert np.array_equal(out, want)
    elif kind == ga.Q_PPR:
        assert float(np.abs(out.astype(np.float64) - want).sum()) <= PPR_L1
    else:
        assert np.max(np.abs(out - want)) <= PR_MAXABS


def test_mix_equals_alone(fleet):
    rng = np.random.default_rng(7)
    queries = []
    for i in range(240):
        gi = int(rng.integers(0, len(fleet)))
        n = fleet[gi][1]
        u = rng.random()
        kind = ga.Q_BFS if u < 0.7 else ga.Q_PPR if u < 0.85 else ga.Q_WCC if u < 0.95 else ga.Q_PAGERANK
        arg = int(rng.integers(0, n)) if kind in (ga.Q_BFS, ga.Q_PPR) else int(rng.integers(0, 2))
        queries.append((gi, kind, arg))
    with ga.Server(workers=4) as srv:
        ids = [srv.add_graph(f[3]) for f in fleet]
        outs = {}
        for qi, (gi, kind, arg) in enumerate(queries):
            n = fleet[gi][1]
            out = np.empty(n, np.uint8) if kind == ga.Q_BFS else np.empty(n, np.float32 if kind != ga.Q_WCC else np.uint32)
            qid = srv.submit(ids[gi], kind, arg, out)
            outs[qid] = (qi, out)
        srv.drain()
        res = srv.poll(4096)
    assert len(res) == len(queries)
    assert all(r["status"] == 0 for r in res)
    assert max(r["batch_size"] for r in res if r["kind"] == ga.Q_BFS) > 1, "no BFS batching happened"
    ora = _Oracle(fleet)
    for r in res:
        qi, out = outs[r["id"]]
        gi, kind, arg = queries[qi]
        _check_oracle(ora(gi, kind, arg), kind, arg, out)
        alone = _alone(fleet[gi][3], kind, arg)
        if kind in (ga.Q_BFS, ga.Q_WCC):
            assert np.array_equal(out, alone), (qi, kind, arg)
        else:
            assert np.max(np.abs(out - alone)) <= PR_MAXABS, (qi, kind, arg)
        assert r["t_done_ms"] >= r["t_start_ms"] >= r["t_submit_ms"]
    _say("mix ok, batches up to", max(r["batch_size"] for r in res))


def test_lone_bfs_takes_the_single_root_path(fleet):
    """A BFS with no other BFS of its graph waiting runs alone on the
    single-root direction-optimizing path (batch size 1, same depths); the
    batching rule measured at add time is reported."""
    ora = _Oracle(fleet)
    with ga.Server(workers=2) as srv:
        ids = [srv.add_graph(f[3]) for f in fleet]
        for gi, gid in enumerate(ids):
            pol = srv.bfs_policy(gid)
            _say("policy", fleet[gi][0], pol)
            assert 1 <= pol["batch_min"] <= 65 and pol["single_ms"] > 0 and pol["batch64_ms"] > 0
            n = fleet[gi][1]
            for root in (0, 1, n - 1):
                out = np.empty(n, np.uint8)
                srv.submit(gid, ga.Q_BFS, root, out)
                srv.drain()
                (r,) = srv.poll(16)
                assert r["status"] == 0 and r["batch_size"] == 1
                assert np.array_equal(out, ora(gi, ga.Q_BFS, root))
                assert r["value"] == np.count_nonzero(out)
        # batch_min 1: every BFS batches, even alone
        srv.set_batch_min(ids[0], 1)
        out = np.empty(fleet[0][1], np.uint8)
        srv.submit(ids[0], ga.Q_BFS, 3, out)
        srv.drain()
        (r,) = srv.poll(16)
        assert r["status"] == 0 and np.array_equal(out, ora(0, ga.Q_BFS, 3))
        mem = srv.memory()
        assert mem["graph_bytes"] >= sum(f[3].E * 4 for f in fleet)
        assert srv.last_error() == ""


def test_budget_admission(fleet):
    name, n, arrays, g = fleet[1]
    # one BFS with its depths over the index: lane masks 24n, queue + degrees 8n,
    # prefix 8(n+1) + block sums, depth rows 64n, lane-major depth n, small blocks
    need_bfs_one = 24 * n + 8 * n + 8 * (n + 1) + 8 * (n // 1024 + 1) + 64 * n + n + (1 << 20)
    budget = 2 * need_bfs_one         # room for about two small BFS workspaces at a time
    with ga.Server(hbm_budget=budget, workers=4) as srv:
        gid = srv.add_graph(g)
        rng = np.random.default_rng(3)
        roots = rng.integers(0, n, 100)
        outs = {}
        for r in roots:
            out = np.empty(n, np.uint8)
            outs[srv.submit(gid, ga.Q_BFS, int(r), out)] = (int(r), out)
        peak = 0
        got = []
        while len(got) < len(roots):
            got += srv.poll(4096, wait_ms=50)
            peak = max(peak, srv.info()["held_bytes"])
        assert peak <= budget
        for res in got:
            r, out = outs[res["id"]]
            assert np.array_equal(out, g.bfs(r).depth)
    with ga.Server(hbm_budget=4 * n, workers=1) as tiny:
        gid = tiny.add_graph(g)
        with pytest.raises(ga.GStoreError):
            tiny.submit(gid, ga.Q_BFS, 0)          # needs more than the whole budget
