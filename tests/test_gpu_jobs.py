"""The reference's remaining jobs and tool on the device, against the oracle:
-j 30 asynchronous BFS (final depths bit-exact; they equal the plain BFS's),
-j 4 single colour propagation (colour array bit-exact), -c 4 text edge list
to binary pairs (pairs, minimum, maximum and token count exact, including the
reference's quirks).  Per-sweep counts of the two jobs depend on the order
updates land (in the reference too), so they are checked by invariants."""
import json
import os
import zlib

import numpy as np
import pytest

import gstore_amd as ga
import oracle
from tests import graphs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[gpu-jobs]", *a, flush=True)


def _cases():
    out = [(name, *fn()) for name, fn in graphs.HAND.items()]
    out.append(("two_groups",) + graphs.two_groups())
    for s in (10, 14, 16):
        out.append((f"kron{s}",) + graphs.kron(s))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_bfs_async_matches_oracle(name, n, e, directed):
    start, grid, deg = oracle.convert(e, n, directed)
    with ga.Graph.from_arrays(n, start, grid, deg, directed=directed) as g:
        for root in sorted({0, 1, n - 1, int(np.argmax(deg))}):
            _say("bfs_async", name, directed, root)
            got, sweeps = g.bfs_async(root)
            want, _ = oracle.bfs2(n, start, grid, root, directed)
            assert np.array_equal(got, want), (name, root)
            assert np.array_equal(got, oracle.bfs(n, start, grid, root, directed)[0])
            assert sweeps[-1] == 0 and all(sweeps[:-1])
            assert sum(sweeps) >= int(np.count_nonzero(got)) - 1


@pytest.mark.parametrize("name,n,e", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("directed", [False, True])
def test_traverse_matches_oracle(name, n, e, directed):
    start, grid, deg = oracle.convert(e, n, directed)
    with ga.Graph.from_arrays(n, start, grid, deg, directed=directed) as g:
        for pivot in sorted({0, 1, n // 2, int(np.argmax(deg))}):
            _say("traverse", name, directed, pivot)
            cid, sweeps, colored = g.traverse(pivot, color=5)
            want, _ = oracle.traverse(n, start, grid, pivot, 5, directed)
            assert np.array_equal(cid, want), (name, pivot)
            assert colored == int(np.count_nonzero(want == 5)) == 1 + sum(sweeps)
            assert sweeps[-1] == 0


def test_jobs_reject_bad_arguments():
    n, e = graphs.path(10)
    with ga.Graph.from_arrays(n, *oracle.convert(e, n)) as g:
        with pytest.raises(ga.GStoreError):
            g.traverse(0, color=0xFFFFFFFF)
        with pytest.raises(ga.GStoreError):
            g.bfs_async(n)
        n2, e2 = graphs.path(300)
    with ga.Graph.from_arrays(n2, *oracle.convert(e2, n2)) as g:
        with pytest.raises(ga.GStoreError) as ei:
            g.bfs_async(0)                       # depth 300 does not fit the u8 encoding
        assert ei.value.status == 7
        assert np.array_equal(g.bfs_async(150)[0], oracle.bfs(n2, *oracle.convert(e2, n2)[:2], 150)[0])


def test_streamed_jobs_match_oracle():
    n, e = graphs.kron(15)
    start, grid, deg = oracle.convert(e, n)
    with ga.Graph.from_arrays(n, start, grid, deg) as g:
        g.set_device_budget(2 * 3 * 4096 * 4 + 4096 * 4, 3 * 4096 * 4)
        assert not g.resident
        assert np.array_equal(g.bfs_async(7)[0], oracle.bfs(n, start, grid, 7)[0])
        assert np.array_equal(g.traverse(7, 1)[0], oracle.traverse(n, start, grid, 7, 1)[0])


def test_bfs_async_at_kron25_matches_fixture():
    path = os.path.join(GOLDEN, "scale_kron25.json")
    fx = json.load(open(path))
    with ga.Graph.kron(25) as g:
        for rec in fx["bfs"][:3] + fx["bfs"][-1:]:
            d, sweeps = g.bfs_async(rec["root"])
            assert zlib.crc32(d.view(np.uint8)) == rec["depth_crc32"], rec["root"]
            _say("kron25 bfs_async", rec["root"], "sweeps", len(sweeps), "ms", g.stats()["ms"])


TEXTS = [b"1 2\n3 4\n", b"%%MatrixMarket\n% c\n0 5\n5 9\n", b"3\t4\r\n5 6\r\n", b" 5 6\n7\t8", b"2 3\n4",
         b"1000001 1000002\n1000003 1000004\n", b"7 -1\n", b"", b"\n\n", b"%only a header\n", b"12",
         b"99999999999999999999 3\n-99999999999999999999 4\n", b"+5 6\n\t 7 8 \n"]


@pytest.mark.parametrize("text", TEXTS, ids=[f"t{i}" for i in range(len(TEXTS))])
def test_text_to_bin_known_texts(text):
    got = ga.text_to_bin(text)
    want = oracle.text_to_bin(text)
    assert np.array_equal(got[0], want[0]) and got[1:] == want[1:]


def test_text_to_bin_large():
    rng = np.random.default_rng(9)
    e = rng.integers(0, 1 << 30, (1 << 20, 2)).astype(np.uint32)
    seps = [b" ", b"\t", b"   ", b" \t "]
    parts = [b"% header line\n"]
    for i, (a, b) in enumerate(e):
        parts.append(str(a).encode() + seps[i & 3] + str(b).encode() + (b"\r\n" if i % 5 == 0 else b"\n"))
    text = b"".join(parts)
    got = ga.text_to_bin(text)
    want = oracle.text_to_bin(text)
    assert np.array_equal(got[0], want[0]) and got[1:] == want[1:]
    assert got[3] == 2 * len(e)
