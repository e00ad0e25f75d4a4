"""gstoreu / gstored drop-in command line: convert (-c 1), generate (-g), and
jobs -j 0/1/2/3 (and their 1x/2x variants), checked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from tests import graphs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GSU = os.path.join(ROOT, "gstore_amd", "bin", "gstoreu")
GSD = os.path.join(ROOT, "gstore_amd", "bin", "gstored")


def run(*args, exe=GSU):
    print("[gpu-cli]", os.path.basename(exe), *args, flush=True)
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


@pytest.fixture(scope="module")
def converted(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    n, e = graphs.kron(13)
    edges = str(d / "edges.bin")
    e.astype(np.uint32).tofile(edges)
    pre = str(d / "k13")
    out = run("-v", n, "-i", edges, "-o", pre, "-c", 1)
    assert "Total Tiles Count (in One Direction) = 256" in out
    return n, e, pre, d


def test_convert_matches_oracle(converted):
    n, e, pre, _ = converted
    start, grid, deg = oracle.convert(e, n)
    assert np.array_equal(np.fromfile(pre + ".start", np.uint64), start)
    assert np.array_equal(np.fromfile(pre + ".degree", np.int32), deg)


def test_jobs_match_oracle(converted):
    n, e, pre, d = converted
    start, grid, deg = oracle.convert(e, n)
    # BFS -j 0 -a 5 and -j 20 (root 1, gstore.cpp:2634)
    for job, root in ((0, 5), (20, 1), (10, 1)):
        out = run("-v", n, "-i", pre, "-j", job, "-a", 5, "-p", str(d / "depth.bin"))
        assert f"Root = {root}" in out
        depth, counts = oracle.bfs(n, start, grid, root)
        got = [int(m) for m in re.findall(r"level: \d+ frontier count: (\d+)", out)]
        assert got == counts
        assert np.array_equal(np.fromfile(str(d / "depth.bin"), np.uint8), depth)
    # PageRank -j 1 -a 3 (compat arithmetic, bit-exact)
    out = run("-v", n, "-i", pre, "-j", 21, "-a", 3, "-p", str(d / "rank.bin"))
    assert "Page-rank_mmap time" in out
    assert np.array_equal(np.fromfile(str(d / "rank.bin"), np.float32), oracle.pagerank(n, start, grid, deg, 3, 0))
    # k-core -j 2 -a 4 / -j 22
    _, ne0, _, _ = oracle.kcore(n, start, grid, deg, 4)
    for job in (2, 22):
        out = run("-v", n, "-i", pre, "-j", job, "-a", 4)
        assert f">= 4vertices count = {ne0}" in out
    # WCC -j 3
    _, nc = oracle.wcc(n, start, grid)
    out = run("-v", n, "-i", pre, "-j", 3)
    assert f"WCC count = {nc}" in out


def test_generate_flag(tmp_path):
    pre = str(tmp_path / "g11")
    run("-g", 11, "-o", pre)
    start, grid, deg = oracle.convert(oracle.kron(11, 0, 16 << 11), 1 << 11)
    assert np.array_equal(np.fromfile(pre + ".start", np.uint64), start)
    out = run("-s", 11, "-i", pre, "-j", 3)
    assert f"WCC count = {oracle.wcc(1 << 11, start, grid)[1]}" in out


def test_directed_tool(tmp_path):
    n, e = graphs.kron(11)
    edges = str(tmp_path / "e.bin")
    e.astype(np.uint32).tofile(edges)
    pre = str(tmp_path / "d11")
    run("-v", n, "-i", edges, "-o", pre, "-c", 1, exe=GSD)
    start, grid, deg = oracle.convert(e, n, directed=True)
    assert np.array_equal(np.fromfile(pre + ".start", np.uint64), start)
    out = run("-v", n, "-i", pre, "-j", 0, "-a", 3, exe=GSD)
    got = [int(m) for m in re.findall(r"level: \d+ frontier count: (\d+)", out)]
    assert got == oracle.bfs(n, start, grid, 3, directed=True)[1]
    out = run("-v", n, "-i", pre, "-j", 2, "-a", 4, exe=GSD)
    assert "Kcore not applicable for directed graphs" in out


def test_fast_bfs_default_and_tile_flag(converted):
    """gstoreu runs the direction-optimizing BFS on a resident undirected grid
    and prints the index build on its own line; -b 0 keeps the tile sweep.
    Both print the oracle's level counts."""
    n, e, pre, d = converted
    start, grid, deg = oracle.convert(e, n)
    depth, counts = oracle.bfs(n, start, grid, 7)
    for flag in ([], ["-b", 0]):
        out = run("-v", n, "-i", pre, "-j", 0, "-a", 7, *flag, "-p", str(d / "d.bin"))
        assert [int(m) for m in re.findall(r"level: \d+ frontier count: (\d+)", out)] == counts
        assert np.array_equal(np.fromfile(str(d / "d.bin"), np.uint8), depth)
        assert ("Index build time = " in out) == (not flag)


def test_async_bfs_and_traverse_jobs(converted):
    n, e, pre, d = converted
    start, grid, deg = oracle.convert(e, n)
    out = run("-v", n, "-i", pre, "-j", 30, "-p", str(d / "d30.bin"))
    assert "Root = 1" in out and "Half BFS time" in out
    assert np.array_equal(np.fromfile(str(d / "d30.bin"), np.uint8), oracle.bfs(n, start, grid, 1)[0])
    sweeps = [int(m) for m in re.findall(r"level: \d+ frontier count: (\d+)", out)]
    assert sweeps[-1] == 0 and all(sweeps[:-1])
    out = run("-v", n, "-i", pre, "-j", 4, "-a", 3, "-p", str(d / "c4.bin"))
    assert "Single color propagation time" in out
    cid, counts = oracle.traverse(n, start, grid, 3, 0)
    assert np.array_equal(np.fromfile(str(d / "c4.bin"), np.uint32), cid)


def test_text_to_bin_tool(tmp_path):
    rng = np.random.default_rng(2)
    e = rng.integers(5, 1 << 18, (20000, 2)).astype(np.uint32)
    text = b"%%MatrixMarket comment\n" + b"".join(f"{a}\t{b}\n".encode() for a, b in e)
    p = tmp_path / "g.txt"
    p.write_bytes(text)
    r = subprocess.run([GSU, "-i", str(p), "-c", "4"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want, vmin, vmax, ntok = oracle.text_to_bin(text)
    got = np.fromfile(str(p) + ".edge", np.uint32).reshape(-1, 2)
    assert np.array_equal(got, want) and np.array_equal(got, e - vmin)
    assert f"min vertex id: {vmin}" in r.stderr and f"edge count: {len(e)}" in r.stderr
