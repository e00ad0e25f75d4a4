"""CPU oracle for the G-Store hot path -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers over oracle/liboracle.so (gs_oracle.c, a C restatement of the
reference's algorithms; see that file's header for what pins each function).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline; the engine
(gstore_amd/) never imports it.

``ref_generator()`` loads the reference's own Graph500 generator compiled from
its source files into oracle/_ref/ (make -C oracle ref), when present.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_GEN = os.path.join(HERE, "_ref", "libref_graph500.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-B", "liboracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        u64, i64, p, i = ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
        L.or_kron_range.argtypes = [i, u64, u64, i64, i64, p]
        L.or_convert_count.argtypes = [p, u64, u64, i, p, p]
        L.or_convert_count.restype = i64
        L.or_convert_scatter.argtypes = [p, u64, u64, i, p, p]
        L.or_pagerank.argtypes = [u64, i, p, p, p, i, i, i, p]
        L.or_ppr.argtypes = [u64, i, p, p, p, u64, i, i, p]
        L.or_pagerank64.argtypes = [u64, i, p, p, p, i, i, p]
        L.or_kron_convert_count.argtypes = [i, i, u64, u64, i, u64, p, p]
        L.or_kron_convert_count.restype = i64
        L.or_kron_convert_scatter.argtypes = [i, i, u64, u64, i, u64, p, p]
        L.or_grid_checksum.argtypes = [u64, p, p]
        L.or_bfs2.argtypes = [u64, i, p, p, u64, p, p, i]
        L.or_traverse.argtypes = [u64, i, p, p, u64, ctypes.c_uint32, p, p, i]
        L.or_text_to_bin.argtypes = [ctypes.c_char_p, u64, p, u64, p, p, p]
        L.or_text_to_bin.restype = i64
        L.or_bfs_mt.argtypes = [u64, i, p, p, u64, p, p, i, i]
        L.or_wcc_mt.argtypes = [u64, i, p, p, p, p, i]
        L.or_kcore_mt.argtypes = [u64, p, p, p, i, p, p, p, i, p]
        L.or_grid_checksum.restype = u64
        L.or_bfs.argtypes = [u64, i, p, p, u64, p, p, i]
        L.or_bfs_check.argtypes = [u64, i, p, p, u64, p, i]
        L.or_bfs_check.restype = i64
        L.or_pr_step64_sample.argtypes = [u64, i, p, p, p, p, p, u64, i, p]
        L.or_wcc.argtypes = [u64, i, p, p, p, p]
        L.or_kcore.argtypes = [u64, p, p, p, i, p, p, p]
        L.or_total_tiles.argtypes = [u64, i]
        L.or_total_tiles.restype = u64
        L.or_tile_slot.argtypes = [u64, u64, u64, i]
        L.or_tile_slot.restype = u64
        L.or_groups.argtypes = [u64]
        L.or_groups.restype = u64
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


def kron(scale: int, e0: int, e1: int, seed1: int = 1, seed2: int = 2) -> np.ndarray:
    out = np.empty((e1 - e0, 2), np.uint32)
    lib().or_kron_range(scale, seed1, seed2, e0, e1, _p(out))
    return out


def convert(edges: np.ndarray, n: int, directed: bool = False):
    """-> (start u64[T+1], grid u32[E], degree i32[N]); in-tile order = input order."""
    edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    L = lib()
    T = L.or_total_tiles(256 * L.or_groups(n), int(directed))
    start = np.empty(T + 1, np.uint64)
    degree = np.empty(n, np.int32)
    E = L.or_convert_count(_p(edges), edges.shape[0], n, int(directed), _p(start), _p(degree))
    if E < 0:
        raise ValueError("edge list holds a vertex id >= N")
    grid = np.empty(E, np.uint32)
    L.or_convert_scatter(_p(edges), edges.shape[0], n, int(directed), _p(start), _p(grid))
    return start, grid, degree


def pagerank(n, start, grid, degree, iters=5, mode=0, nthreads=1, directed=False) -> np.ndarray:
    out = np.empty(n, np.float32)
    rc = lib().or_pagerank(n, int(directed), _p(start), _p(grid), _p(degree), iters, mode, nthreads, _p(out))
    if rc:
        raise RuntimeError("or_pagerank failed")
    return out


def pagerank64(n, start, grid, degree, iters=5, nthreads=1, directed=False) -> np.ndarray:
    """FP64 PageRank of the intended form (exact 1/deg factor)."""
    out = np.empty(n, np.float64)
    if lib().or_pagerank64(n, int(directed), _p(start), _p(grid), _p(degree), iters, nthreads, _p(out)):
        raise RuntimeError("or_pagerank64 failed")
    return out


def ppr(n, start, grid, degree, source, iters=20, directed=False, nthreads=1) -> np.ndarray:
    """FP64 personalised PageRank."""
    out = np.empty(n, np.float64)
    if lib().or_ppr(n, int(directed), _p(start), _p(grid), _p(degree), source, iters, nthreads, _p(out)):
        raise RuntimeError("or_ppr failed")
    return out


def kron_convert(scale: int, edge_factor: int = 16, seed1: int = 1, seed2: int = 2, directed: bool = False,
                 chunk: int = 1 << 26):
    """Generate and convert in chunks (the tuple list is never held whole)
    -> (start, grid, degree), equal to convert(kron(scale, 0, M))."""
    L = lib()
    n = 1 << scale
    T = L.or_total_tiles(256 * L.or_groups(n), int(directed))
    start = np.empty(T + 1, np.uint64)
    degree = np.empty(n, np.int32)
    E = L.or_kron_convert_count(scale, edge_factor, seed1, seed2, int(directed), chunk, _p(start), _p(degree))
    if E < 0:
        raise RuntimeError("or_kron_convert_count failed")
    grid = np.empty(E, np.uint32)
    if L.or_kron_convert_scatter(scale, edge_factor, seed1, seed2, int(directed), chunk, _p(start), _p(grid)):
        raise RuntimeError("or_kron_convert_scatter failed")
    return start, grid, degree


def grid_checksum(start, grid) -> int:
    """Order-independent checksum of the (tile, edge word) multiset."""
    return int(lib().or_grid_checksum(start.size - 1, _p(start), _p(grid)))


def bfs(n, start, grid, root, directed=False, nthreads=1):
    depth = np.empty(n, np.uint8)
    counts = np.zeros(256, np.uint64)
    if nthreads > 1:
        nl = lib().or_bfs_mt(n, int(directed), _p(start), _p(grid), root, _p(depth), _p(counts), 256, nthreads)
    else:
        nl = lib().or_bfs(n, int(directed), _p(start), _p(grid), root, _p(depth), _p(counts), 256)
    if nl < 0:
        raise RuntimeError(f"or_bfs failed ({nl})")
    return depth, [int(c) for c in counts[:nl]]


def bfs_check(n, start, grid, root, depth, directed=False, nthreads=1) -> int:
    """Violations of the conditions that fix BFS depths (0: depth is the BFS
    depth array from root)."""
    d = np.ascontiguousarray(depth, np.uint8)
    if d.size != n:
        raise ValueError("depth array of the wrong size")
    bad = lib().or_bfs_check(n, int(directed), _p(start), _p(grid), root, _p(d), nthreads)
    if bad < 0:
        raise RuntimeError("or_bfs_check failed")
    return int(bad)


def pr_step64_sample(n, start, grid, degree, prior, idx, directed=False, nthreads=1) -> np.ndarray:
    """One FP64 PageRank step (intended form) from the iterate `prior`,
    evaluated at the sorted distinct vertices idx."""
    pr = np.ascontiguousarray(prior, np.float64)
    ix = np.ascontiguousarray(idx, np.uint32)
    out = np.empty(ix.size, np.float64)
    if pr.size != n or lib().or_pr_step64_sample(n, int(directed), _p(start), _p(grid), _p(degree), _p(pr),
                                                 _p(ix), ix.size, nthreads, _p(out)):
        raise RuntimeError("or_pr_step64_sample failed")
    return out


def wcc(n, start, grid, directed=False, nthreads=1):
    label = np.empty(n, np.uint32)
    nc = np.zeros(1, np.uint64)
    if nthreads > 1:
        rc = lib().or_wcc_mt(n, int(directed), _p(start), _p(grid), _p(label), _p(nc), nthreads)
    else:
        rc = lib().or_wcc(n, int(directed), _p(start), _p(grid), _p(label), _p(nc))
    if rc:
        raise RuntimeError("or_wcc failed")
    return label, int(nc[0])


def kcore(n, start, grid, degree, k, nthreads=1, scanned=None):
    """-> (deg, count_ne0, count_pos, rounds); with nthreads > 1 and a list
    `scanned`, appends the stored edges its edge passes visited (tile skip rule)."""
    deg = np.empty(n, np.int32)
    c = np.zeros(3, np.uint64)
    if nthreads > 1:
        rounds = lib().or_kcore_mt(n, _p(start), _p(grid), _p(degree), k, _p(deg), _p(c[:1]),
                                   ctypes.c_void_p(c.ctypes.data + 8), nthreads, ctypes.c_void_p(c.ctypes.data + 16))
        if scanned is not None:
            scanned.append(int(c[2]))
    else:
        rounds = lib().or_kcore(n, _p(start), _p(grid), _p(degree), k, _p(deg), _p(c[:1]),
                                ctypes.c_void_p(c.ctypes.data + 8))
    if rounds < 0:
        raise RuntimeError("or_kcore failed")
    return deg, int(c[0]), int(c[1]), rounds


def bfs2(n, start, grid, root, directed=False):
    """-j 30: asynchronous BFS -> (depth u8, per-sweep update counts)."""
    depth = np.empty(n, np.uint8)
    counts = np.zeros(4096, np.uint64)
    ns = lib().or_bfs2(n, int(directed), _p(start), _p(grid), root, _p(depth), _p(counts), 4096)
    if ns < 0:
        raise RuntimeError(f"or_bfs2 failed ({ns})")
    return depth, [int(c) for c in counts[:ns]]


def traverse(n, start, grid, pivot, color=0, directed=False):
    """-j 4: single colour propagation -> (cid u32, per-sweep coloured counts)."""
    cid = np.empty(n, np.uint32)
    counts = np.zeros(1 << 16, np.uint64)
    ns = lib().or_traverse(n, int(directed), _p(start), _p(grid), pivot, color, _p(cid), _p(counts), 1 << 16)
    if ns < 0:
        raise RuntimeError(f"or_traverse failed ({ns})")
    return cid, [int(c) for c in counts[:ns]]


def text_to_bin(text: bytes):
    """-c 4: text edge list -> (pairs (M, 2) u32, v_min, v_max, tokens)."""
    L = lib()
    vmin, vmax, ntok = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    m = L.or_text_to_bin(text, len(text), None, 0, ctypes.byref(vmin), ctypes.byref(vmax), ctypes.byref(ntok))
    out = np.empty((m, 2), np.uint32)
    L.or_text_to_bin(text, len(text), _p(out), m, ctypes.byref(vmin), ctypes.byref(vmax), ctypes.byref(ntok))
    return out, vmin.value, vmax.value, ntok.value


def ref_generator():
    """The reference's generator library built from source, or None."""
    if not os.path.exists(REF_GEN):
        return None
    L = ctypes.CDLL(REF_GEN)
    L.make_mrg_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
    L.generate_kronecker_range.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_void_p]
    return L


def ref_kron(L, scale: int, e0: int, e1: int, seed1: int = 1, seed2: int = 2) -> np.ndarray:
    seed = (ctypes.c_uint64 * 5)()   # uint_fast32_t is 64-bit on x86-64 glibc
    L.make_mrg_seed(seed1, seed2, seed)
    out = np.empty((e1 - e0, 2), np.uint32)
    L.generate_kronecker_range(seed, scale, e0, e1, _p(out))
    return out
