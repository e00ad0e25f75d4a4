"""gstore_amd -- MI355X G-Store engine, Python host side.

A thin ctypes layer over the C ABI in include/gstore_amd.h (libgstore_amd.so,
built by gstore_amd.build).  It mirrors the reference's job surface
(gstore.cpp:235-333): ``bfs`` is -j 0/20, ``pagerank`` -j 1/21, ``kcore``
-j 2/22, ``wcc`` -j 3/23, with the reference's argument defaults (PageRank 5
iterations, BFS root 0, k = 10; gstore.cpp:2194-2196, 2307, 2353).

There is no CPU fallback: if the HIP library is missing, importing works but
every call raises GStoreError.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import tiles

__all__ = ["Graph", "Server", "QueryResult", "Q_BFS", "Q_PPR", "Q_WCC", "Q_PAGERANK", "GStoreError", "kron_edges", "convert", "text_to_bin", "text_file_to_bin", "load_library", "tiles", "LIB_PATH",
           "PR_COMPAT", "PR_INTENDED"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgstore_amd.so")

PR_COMPAT = 0
PR_INTENDED = 1
PR_SCATTER, PR_PULL, PR_BLOCKED = 0, 1, 2  # PageRank methods (set_pagerank_method)

GS_OK = 0
_STATUS = {1: "GS_ERR_ARG", 2: "GS_ERR_IO", 3: "GS_ERR_FORMAT", 4: "GS_ERR_NOMEM",
           5: "GS_ERR_DEVICE", 6: "GS_ERR_UNSUPPORTED", 7: "GS_ERR_DEPTH_OVERFLOW"}
KERNELS = ["pr_scatter", "pr_finalize", "bfs_expand", "bfs_level_end", "wcc_hook", "wcc_compress",
           "wcc_check", "kcore_peel", "kcore_scan", "kron_gen", "convert", "h2d_copy", "pr_pull", "index_build",
           "pb_gather", "pb_accum", "pb_reduce", "bfs_topdown", "bfs_bottomup"]


class GStoreError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class _Info(ctypes.Structure):
    _fields_ = [("vert_count", ctypes.c_uint64), ("edge_count", ctypes.c_uint64),
                ("tile_slots", ctypes.c_uint64), ("nonempty_tiles", ctypes.c_uint64),
                ("groups", ctypes.c_uint64), ("tiles_per_side", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("resident_edges", ctypes.c_uint64),
                ("segment_edges", ctypes.c_uint64), ("directed", ctypes.c_int32),
                ("resident", ctypes.c_int32), ("index_entries", ctypes.c_uint64),
                ("index_vertices", ctypes.c_uint64), ("index_ms", ctypes.c_double)]


class _Stats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double), ("iterations", ctypes.c_uint64),
                ("edges_scanned", ctypes.c_uint64), ("tiles_scanned", ctypes.c_uint64),
                ("h2d_bytes", ctypes.c_uint64), ("kern_ms", ctypes.c_double * len(KERNELS)),
                ("kern_calls", ctypes.c_uint64 * len(KERNELS))]


class _LevelStat(ctypes.Structure):
    _fields_ = [("level", ctypes.c_uint32), ("direction", ctypes.c_int32), ("frontier", ctypes.c_uint64),
                ("edges", ctypes.c_uint64), ("tiles", ctypes.c_uint64), ("reached", ctypes.c_uint64),
                ("ms", ctypes.c_double)]


class _Tune(ctypes.Structure):
    _fields_ = [("chunk", ctypes.c_uint32), ("xcd_order", ctypes.c_int32), ("from_cache", ctypes.c_int32),
                ("ncand", ctypes.c_int32), ("best_ms", ctypes.c_double), ("cand_chunk", ctypes.c_uint32 * 6),
                ("cand_order", ctypes.c_int32 * 6), ("cand_ms", ctypes.c_double * 6),
                ("do_alpha", ctypes.c_int32), ("do_beta", ctypes.c_int32), ("do_ms", ctypes.c_double)]


class _IoStats(ctypes.Structure):
    _fields_ = [("disk_bytes", ctypes.c_uint64), ("direct_bytes", ctypes.c_uint64), ("read_ms", ctypes.c_double),
                ("wait_ms", ctypes.c_double), ("io_threads", ctypes.c_int32), ("staging_buffers", ctypes.c_int32),
                ("external", ctypes.c_int32), ("read_mode", ctypes.c_int32)]


class QueryResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("arg", ctypes.c_uint64), ("value", ctypes.c_uint64),
                ("aux", ctypes.c_uint64), ("reserved", ctypes.c_uint64), ("t_submit_ms", ctypes.c_double),
                ("t_start_ms", ctypes.c_double), ("t_done_ms", ctypes.c_double), ("graph", ctypes.c_int32),
                ("kind", ctypes.c_int32), ("status", ctypes.c_int32), ("batch_size", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


Q_BFS, Q_PPR, Q_WCC, Q_PAGERANK = 0, 1, 2, 3

_lib: Optional[ctypes.CDLL] = None

_P = ctypes.c_void_p
_SIGS = {
    "gs_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "gs_open_arrays": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, _P, _P, _P, ctypes.POINTER(_P)]),
    "gs_open_kron": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(_P)]),
    "gs_kron_edges": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                     ctypes.c_int64, _P]),
    "gs_convert": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_char_p]),
    "gs_save": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gs_get_info": (ctypes.c_int, [_P, ctypes.POINTER(_Info)]),
    "gs_set_device_budget": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64]),
    "gs_get_arrays": (ctypes.c_int, [_P, _P, _P, _P]),
    "gs_get_host_grid": (ctypes.c_int, [_P, _P, _P]),
    "gs_open_external": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                        ctypes.c_uint64, ctypes.POINTER(_P)]),
    "gs_get_io_stats": (ctypes.c_int, [_P, ctypes.POINTER(_IoStats)]),
    "gs_pagerank": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _P]),
    "gs_ppr": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_int, _P]),
    "gs_bfs": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gs_bfs_batch": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gs_wcc": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "gs_kcore": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_uint64),
                                ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int)]),
    "gs_bfs_async": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gs_traverse": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint32, _P, _P, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]),
    "gs_text_to_bin": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint64)]),
    "gs_text_to_bin_mem": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, _P, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]),
    "gs_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(_Stats)]),
    "gs_get_level_stats": (ctypes.c_int, [_P, ctypes.POINTER(_LevelStat), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]),
    "gs_set_kernel_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_debug_fail_alloc": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_sweep_order": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_chunk": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gs_set_pagerank_heavy": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_pagerank_method": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_pagerank_blocking": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                                ctypes.c_uint32, ctypes.c_uint32]),
    "gs_set_wcc_method": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_pagerank_gather": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gs_set_kcore_method": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_batch_bfs_method": (ctypes.c_int, [_P, ctypes.c_int]),
    "gs_set_bfs_mode": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gs_tune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(_Tune)]),
    "gs_kernel_name": (ctypes.c_char_p, [ctypes.c_int]),
    "gs_close": (None, [_P]),
    "gs_server_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_P)]),
    "gs_server_add_graph": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int)]),
    "gs_server_set": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gs_server_submit": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _P,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "gs_server_poll": (ctypes.c_int, [_P, ctypes.POINTER(QueryResult), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int)]),
    "gs_server_drain": (ctypes.c_int, [_P]),
    "gs_server_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64)]),
    "gs_server_destroy": (None, [_P]),
    "gs_server_memory": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "gs_server_bfs_policy": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]),
    "gs_server_set_batch_min": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "gs_server_last_error": (ctypes.c_size_t, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "gs_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
}


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libgstore_amd.so (raises GStoreError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GStoreError(5, f"{path} is not built; run gstore_amd.build.build_all()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != GS_OK:
        buf = ctypes.create_string_buffer(1024)
        _lib.gs_last_error(buf, 1024)
        raise GStoreError(status, buf.value.decode(errors="replace"))


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class BfsResult:
    depth: np.ndarray          # u8, 0 = unreached, root = 1
    level_counts: list         # vertices reached at level 1, 2, ... (last entry 0)


class Graph:
    """A tile grid resident in HBM (one gs_graph handle)."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle
        info = _Info()
        _check(_lib.gs_get_info(self._h, ctypes.byref(info)))
        self.n = int(info.vert_count)
        self.E = int(info.edge_count)
        self.T = int(info.tile_slots)
        self.nonempty_tiles = int(info.nonempty_tiles)
        self.p = int(info.groups)
        self.p_s = int(info.tiles_per_side)
        self.directed = bool(info.directed)
        self.resident = bool(info.resident)
        self.resident_edges = int(info.resident_edges)
        self.segment_edges = int(info.segment_edges)

    def index_info(self) -> dict:
        """Neighbour-list index: entries (2E), vertices with edges, build ms (zeros until built)."""
        info = _Info()
        _check(_lib.gs_get_info(self._h, ctypes.byref(info)))
        return {"entries": int(info.index_entries), "vertices": int(info.index_vertices),
                "build_ms": float(info.index_ms), "device_bytes": int(info.device_bytes)}

    # --- construction -------------------------------------------------------
    @classmethod
    def open(cls, prefix: str, vert_count: int, directed: bool = False, device_budget: int = 0) -> "Graph":
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.gs_open(prefix.encode(), vert_count, int(directed), device_budget, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def open_external(cls, prefix: str, vert_count: int, device_budget: int, directed: bool = False,
                      segment_bytes: int = 0) -> "Graph":
        """External mode: the grid stays in PREFIX.grid and streams from the file
        every pass (beyond a resident prefix of the device budget)."""
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.gs_open_external(prefix.encode(), vert_count, int(directed), device_budget, segment_bytes,
                                     ctypes.byref(h)))
        return cls(h)

    def io_stats(self) -> dict:
        """File reads of an external-mode graph since it was opened."""
        st = _IoStats()
        _check(_lib.gs_get_io_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _IoStats._fields_}

    @classmethod
    def from_arrays(cls, n: int, start: np.ndarray, grid: np.ndarray, degree: np.ndarray,
                    directed: bool = False) -> "Graph":
        load_library()
        start = np.ascontiguousarray(start, dtype=np.uint64)
        grid = np.ascontiguousarray(grid, dtype=np.uint32)
        degree = np.ascontiguousarray(degree, dtype=np.int32)
        T = tiles.total_slots(tiles.tiles_per_side(n), directed)
        if start.size != T + 1 or degree.size != n or grid.size != int(start[-1]):
            raise GStoreError(1, f"array sizes do not match N={n}: start {start.size} (want {T + 1}), "
                                 f"grid {grid.size} (want {int(start[-1])}), degree {degree.size}")
        h = ctypes.c_void_p()
        _check(_lib.gs_open_arrays(n, int(directed), _ptr(start), _ptr(grid) if grid.size else None,
                                   _ptr(degree), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def kron(cls, scale: int, edge_factor: int = 16, seed1: int = 1, seed2: int = 2,
             directed: bool = False) -> "Graph":
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.gs_open_kron(scale, edge_factor, seed1, seed2, int(directed), ctypes.byref(h)))
        return cls(h)

    def close(self) -> None:
        if self._h:
            _lib.gs_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- data ---------------------------------------------------------------
    def arrays(self, with_grid: bool = True):
        """(start, grid, degree) copied to the host; grid is None with with_grid=False."""
        start = np.empty(self.T + 1, np.uint64)
        grid = np.empty(self.E, np.uint32) if with_grid else None
        degree = np.empty(self.n, np.int32)
        _check(_lib.gs_get_arrays(self._h, _ptr(start), _ptr(grid) if with_grid and self.E else None,
                                  _ptr(degree)))
        return start, grid, degree

    def host_grid(self) -> np.ndarray:
        """Semi-external mode: a read-only view of the pinned host grid (no copy),
        valid while the graph is open and its device budget unchanged."""
        p = ctypes.c_void_p()
        e = ctypes.c_uint64(0)
        _check(_lib.gs_get_host_grid(self._h, ctypes.byref(p), ctypes.byref(e)))
        if not e.value:
            return np.empty(0, np.uint32)
        view = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), shape=(e.value,))
        view.flags.writeable = False
        return view

    def set_device_budget(self, budget_bytes: int, segment_bytes: int = 0) -> None:
        """Semi-external mode: keep at most budget_bytes of the grid in HBM."""
        _check(_lib.gs_set_device_budget(self._h, budget_bytes, segment_bytes))
        self.__init__(self._h)

    def degree(self) -> np.ndarray:
        degree = np.empty(self.n, np.int32)
        _check(_lib.gs_get_arrays(self._h, None, None, _ptr(degree)))
        return degree

    def save(self, prefix: str) -> None:
        _check(_lib.gs_save(self._h, prefix.encode()))

    def tune(self, cache_path: Optional[str] = None) -> dict:
        """Autotune edges-per-workgroup and sweep order (gs_tune)."""
        r = _Tune()
        _check(_lib.gs_tune(self._h, cache_path.encode() if cache_path else None, ctypes.byref(r)))
        return {"chunk": r.chunk, "xcd_order": r.xcd_order, "from_cache": bool(r.from_cache),
                "best_ms": r.best_ms, "do_alpha": r.do_alpha, "do_beta": r.do_beta, "do_ms": r.do_ms,
                "candidates": [{"chunk": r.cand_chunk[i], "xcd_order": r.cand_order[i], "ms": r.cand_ms[i]}
                               for i in range(r.ncand)]}

    def set_bfs_mode(self, mode: int, alpha: int = 0, beta: int = 0) -> None:
        """0: tile sweep; 1: direction-optimizing over a neighbour-list index
        (alpha/beta switch thresholds; 0 keeps the current ones)."""
        _check(_lib.gs_set_bfs_mode(self._h, mode, alpha, beta))

    def set_pagerank_method(self, method) -> None:
        """PR_BLOCKED (2, default): two streamed passes over the blocked copy of the
        neighbour-list index; PR_PULL (1): pull over the lists; PR_SCATTER (0):
        tile scatter with atomics.  On directed and semi-external grids 1 and 2
        select the binned scatter (no global atomics), 0 the atomic one.
        True / False mean PR_PULL / PR_SCATTER."""
        _check(_lib.gs_set_pagerank_method(self._h, int(method)))

    def set_pagerank_blocking(self, gather_entries: int = 1 << 16, gather_runs: int = 2048,
                              accum_entries: int = 1 << 21, dense_entries: int = 1 << 30,
                              compact_permille: int = 800) -> None:
        """Work split and layout of the blocked PageRank (small values exercise the split paths)."""
        _check(_lib.gs_set_pagerank_blocking(self._h, gather_entries, gather_runs, accum_entries, dense_entries,
                                             compact_permille))

    def set_pagerank_gather(self, units_per_grab: int) -> None:
        """Blocked PageRank gather: 0 = a workgroup per unit; n = resident workgroups taking n units per grab."""
        _check(_lib.gs_set_pagerank_gather(self._h, units_per_grab))

    def set_wcc_method(self, method: int) -> None:
        """1 (default): sampled linking over the neighbour-list index; 0: hook passes over the tiles."""
        _check(_lib.gs_set_wcc_method(self._h, int(method)))

    def level_stats(self) -> list:
        """Per-level record of the last BFS: direction (tile / top-down / bottom-up),
        frontier, edges, tiles, reached, ms."""
        buf = (_LevelStat * 256)()
        n = ctypes.c_int(0)
        _check(_lib.gs_get_level_stats(self._h, buf, 256, ctypes.byref(n)))
        names = {0: "tile", 1: "top-down", 2: "bottom-up"}
        return [{"level": int(b.level), "direction": names.get(int(b.direction), "?"), "frontier": int(b.frontier),
                 "edges": int(b.edges), "tiles": int(b.tiles), "reached": int(b.reached), "ms": float(b.ms)}
                for b in buf[:min(n.value, 256)]]

    def set_batch_bfs_method(self, method: int) -> None:
        """2 (default): lanes over the neighbour-list index, levels driven from the
        device, four levels per hipGraph launch; 1: the same with a host round trip
        per level; 0: lanes over the tile sweep."""
        _check(_lib.gs_set_batch_bfs_method(self._h, int(method)))

    def set_kcore_method(self, method: int) -> None:
        """1 (default): push the removed vertices' lists over the index; 0: rescan active tiles."""
        _check(_lib.gs_set_kcore_method(self._h, int(method)))

    def set_pagerank_heavy(self, on: bool) -> None:
        _check(_lib.gs_set_pagerank_heavy(self._h, int(on)))

    def set_chunk(self, edges: int) -> None:
        _check(_lib.gs_set_chunk(self._h, edges))

    def set_sweep_order(self, xcd_contiguous: bool) -> None:
        _check(_lib.gs_set_sweep_order(self._h, int(xcd_contiguous)))

    def debug_fail_alloc(self, nth: int) -> None:
        """Test hook: the nth device allocation of the next index build fails."""
        _check(_lib.gs_debug_fail_alloc(self._h, nth))

    def set_kernel_timing(self, on: bool) -> None:
        _check(_lib.gs_set_kernel_timing(self._h, int(on)))

    def stats(self) -> dict:
        st = _Stats()
        _check(_lib.gs_get_stats(self._h, ctypes.byref(st)))
        return {"ms": st.ms, "iterations": int(st.iterations), "edges_scanned": int(st.edges_scanned),
                "tiles_scanned": int(st.tiles_scanned), "h2d_bytes": int(st.h2d_bytes),
                "kernels": {KERNELS[i]: {"ms": st.kern_ms[i], "calls": int(st.kern_calls[i])}
                            for i in range(len(KERNELS)) if st.kern_calls[i]}}

    # --- jobs -----------------------------------------------------------------
    def pagerank(self, iters: int = 5, mode: int = PR_COMPAT, copy_out: bool = True) -> Optional[np.ndarray]:
        out = np.empty(self.n, np.float32) if copy_out else None
        _check(_lib.gs_pagerank(self._h, iters, mode, _ptr(out) if copy_out else None))
        return out

    def ppr(self, source: int, iters: int = 20, copy_out: bool = True) -> Optional[np.ndarray]:
        """Personalised PageRank from `source` (FP32 ranks, damping 0.85)."""
        out = np.empty(self.n, np.float32) if copy_out else None
        _check(_lib.gs_ppr(self._h, source, iters, _ptr(out) if copy_out else None))
        return out

    def bfs(self, root: int = 0, copy_out: bool = True, max_levels: int = 256) -> BfsResult:
        depth = np.empty(self.n, np.uint8) if copy_out else None
        counts = np.zeros(max_levels, np.uint64)
        nlev = ctypes.c_int(0)
        _check(_lib.gs_bfs(self._h, root, _ptr(depth) if copy_out else None, _ptr(counts), max_levels,
                           ctypes.byref(nlev)))
        return BfsResult(depth, [int(c) for c in counts[:min(nlev.value, max_levels)]])

    def bfs_batch(self, roots, copy_out: bool = True, max_levels: int = 256):
        """Up to 64 roots in one multi-source sweep -> (depth[n_roots, N] u8, level_counts per root)."""
        r = np.ascontiguousarray(roots, dtype=np.uint64)
        depth = np.empty((r.size, self.n), np.uint8) if copy_out else None
        counts = np.zeros((r.size, max_levels), np.uint64)
        nlev = ctypes.c_int(0)
        _check(_lib.gs_bfs_batch(self._h, _ptr(r), r.size, _ptr(depth) if copy_out else None, _ptr(counts),
                                 max_levels, ctypes.byref(nlev)))
        lv = min(nlev.value, max_levels)
        return depth, [[int(c) for c in counts[i, :lv]] for i in range(r.size)]

    def wcc(self, copy_out: bool = True):
        labels = np.empty(self.n, np.uint32) if copy_out else None
        nc = ctypes.c_uint64(0)
        _check(_lib.gs_wcc(self._h, _ptr(labels) if copy_out else None, ctypes.byref(nc)))
        return labels, int(nc.value)

    def bfs_async(self, root: int = 1, copy_out: bool = True, max_sweeps: int = 4096):
        """-j 30: asynchronous BFS -> (depth u8 or None, per-sweep store counts)."""
        depth = np.empty(self.n, np.uint8) if copy_out else None
        counts = np.zeros(max_sweeps, np.uint64)
        ns = ctypes.c_int(0)
        _check(_lib.gs_bfs_async(self._h, root, _ptr(depth) if copy_out else None, _ptr(counts), max_sweeps,
                                 ctypes.byref(ns)))
        return depth, [int(c) for c in counts[:min(ns.value, max_sweeps)]]

    def traverse(self, pivot: int = 1, color: int = 0, copy_out: bool = True, max_sweeps: int = 1 << 16):
        """-j 4: single colour propagation -> (cid u32 or None, per-sweep counts, coloured vertices)."""
        cid = np.empty(self.n, np.uint32) if copy_out else None
        counts = np.zeros(max_sweeps, np.uint64)
        ns, col = ctypes.c_int(0), ctypes.c_uint64(0)
        _check(_lib.gs_traverse(self._h, pivot, color, _ptr(cid) if copy_out else None, _ptr(counts), max_sweeps,
                                ctypes.byref(ns), ctypes.byref(col)))
        return cid, [int(c) for c in counts[:min(ns.value, max_sweeps)]], int(col.value)

    def kcore(self, k: int = 10, copy_out: bool = True):
        deg = np.empty(self.n, np.int32) if copy_out else None
        ne0, pos, rounds = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_int(0)
        _check(_lib.gs_kcore(self._h, k, _ptr(deg) if copy_out else None, ctypes.byref(ne0), ctypes.byref(pos),
                             ctypes.byref(rounds)))
        return deg, int(ne0.value), int(pos.value), int(rounds.value)


class Server:
    """Continuous-batching query server over resident graphs (gs_server_*)."""

    def __init__(self, hbm_budget: int = 0, workers: int = 4):
        load_library()
        self._h = ctypes.c_void_p()
        _check(_lib.gs_server_create(hbm_budget, workers, ctypes.byref(self._h)))
        self._graphs = []
        self._keep = {}   # query id -> output buffer kept alive until polled

    def add_graph(self, g: "Graph") -> int:
        gid = ctypes.c_int(0)
        _check(_lib.gs_server_add_graph(self._h, g._h, ctypes.byref(gid)))
        self._graphs.append(g)
        return gid.value

    def configure(self, max_batch: int = 64, ppr_iters: int = 20, pr_iters: int = 20) -> None:
        _check(_lib.gs_server_set(self._h, max_batch, ppr_iters, pr_iters))

    def submit(self, graph: int, kind: int, arg: int = 0, out: Optional[np.ndarray] = None) -> int:
        qid = ctypes.c_uint64(0)
        _check(_lib.gs_server_submit(self._h, graph, kind, arg, _ptr(out) if out is not None else None,
                                     ctypes.byref(qid)))
        if out is not None:
            self._keep[qid.value] = out
        return qid.value

    def poll(self, max_results: int = 4096, wait_ms: int = 0) -> list:
        buf = (QueryResult * max_results)()
        n = ctypes.c_int(0)
        _check(_lib.gs_server_poll(self._h, buf, max_results, wait_ms, ctypes.byref(n)))
        out = [buf[i].as_dict() for i in range(n.value)]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def drain(self) -> None:
        _check(_lib.gs_server_drain(self._h))

    def info(self) -> dict:
        h, p, i = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.gs_server_info(self._h, ctypes.byref(h), ctypes.byref(p), ctypes.byref(i)))
        return {"held_bytes": h.value, "pending": p.value, "inflight": i.value}

    def memory(self) -> dict:
        """Workspace bytes held by queries (bounded by the HBM budget) and the graphs' own bytes."""
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.gs_server_memory(self._h, ctypes.byref(a), ctypes.byref(b)))
        return {"workspace_bytes": a.value, "graph_bytes": b.value}

    def bfs_policy(self, graph: int) -> dict:
        """Batching rule measured when the graph was added."""
        k = ctypes.c_int(0)
        a, b, c = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        _check(_lib.gs_server_bfs_policy(self._h, graph, ctypes.byref(k), ctypes.byref(a), ctypes.byref(b),
                                         ctypes.byref(c)))
        return {"batch_min": k.value, "single_ms": a.value, "batch1_ms": b.value, "batch64_ms": c.value}

    def set_batch_min(self, graph: int, batch_min: int) -> None:
        _check(_lib.gs_server_set_batch_min(self._h, graph, batch_min))

    def last_error(self) -> str:
        buf = ctypes.create_string_buffer(1024)
        _lib.gs_server_last_error(self._h, buf, 1024)
        return buf.value.decode(errors="replace")

    def close(self) -> None:
        if self._h:
            _lib.gs_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def text_to_bin(text: bytes):
    """-c 4 on the device: text edge list -> (pairs (M, 2) u32, v_min, v_max, tokens)."""
    load_library()
    pairs, vmin, vmax, ntok = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    _check(_lib.gs_text_to_bin_mem(text, len(text), None, 0, ctypes.byref(pairs), ctypes.byref(vmin),
                                   ctypes.byref(vmax), ctypes.byref(ntok)))
    out = np.empty((pairs.value, 2), np.uint32)
    if pairs.value:
        _check(_lib.gs_text_to_bin_mem(text, len(text), _ptr(out), pairs.value, ctypes.byref(pairs),
                                       ctypes.byref(vmin), ctypes.byref(vmax), ctypes.byref(ntok)))
    return out, vmin.value, vmax.value, ntok.value


def text_file_to_bin(path: str, out_path: Optional[str] = None) -> dict:
    """-c 4 on a file: writes out_path (default path + '.edge')."""
    load_library()
    pairs, vmin, vmax, ntok = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    _check(_lib.gs_text_to_bin(path.encode(), out_path.encode() if out_path else None, ctypes.byref(pairs),
                               ctypes.byref(vmin), ctypes.byref(vmax), ctypes.byref(ntok)))
    return {"pairs": pairs.value, "vmin": vmin.value, "vmax": vmax.value, "tokens": ntok.value}


def kron_edges(scale: int, e0: int, e1: int, seed1: int = 1, seed2: int = 2) -> np.ndarray:
    """Graph500 tuples [e0, e1) generated on the device, shape (e1-e0, 2) u32."""
    load_library()
    out = np.empty((e1 - e0, 2), np.uint32)
    _check(_lib.gs_kron_edges(scale, seed1, seed2, e0, e1, _ptr(out)))
    return out


def convert(edges: np.ndarray, vert_count: int, prefix: str, directed: bool = False) -> None:
    """-c 1: tuples (M, 2) u32 -> PREFIX.grid/.start/.degree, converted on the device."""
    load_library()
    edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    _check(_lib.gs_convert(_ptr(edges) if edges.size else None, edges.shape[0], vert_count, int(directed),
                           prefix.encode()))
