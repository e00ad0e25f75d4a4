"""Tile-grid geometry of the G-Store file format, host side.

Restates gstore.h:63-141 (undirected upper-triangle grid, HALF_GRID) and
gstore.h:176-198 (directed full grid): vertices are cut into 2^16-wide tiles,
256 x 256 tiles form a physical group, and .start lists one u64 offset per
tile slot (plus a sentinel) in group-row-major order.
"""
from __future__ import annotations

import numpy as np

TILE_BITS = 16      # bit_shift2, gstore.h:77
GROUP_TILES = 256   # p_p, gstore.h:89
GROUP_BITS = 24     # bit_shift1, gstore.h:96


def groups(n: int) -> int:
    """Physical groups per side, ceil(N / 2^24) (gstore.cpp:152-157)."""
    return (n >> GROUP_BITS) + (1 if n & ((1 << GROUP_BITS) - 1) else 0)


def tiles_per_side(n: int) -> int:
    return GROUP_TILES * groups(n)


def total_slots(p_s: int, directed: bool = False) -> int:
    """calc_total_part (gstore.h:121-125)."""
    return p_s * p_s if directed else p_s * (p_s + 1) // 2


def _tri(i: int, j: int, n: int) -> int:
    return (i * (2 * n - i - 1)) // 2 + j   # calc_index, gstore.h:109-113


def tile_slot(r: int, c: int, p: int, directed: bool = False) -> int:
    """.start index of tile (r, c) in global tile coordinates (r <= c if undirected)."""
    bi, si, bj, sj = r >> 8, r & 255, c >> 8, c & 255
    if directed:
        return ((bi * p + bj) << 16) + si * 256 + sj
    if bi == bj:
        return ((_tri(bi, bi, p) - bi) << 16) + ((bi * 257) << 7) + _tri(si, sj, 256)
    return ((_tri(bi, bj, p) - bi - 1) << 16) + (((bi + 1) * 257) << 7) + si * 256 + sj


def slot_coords(p: int, directed: bool = False):
    """(row, col) arrays of every tile slot in .start order."""
    rows, cols = [], []
    ar = np.arange(256, dtype=np.uint32)
    for bi in range(p):
        for bj in (range(p) if directed else range(bi, p)):
            if not directed and bi == bj:
                si, sj = np.triu_indices(256)
            else:
                si, sj = np.repeat(ar, 256), np.tile(ar, 256)
            rows.append(bi * 256 + si.astype(np.uint32))
            cols.append(bj * 256 + sj.astype(np.uint32))
    return np.concatenate(rows).astype(np.uint32), np.concatenate(cols).astype(np.uint32)


def decode(n: int, start: np.ndarray, grid: np.ndarray, directed: bool = False):
    """Global (v0, v1) of every stored edge, in file order."""
    p = groups(n)
    r, c = slot_coords(p, directed)
    counts = np.diff(start.astype(np.int64))
    tr = np.repeat(r.astype(np.uint64), counts)
    tc = np.repeat(c.astype(np.uint64), counts)
    g = grid.astype(np.uint64)
    return (tr << 16) | (g & 0xFFFF), (tc << 16) | (g >> 16)
