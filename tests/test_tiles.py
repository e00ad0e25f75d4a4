"""Tile geometry: closed forms against enumeration, host helpers against the
oracle's C implementation.  CPU only."""
import numpy as np
import pytest

import oracle
from gstore_amd import tiles


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("directed", [False, True])
def test_slot_is_enumeration_index(p, directed):
    r, c = tiles.slot_coords(p, directed)
    assert r.size == tiles.total_slots(256 * p, directed)
    rng = np.random.default_rng(p)
    for t in rng.integers(0, r.size, 4000).tolist() + [0, r.size - 1]:
        assert tiles.tile_slot(int(r[t]), int(c[t]), p, directed) == t
    if not directed:
        assert np.all(r <= c)


@pytest.mark.parametrize("p", [1, 2, 5, 16])
def test_group_offsets_closed_form(p):
    """beg_edge_offset1/2 (gstore.h:128-141): groups before (i, j) in upper-
    triangle order, 65536 slots per off-diagonal and 32896 per diagonal group."""
    for i in range(p):
        for j in range(i, p):
            before = [(a, b) for a in range(p) for b in range(a, p) if (a, b) < (i, j)]
            want = sum(32896 if a == b else 65536 for a, b in before)
            assert tiles.tile_slot(256 * i, 256 * j, p) == want
            if p > 5:
                break


def test_python_and_c_geometry_agree():
    L = oracle.lib()
    rng = np.random.default_rng(5)
    for p in (1, 2, 7, 64):
        for directed in (0, 1):
            assert L.or_total_tiles(256 * p, directed) == tiles.total_slots(256 * p, bool(directed))
            for _ in range(300):
                r, c = sorted(rng.integers(0, 256 * p, 2).tolist())
                if directed and rng.random() < 0.5:
                    r, c = c, r
                assert L.or_tile_slot(r, c, p, directed) == tiles.tile_slot(r, c, p, bool(directed))


@pytest.mark.parametrize("n,p", [(1, 1), (1 << 20, 1), ((1 << 24), 1), ((1 << 24) + 1, 2), (1 << 25, 2),
                                 (1 << 28, 16), (1 << 30, 64)])
def test_groups(n, p):
    assert tiles.groups(n) == p == oracle.lib().or_groups(n)


def test_survey_table_sizes():
    # T for s25/s28/s30: 131,328 / 8,390,656 / 134,225,920
    assert tiles.total_slots(tiles.tiles_per_side(1 << 25)) == 131_328
    assert tiles.total_slots(tiles.tiles_per_side(1 << 28)) == 8_390_656
    assert tiles.total_slots(tiles.tiles_per_side(1 << 30)) == 134_225_920
