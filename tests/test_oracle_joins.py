"""Block joins through every pair of faces (getSwapIndex, xyExchange and
edgeExchange, clGhostRoutines.c:47-296, :407-538), pinned geometrically:
for each of the sixteen (this face, that face) pairs, block 1 is a turned
copy of the unit-square grid beside block 0, and

  * the ghost vertices the host builds beyond the join continue each
    block's own grid (vertex form of the index map);
  * the oracle's edgeExchange fills each ghost cell from the cell of the
    other block with the same centre (cell form of the index map): block 1's
    f is set to a code of the cell centre and read back through the map.

These hold for the oriented joins whose end points come in opposite order
too (the row flip before the map)."""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O
from pyugks_amd import bc as bcd
from pyugks_amd import joins
from pyugks_amd.data import ghost_coordinates

PAIRS = list(itertools.product(range(4), range(4)))


def _turn_point(p, q, n, G, turns, shift):
    x, y = (p - G) / n - 0.5, (q - G) / n - 0.5
    for _ in range(turns % 4):
        x, y = -y, x
    return x + 0.5 + shift[0], y + 0.5 + shift[1]


def _strip(face, NI, NJ, G, vertices):
    """Padded indices of the ghost strip beyond a face (vertices: the rows
    xyExchange writes; cells: the ghost cells edgeExchange fills)."""
    e = 1 if vertices else 0
    if face == bcd.NORTH:
        return [(i, j) for i in range(G, NI - G + e) for j in range(NJ - G + e, NJ + e)]
    if face == bcd.SOUTH:
        return [(i, j) for i in range(G, NI - G + e) for j in range(0, G)]
    if face == bcd.EAST:
        return [(i, j) for i in range(NI - G + e, NI + e) for j in range(G, NJ - G + e)]
    return [(i, j) for i in range(0, G) for j in range(G, NJ - G + e)]


@pytest.mark.parametrize("method", ["vanLeer", "WENO5"])
@pytest.mark.parametrize("this_face,turns", PAIRS)
def test_ghost_vertices_continue_the_grid(this_face, turns, method):
    n = 5
    g, that_face = joins.joined_pair(this_face, turns, n=n, flux_method=method)
    G = 3 if method == "WENO5" else 2
    xy = ghost_coordinates(g.blocks, G)
    NI = NJ = n + 2 * G
    for blk, face, t, shift in ((0, this_face, 0, (0.0, 0.0)), (1, that_face, turns, joins._SHIFT[this_face])):
        for p, q in _strip(face, NI, NJ, G, True):
            ex = _turn_point(p, q, n, G, t, shift)
            assert np.allclose(xy[blk][p, q], ex, atol=1e-13, rtol=0), (this_face, that_face, blk, p, q)


@pytest.mark.parametrize("this_face,turns", PAIRS)
def test_exchange_fills_ghost_cells_from_the_matching_cells(this_face, turns):
    n = 5
    g, that_face = joins.joined_pair(this_face, turns, n=n, flux_method="WENO5")
    o = O.OracleSim(g, threads=1)
    a, b = o.blocks
    G = 3
    NI = NJ = n + 2 * G
    code = lambda c: c[..., 0] + 7.0 * c[..., 1]
    fb = b["f"]
    fb[..., 0] = code(b["centre"])[..., None]
    fb[..., 1] = -code(b["centre"])[..., None]
    L = O.lib()
    L.or_edgeExchange.argtypes = [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 2 + [C.c_int] * 5
    ori = g.blocks[0].bc_list[this_face].orientation
    L.or_edgeExchange(C.addressof(a.block.p), a["f"].ctypes.data, a["xy"].ctypes.data, this_face, fb.ctypes.data,
                      b["xy"].ctypes.data, NI, NJ, that_face, ori, bcd.NO_FLIP)
    for p, q in _strip(this_face, NI, NJ, G, False):
        cx, cy = _turn_point(p + 0.5, q + 0.5, n, G, 0, (0.0, 0.0))
        assert np.allclose(a["f"][p, q, :, 0], cx + 7.0 * cy, atol=1e-12), (this_face, that_face, p, q)
        assert np.allclose(a["f"][p, q, :, 1], -(cx + 7.0 * cy), atol=1e-12)
