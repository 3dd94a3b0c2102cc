"""Two-block cases joined through any pair of faces.

``joined_pair(this_face, turns, ...)`` builds block 0 on the unit square (i
along +x, j along +y) and block 1 on the neighbouring unit square beyond
block 0's ``this_face``, its grid turned by ``turns`` quarter turns
(counter-clockwise, so every cell keeps a positive area).  The face of
block 1 that lands on the join decides the pair (this_face, that_face);
the four turns of each of the four faces give all sixteen pairs of
getSwapIndex (clGhostRoutines.c:47-138), half of them with the join's end
points in opposite order (the misaligned case of edgeExchange, :498-512).
The orientation flag is the one such a conforming join needs: a face
joined to the same face of the other block (N-N, E-E, ...) runs opposite
and is flipped (NO_HOLD); a quarter-turn join (N-W, E-S, ...) whose end
points run opposite is already reversed by the quarter turn of the index
map, so it is HOLD.  Every other face is a diffuse wall, the top one moving."""
from __future__ import annotations

import copy

import numpy as np

from . import bc as bcd
from .data import BlockDef, FlowCondition, UGKSData

_SHIFT = {bcd.NORTH: (0.0, 1.0), bcd.EAST: (1.0, 0.0), bcd.SOUTH: (0.0, -1.0), bcd.WEST: (-1.0, 0.0)}


def _face_midpoints(x, y):
    return {bcd.NORTH: (0.5 * (x[0, -1] + x[-1, -1]), 0.5 * (y[0, -1] + y[-1, -1])),
            bcd.EAST: (0.5 * (x[-1, 0] + x[-1, -1]), 0.5 * (y[-1, 0] + y[-1, -1])),
            bcd.SOUTH: (0.5 * (x[0, 0] + x[-1, 0]), 0.5 * (y[0, 0] + y[-1, 0])),
            bcd.WEST: (0.5 * (x[0, 0] + x[0, -1]), 0.5 * (y[0, 0] + y[0, -1]))}


def turned_grid(n: int, turns: int, shift=(0.0, 0.0)):
    """Vertices of an n x n grid of the unit square turned by ``turns``
    quarter turns about its centre, then shifted."""
    k = np.arange(n + 1, dtype=np.float64) / n
    I, J = np.meshgrid(k, k, indexing="ij")
    x, y = I - 0.5, J - 0.5
    for _ in range(turns % 4):
        x, y = -y, x
    return x + 0.5 + shift[0], y + 0.5 + shift[1]


def joined_pair(this_face: int, turns: int, n: int = 6, flux_method: str = "vanLeer", nc: int = 8,
                Kn: float = 0.05, lid: float = 0.2, wall_T: float = 1.0) -> tuple:
    """(UGKSData, that_face) for block 0's ``this_face`` joined to block 1."""
    from .cases import _base, newton_cotes
    g = _base(flux_method, Kn)
    newton_cotes(g, nc)
    x0, y0 = turned_grid(n, 0)
    x1, y1 = turned_grid(n, turns, _SHIFT[this_face])
    jx, jy = _face_midpoints(x0, y0)[this_face]
    mids = _face_midpoints(x1, y1)
    that_face = min(mids, key=lambda f: (mids[f][0] - jx) ** 2 + (mids[f][1] - jy) ** 2)
    top = max(np.max(y0), np.max(y1))
    ends = {bcd.NORTH: ((0, -1), (-1, -1)), bcd.EAST: ((-1, 0), (-1, -1)), bcd.SOUTH: ((0, 0), (-1, 0)),
            bcd.WEST: ((0, 0), (0, -1))}
    a1 = (x0[ends[this_face][0]], y0[ends[this_face][0]])
    b1 = (x1[ends[that_face][0]], y1[ends[that_face][0]])
    reversed_ends = abs(a1[0] - b1[0]) + abs(a1[1] - b1[1]) > 1e-12
    quarter = (this_face in (bcd.NORTH, bcd.SOUTH)) != (that_face in (bcd.NORTH, bcd.SOUTH))
    ori = bcd.HOLD if (reversed_ends and quarter) else bcd.NO_HOLD

    def walls(x, y, joined, other, other_face):
        out = []
        m = _face_midpoints(x, y)
        for f in (bcd.NORTH, bcd.EAST, bcd.SOUTH, bcd.WEST):
            if f == joined:
                out.append(bcd.AdjacentBC(other, other_face, orientation=ori))
            else:
                # the wall at the top of the two squares moves along +x (global frame)
                moving = abs(m[f][1] - top) < 1e-12 and abs(y[0, 0] - y[-1, -1]) > 0
                out.append(bcd.DiffuseBC(D=1.0, T=wall_T, U=lid if moving else 0.0, V=0.0))
        return out
    fill = FlowCondition(D=1.0, U=0.0, V=0.0, T=1.0)
    b0 = BlockDef(x=x0, y=y0, bc_list=walls(x0, y0, this_face, 1, that_face), fill_condition=copy.deepcopy(fill),
                  label="pair-0")
    b1 = BlockDef(x=x1, y=y1, bc_list=walls(x1, y1, that_face, 0, this_face), fill_condition=copy.deepcopy(fill),
                  label="pair-1")
    g.blocks = [b0, b1]
    return g, that_face
