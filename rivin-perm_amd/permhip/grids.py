"""Biadjacency matrices of Cartesian grid graphs (dimer coverings).

Vertices are the points of range(shape[0]) x ... x range(shape[-1]); the
bipartition is by parity of the coordinate sum, rows are the even points and
columns the odd points, each in lexicographic order.  Two points are joined
when they differ by one along a single axis, or wrap around an axis listed in
``periodic_axes`` (which must have even length >= 4 so the wrap keeps the
bipartition and adds no parallel bond).  The permanent counts perfect
matchings, i.e. dimer coverings (the reference's
perm_binary_applications.grid_biadjacency, perm_binary_applications.py:106-162;
row/column order does not change the permanent).
"""

from __future__ import annotations

from itertools import product
from typing import Iterable, List, Sequence


def grid_biadjacency(shape: Sequence[int], *, periodic_axes: Iterable[int] = ()) -> List[List[int]]:
    shape = tuple(int(s) for s in shape)
    if not shape or any(s < 1 for s in shape):
        raise ValueError("shape must be a non-empty sequence of positive extents")
    periodic = set(int(a) for a in periodic_axes)
    for a in periodic:
        if not 0 <= a < len(shape):
            raise ValueError("periodic axis out of range")
        if shape[a] % 2 or shape[a] < 4:
            raise ValueError("a periodic axis must have even length at least 4")
    points = list(product(*(range(s) for s in shape)))
    even = [pt for pt in points if sum(pt) % 2 == 0]
    odd = [pt for pt in points if sum(pt) % 2 == 1]
    if len(even) != len(odd):
        raise ValueError("unequal parity classes: no square biadjacency matrix")
    col_of = {pt: j for j, pt in enumerate(odd)}
    matrix = [[0] * len(odd) for _ in even]
    for i, pt in enumerate(even):
        for axis, extent in enumerate(shape):
            for step in (-1, 1):
                q = list(pt)
                q[axis] += step
                if not 0 <= q[axis] < extent:
                    if axis not in periodic:
                        continue
                    q[axis] %= extent
                matrix[i][col_of[tuple(q)]] = 1
    return matrix
