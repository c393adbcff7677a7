"""Triangle meshes of Manhattan copper (rectangles on one interface) for the RWG basis.

A tensor grid: along each axis the copper's edges, the lines that ports and elements must follow and the via axes
are breakpoints; an interval between two edges of a strip no wider than `narrow` gets ncross cells graded toward
both edges (the current crowds there), every other interval cells of at most dmax. The grid cells whose centres lie
in the copper become two triangles each, the diagonal alternating cell to cell (a chequerboard, so no direction is
preferred), counter-clockwise. A via axis becomes a grid line only where the grid has none already, so boards
without vias mesh the same with or without them nearby.

The reference meshes arbitrary polygons with a constrained Delaunay mesher (shapely + CDT) and rectangles with a
tensor grid of the same kind (src/mom25d/multilayer.py:322-349); polygons are outside this build's scope.
"""
import numpy as np

from .base import graded

TOL = 1e-9           # coordinates closer than this (m) are the same line


def _distinct(vals):
    """Sorted values with those closer than TOL to their predecessor dropped."""
    v = np.sort(np.round(np.asarray(vals, float).ravel(), 12))
    if not len(v):
        return v
    keep = np.concatenate([[True], np.diff(v) > TOL])
    return v[keep]


def grid_axis(edges, forced, dmax, ncross, narrow, grade=1.0):
    """Grid coordinates along one axis from the copper's edge coordinates and the forced lines."""
    e, f = _distinct(edges), _distinct(forced)
    coords = [e[:1]]
    for a, b in zip(e[:-1], e[1:]):
        cuts = f[(f > a + TOL) & (f < b - TOL)]
        if not len(cuts) and b - a <= narrow + 1e-12:
            coords.append(a + (b - a) * graded(np.linspace(0, 1, ncross + 1)[1:], grade))
            continue
        stops = np.concatenate([[a], cuts, [b]])
        for c, d in zip(stops[:-1], stops[1:]):
            n = max(1, int(np.ceil((d - c) / dmax - 1e-9)))
            coords.append(np.linspace(c, d, n + 1)[1:])
    return np.concatenate(coords)


def _lines(edges, forced, want, dmax, ncross, narrow, grade):
    """grid_axis with every coordinate of `want` a line: one missing is forced and the grid drawn again."""
    forced = list(forced)
    lines = grid_axis(edges, forced, dmax, ncross, narrow, grade)
    for _ in range(len(want) + 1):
        missing = [w for w in want if np.min(np.abs(lines - w)) > TOL]
        if not missing:
            return lines
        forced += missing
        lines = grid_axis(edges, forced, dmax, ncross, narrow, grade)
    raise RuntimeError("tensor_mesh: a via axis could not be made a grid line")


def tensor_mesh(rects, dmax, ncross, narrow, xlines=(), ylines=(), grade=1.0, points=()):
    """(nodes, tris) of the union of rectangles (x0, y0, x1, y1): counter-clockwise triangles, nodes numbered in
    order of first use. xlines / ylines: coordinates that must be grid lines (ports, elements); points: (x, y) that
    must be grid nodes (via axes)."""
    R = np.asarray(rects, float).reshape(-1, 4)
    P = np.asarray(points, float).reshape(-1, 2)
    X = _lines(R[:, [0, 2]], xlines, P[:, 0], dmax, ncross, narrow, grade)
    Y = _lines(R[:, [1, 3]], ylines, P[:, 1], dmax, ncross, narrow, grade)
    cx, cy = 0.5 * (X[:-1] + X[1:]), 0.5 * (Y[:-1] + Y[1:])
    inside = np.zeros((len(cx), len(cy)), bool)
    for x0, y0, x1, y1 in R:
        inside |= ((cx > x0) & (cx < x1))[:, None] & ((cy > y0) & (cy < y1))[None, :]
    I, J = np.nonzero(inside)
    if not len(I):
        return np.zeros((0, 2)), np.zeros((0, 3), int)
    ny = len(Y)
    c00, c10, c11, c01 = I * ny + J, (I + 1) * ny + J, (I + 1) * ny + J + 1, I * ny + J + 1
    even = (I + J) % 2 == 0
    first = np.where(even[:, None], np.stack([c00, c10, c11], 1), np.stack([c00, c10, c01], 1))
    second = np.where(even[:, None], np.stack([c00, c11, c01], 1), np.stack([c10, c11, c01], 1))
    grid_tris = np.stack([first, second], 1).reshape(-1, 3)
    # nodes numbered as the cells' corners (counter-clockwise from the lower left) first appear, cell by cell
    corners = np.stack([c00, c10, c11, c01], 1).ravel()
    used, at = np.unique(corners, return_index=True)
    rank = np.empty(len(used), int)
    rank[np.argsort(at, kind='stable')] = np.arange(len(used))
    tris = rank[np.searchsorted(used, grid_tris)]
    order = np.argsort(rank)
    nodes = np.stack([X[used[order] // ny], Y[used[order] % ny]], 1)
    return nodes, tris


def find_node(nodes, x, y, tol=TOL):
    """Index of the node at (x, y), or -1."""
    if not len(nodes):
        return -1
    d = np.hypot(nodes[:, 0] - x, nodes[:, 1] - y)
    i = int(np.argmin(d))
    return i if d[i] <= tol else -1


def contains(rects, x, y, tol=TOL):
    """(x, y) on or inside one of the rectangles."""
    return any(x0 - tol <= x <= x1 + tol and y0 - tol <= y <= y1 + tol for x0, y0, x1, y1 in rects)
