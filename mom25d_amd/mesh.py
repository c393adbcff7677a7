"""The tensor mesh of Manhattan rectangles (src/mom25d/multilayer.py:322-349): a non-uniform grid, ncross graded
cells across narrow strips and cells of at most dmax along them, two triangles per cell inside the copper. Extra
grid lines put a node at every via axis and on every port or element line, so a via is a mesh vertex (its fan the
triangles around it) and a gap port's edges lie on its line.

The reference meshes arbitrary polygons with a constrained Delaunay mesher (shapely + CDT); that front end is out
of this build's scope, its boards being rectangles."""
import numpy as np

from .base import graded


def _axis_lines(edge_vals, port_vals, dmax, ncross, narrow, grade=0.0):
    e = np.unique(np.round(np.asarray(edge_vals, float), 12))
    pv = np.unique(np.round(np.asarray(port_vals, float), 12))
    out = [e[0]]
    for a, b in zip(e[:-1], e[1:]):
        inner = [p for p in pv if a + 1e-9 < p < b - 1e-9]
        if b - a <= narrow + 1e-12 and not inner:
            out.extend(a + (b - a) * graded(np.linspace(0, 1, ncross + 1)[1:], grade))
            continue
        bps = [a] + inner + [b]
        for c, d in zip(bps[:-1], bps[1:]):
            n = max(1, int(np.ceil((d - c) / dmax - 1e-9)))
            out.extend(np.linspace(c, d, n + 1)[1:])
    return np.array(out)


def tensor_mesh(rects, dmax, ncross, narrow, xlines=(), ylines=(), grade=1.0, points=()):
    """(nodes, tris) of the rectangles' union, counter-clockwise triangles, nodes in order of first use. points:
    (x, y) that must be grid nodes (via axes): their coordinates are added as lines where the grid has none."""
    X, Y = (_axis_lines([r[a] for r in rects] + [r[a + 2] for r in rects], ls, dmax, ncross, narrow, grade)
            for a, ls in ((0, xlines), (1, ylines)))
    if len(points):
        px = [x for x, _ in points if np.min(np.abs(X - x)) > 1e-9]
        py = [y for _, y in points if np.min(np.abs(Y - y)) > 1e-9]
        if px or py:
            X, Y = (_axis_lines([r[a] for r in rects] + [r[a + 2] for r in rects], list(ls) + extra, dmax, ncross,
                                narrow, grade) for a, ls, extra in ((0, xlines, px), (1, ylines, py)))
    cx, cy = 0.5 * (X[:-1] + X[1:]), 0.5 * (Y[:-1] + Y[1:])
    inside = np.zeros((len(cx), len(cy)), bool)
    for x0, y0, x1, y1 in rects:
        inside |= ((cx[:, None] > x0) & (cx[:, None] < x1) & (cy[None, :] > y0) & (cy[None, :] < y1))
    ids, tris = {}, []
    for i, j in zip(*np.nonzero(inside)):
        n0, n1, n2, n3 = (ids.setdefault(ij, len(ids)) for ij in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)))
        tris += [(n0, n1, n2), (n0, n2, n3)] if (i + j) % 2 == 0 else [(n0, n1, n3), (n1, n2, n3)]
    nodes = np.array([(X[i], Y[j]) for i, j in ids]) if ids else np.zeros((0, 2))
    return nodes, np.array(tris, int).reshape(-1, 3)


def find_node(nodes, x, y, tol=1e-9):
    """Index of the node at (x, y), or -1."""
    if not len(nodes):
        return -1
    d = np.hypot(nodes[:, 0] - x, nodes[:, 1] - y)
    i = int(np.argmin(d))
    return i if d[i] <= tol else -1


def contains(rects, x, y, tol=1e-9):
    """(x, y) on or inside one of the rectangles."""
    return any(x0 - tol <= x <= x1 + tol and y0 - tol <= y <= y1 + tol for x0, y0, x1, y1 in rects)
