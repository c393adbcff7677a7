"""Meshes, RWG basis functions, quadrature points and the kernels' geometry dicts. TEST INFRASTRUCTURE ONLY.

  graded, make_mesh_ml   src/mom25d/base.py:81-83, src/mom25d/multilayer.py:322-349 (tensor mesh of rectangles)
  RWG                    src/mom25d/base.py:16-47 (edges in first-seen order, T+ the first triangle)
  _BAR6/_W6, _BAR3/_W3   src/mom25d/base.py:70-77 (6-point degree-4 and 3-point degree-2 rules)
  geometry_dicts         src/mom25d/fast.py:375-400 (g, q) and src/mom25d/sim.py:543-548 (P3, W3, P6, W6, size)
"""
import numpy as np

_a1, _a2 = 0.445948490915965, 0.091576213509771
_BAR6 = np.array([[1 - 2 * _a1, _a1, _a1], [_a1, 1 - 2 * _a1, _a1], [_a1, _a1, 1 - 2 * _a1],
                  [1 - 2 * _a2, _a2, _a2], [_a2, 1 - 2 * _a2, _a2], [_a2, _a2, 1 - 2 * _a2]])
_W6 = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)
_BAR3 = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
_W3 = np.ones(3) / 3


def graded(u, grade):
    return (1 - grade) * u + grade * (1 - np.cos(np.pi * u)) / 2


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


def make_mesh_ml(rects, dmax, ncross, narrow, xlines=(), ylines=(), grade=1.0):
    X, Y = (_axis_lines([r[a] for r in rects] + [r[a + 2] for r in rects], ls, dmax, ncross, narrow, grade)
            for a, ls in ((0, xlines), (1, ylines)))
    cx, cy = 0.5 * (X[:-1] + X[1:]), 0.5 * (Y[:-1] + Y[1:])
    inside = np.zeros((len(cx), len(cy)), bool)
    for x0, y0, x1, y1 in rects:
        inside |= ((cx[:, None] > x0) & (cx[:, None] < x1) & (cy[None, :] > y0) & (cy[None, :] < y1))
    ids, tris = {}, []
    for i, j in zip(*np.nonzero(inside)):
        n0, n1, n2, n3 = (ids.setdefault(ij, len(ids)) for ij in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)))
        tris += [(n0, n1, n2), (n0, n2, n3)] if (i + j) % 2 == 0 else [(n0, n1, n3), (n1, n2, n3)]
    return np.array([(X[i], Y[j]) for i, j in ids]), np.array(tris)


class RWG:
    """RWG basis functions on interior edges (base.py:16-47)."""
    def __init__(self, nodes, tris):
        self.nodes, self.tris = nodes, tris
        V = nodes[tris]
        e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
        self.area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
        assert (self.area > 0).all(), "triangles must be counter-clockwise"
        self.cent = V.mean(1)
        d = {}
        for t, (a, b, c) in enumerate(tris):
            for i, (u, v) in enumerate(((a, b), (b, c), (c, a))):
                d.setdefault((min(u, v), max(u, v)), []).append((t, i))
        E = [(k, v) for k, v in d.items() if len(v) == 2]
        self.Ne = Ne = len(E)
        u, v = np.array([k for k, _ in E], int).reshape(-1, 2).T
        (t1, i1), (t2, i2) = np.array([s for _, s in E], int).reshape(-1, 2, 2).transpose(1, 2, 0)
        self.tp, self.tm = t1, t2
        self.vp, self.vm = nodes[tris[t1, (i1 + 2) % 3]], nodes[tris[t2, (i2 + 2) % 3]]
        pu, pv = nodes[u], nodes[v]
        self.l = np.linalg.norm(pv - pu, axis=1)
        self.mid, self.ends = 0.5 * (pu + pv), np.stack([pu, pv], 1)
        tau = (pv - pu) / self.l[:, None]
        n = np.c_[tau[:, 1], -tau[:, 0]]
        n[((self.cent[t2] - self.cent[t1]) * n).sum(1) < 0] *= -1
        self.nrm = n


def quad_pts(rwg, bar):
    return np.einsum('qk,tkc->tqc', bar, rwg.nodes[rwg.tris])


def geometry_dicts(rwg, tri_layer, layers):
    """(g, q, size): the compiled kernels' geometry (fast.py:375-400) and the triangles' longest edges."""
    r, Ne = rwg, rwg.Ne
    h_tri = np.concatenate([r.tp, r.tm])
    h_alpha = np.concatenate([r.l / (2 * r.area[r.tp]), -r.l / (2 * r.area[r.tm])])
    h_v = np.vstack([r.vp, r.vm])
    H, fill = -np.ones((len(r.tris), 3), int), np.zeros(len(r.tris), int)
    for h, t in enumerate(h_tri):
        H[t, fill[t]] = h
        fill[t] += 1
    lidx = {k: i for i, k in enumerate(layers)}
    c = np.ascontiguousarray
    g = dict(ep=r.tp.astype(np.int32), em=r.tm.astype(np.int32),
             lt=np.array([lidx[k] for k in tri_layer], np.int32),
             te=np.where(H >= 0, H % Ne, -1).astype(np.int32), ta=np.where(H >= 0, h_alpha[H], 0.0),
             tv=np.where((H >= 0)[..., None], h_v[H], 0.0),
             ap=h_alpha[:Ne].copy(), am=h_alpha[Ne:].copy(), vp=c(h_v[:Ne]), vm=c(h_v[Ne:]), cent=c(r.cent),
             area=r.area.copy())
    P3, W3 = quad_pts(r, _BAR3), r.area[:, None] * _W3[None, :]
    P6, W6 = quad_pts(r, _BAR6), r.area[:, None] * _W6[None, :]
    q = dict(nodes=c(r.nodes), tris=r.tris.astype(np.int32), P3=c(P3), W3=c(W3), P6=c(P6), W6=c(W6))
    V = r.nodes[r.tris]
    size = np.linalg.norm(V - np.roll(V, 1, axis=1), axis=2).max(1)
    return g, q, size


def gram_midpoint(rwg):
    """G_mn = int f_m . f_n dS (dense), each shared triangle by the three-edge-midpoint rule (exact for the
    quadratic integrand) from the RWG functions' definition (base.py:153-157 integrates the same exactly)."""
    G = np.zeros((rwg.Ne, rwg.Ne))
    support = {}
    for e in range(rwg.Ne):
        for t, v, s in ((rwg.tp[e], rwg.vp[e], 1.0), (rwg.tm[e], rwg.vm[e], -1.0)):
            support.setdefault(int(t), []).append((e, v, s * rwg.l[e] / (2 * rwg.area[t])))
    for t, fs in support.items():
        P = rwg.nodes[rwg.tris[t]]
        mids = 0.5 * (P + np.roll(P, -1, axis=0))
        for m, vm, am in fs:
            for n, vn, an in fs:
                G[m, n] += am * an * rwg.area[t] / 3 * ((mids - vm) * (mids - vn)).sum()
    return G
