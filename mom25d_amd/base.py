"""RWG basis functions, quadrature rules and the kernels' geometry (host side, NumPy; once per mesh).

  RWG             src/mom25d/base.py:16-47: interior edges in first-seen order over the triangles' edges (a, b),
                  (b, c), (c, a); T+ the first triangle seen; alpha = +-l / 2A; the free vertices
  _BAR6/_W6, ...  src/mom25d/base.py:70-77: the 6-point degree-4 and 3-point degree-2 rules
  geometry        src/mom25d/fast.py:375-400 and src/mom25d/sim.py:543-548: the kernels' g / q dicts
  gram_entries    the RWG Gram matrix G_mn = int f_m . f_n dS (src/mom25d/base.py:153-157), exactly, as sparse
                  entries (the surface-impedance term Zs G of conductor loss)
"""
import numpy as np

MU0 = 4e-7 * np.pi
C0 = 299792458.0
EPS0 = 1.0 / (MU0 * C0 ** 2)

_a1, _a2 = 0.445948490915965, 0.091576213509771
_BAR6 = np.array([[1 - 2 * _a1, _a1, _a1], [_a1, 1 - 2 * _a1, _a1], [_a1, _a1, 1 - 2 * _a1],
                  [1 - 2 * _a2, _a2, _a2], [_a2, 1 - 2 * _a2, _a2], [_a2, _a2, 1 - 2 * _a2]])
_W6 = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)
_BAR3 = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
_W3 = np.ones(3) / 3


def graded(u, grade):
    """u in 0..1 graded toward both ends (cosine) by `grade` (base.py:81-83)."""
    return (1 - grade) * u + grade * (1 - np.cos(np.pi * u)) / 2


class RWG:
    """RWG basis on the interior edges of a triangle mesh (counter-clockwise triangles)."""

    def __init__(self, nodes, tris):
        nodes, tris = np.asarray(nodes, float), np.asarray(tris, int)
        self.nodes, self.tris = nodes, tris
        V = nodes[tris]
        e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
        self.area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
        if not (self.area > 0).all():
            raise ValueError("RWG: triangles must be counter-clockwise with positive area")
        self.cent = V.mean(1)
        # every triangle side as (min node, max node), in the order (t, side): an edge is interior when exactly two
        # sides share it; edges are numbered by the first side seen, T+ that side's triangle
        nt = len(tris)
        a, b = tris[:, [0, 1, 2]], tris[:, [1, 2, 0]]
        lo, hi = np.minimum(a, b).ravel(), np.maximum(a, b).ravel()
        key = lo.astype(np.int64) * (len(nodes) + 1) + hi
        order = np.argsort(key, kind='stable')               # sides of one edge adjacent, in (t, side) order
        ks = key[order]
        starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
        counts = np.diff(np.r_[starts, len(ks)])
        two = starts[counts == 2]
        first, second = order[two], order[two + 1]          # flat side indices t * 3 + i, first < second
        o = np.argsort(first, kind='stable')                 # first-seen order
        first, second = first[o], second[o]
        t1, i1, t2, i2 = first // 3, first % 3, second // 3, second % 3
        self.Ne = len(first)
        self.tp, self.tm = t1, t2
        self.vp, self.vm = nodes[tris[t1, (i1 + 2) % 3]], nodes[tris[t2, (i2 + 2) % 3]]
        u, v = lo[first], hi[first]
        pu, pv = nodes[u], nodes[v]
        self.l = np.linalg.norm(pv - pu, axis=1)
        self.mid, self.ends = 0.5 * (pu + pv), np.stack([pu, pv], 1)
        tau = (pv - pu) / self.l[:, None]
        n = np.c_[tau[:, 1], -tau[:, 0]]
        n[((self.cent[t2] - self.cent[t1]) * n).sum(1) < 0] *= -1
        self.nrm = n                                         # direction of positive current (T+ -> T-)
        del nt

    def div_apply(self, x):
        """D^T x for per-triangle values x (Nt, ...): the RWG divergence tested on each edge, +l/A on T+,
        -l/A on T- (base.py:47)."""
        ap, am = self.l / self.area[self.tp], -self.l / self.area[self.tm]
        sh = (-1,) + (1,) * (np.ndim(x) - 1)
        return ap.reshape(sh) * x[self.tp] + am.reshape(sh) * x[self.tm]


def quad_pts(rwg, bar):
    return np.einsum('qk,tkc->tqc', bar, rwg.nodes[rwg.tris])


def geometry(rwg, tri_layer, layers):
    """(g, q, size): the kernels' mesh dict (fast.py:389-398), quadrature dict (fast.py:399-400, sim.py:545-546)
    and each triangle's longest side (sim.py:544)."""
    r, Ne, Nt = rwg, rwg.Ne, len(rwg.tris)
    h_tri = np.concatenate([r.tp, r.tm])
    h_alpha = np.concatenate([r.l / (2 * r.area[r.tp]), -r.l / (2 * r.area[r.tm])])
    h_v = np.vstack([r.vp, r.vm])
    # the halves on each triangle in half order (edge h on T+, then h - Ne on T-), slots filled in that order
    H = -np.ones((Nt, 3), int)
    order = np.argsort(h_tri, kind='stable')
    ts = h_tri[order]
    slot = np.arange(len(ts)) - np.searchsorted(ts, ts, 'left')
    H[ts, slot] = order
    lidx = {k: i for i, k in enumerate(layers)}
    c = np.ascontiguousarray
    g = dict(ep=r.tp.astype(np.int32), em=r.tm.astype(np.int32),
             lt=np.array([lidx[k] for k in tri_layer], np.int32),
             te=np.where(H >= 0, H % max(Ne, 1), -1).astype(np.int32), ta=np.where(H >= 0, h_alpha[H], 0.0),
             tv=np.where((H >= 0)[..., None], h_v[H], 0.0),
             ap=h_alpha[:Ne].copy(), am=h_alpha[Ne:].copy(), vp=c(h_v[:Ne]), vm=c(h_v[Ne:]), cent=c(r.cent),
             area=r.area.copy())
    P3, W3 = quad_pts(r, _BAR3), r.area[:, None] * _W3[None, :]
    P6, W6 = quad_pts(r, _BAR6), r.area[:, None] * _W6[None, :]
    q = dict(nodes=c(r.nodes), tris=r.tris.astype(np.int32), P3=c(P3), W3=c(W3), P6=c(P6), W6=c(W6))
    V = r.nodes[r.tris]
    size = np.linalg.norm(V - np.roll(V, 1, axis=1), axis=2).max(1)
    return g, q, size


def gram_entries(g, nodes, tris):
    """(rows, cols, vals) of the nonzero entries of G_mn = int f_m . f_n dS, each (m, n) once: a sum over the
    triangles the two functions share of alpha_m alpha_n int_T (r - v_m) . (r - v_n) dA, with the integral exact
    (a quadratic over the triangle: int_T |r|^2 = A / 12 (sum_k |p_k|^2 + |sum_k p_k|^2), int_T r = A c)."""
    te, ta = np.asarray(g['te']), np.asarray(g['ta'])
    tv = np.asarray(g['tv']).reshape(-1, 3, 2)
    A, c = np.asarray(g['area']), np.asarray(g['cent'])
    P = nodes[tris]
    r2 = A / 12 * ((P ** 2).sum((1, 2)) + (P.sum(1) ** 2).sum(1))
    rows, cols, vals = [], [], []
    for s in range(3):
        for u in range(3):
            on = (te[:, s] >= 0) & (te[:, u] >= 0)
            vs, vu = tv[on, s], tv[on, u]
            v = ta[on, s] * ta[on, u] * (r2[on] - A[on] * ((vs + vu) * c[on]).sum(1) + A[on] * (vs * vu).sum(1))
            rows.append(te[on, s]); cols.append(te[on, u]); vals.append(v)
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    Ne = len(g['ep'])
    key, inv = np.unique(rows.astype(np.int64) * Ne + cols, return_inverse=True)
    return (key // Ne).astype(np.int64), (key % Ne).astype(np.int64), np.bincount(inv, weights=vals)
