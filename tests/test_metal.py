"""Conductor loss (CPU): the RWG Gram matrix against the edge-midpoint rule (exact for quadratics, a formula
independent of the closed-form moments used), its symmetry and positive definiteness, and the surface impedance's
limits: 1 / (sigma t) at DC, (1 + j) / (sigma delta) for a thick sheet, roughness only raising the resistance
(by at most 2x), a finish of the copper itself changing nothing."""
import numpy as np

from mom25d_amd import boards
from mom25d_amd.base import MU0, geometry, gram_entries
from mom25d_amd.sim import Metal, Simulation


def midpoint_gram(r):
    """G_mn by the three-edge-midpoint rule over each shared triangle, from the RWG definitions themselves."""
    Ne = r.Ne
    G = np.zeros((Ne, Ne))
    sup = {}
    for e in range(Ne):
        for t, v, s in ((r.tp[e], r.vp[e], 1.0), (r.tm[e], r.vm[e], -1.0)):
            sup.setdefault(t, []).append((e, v, s * r.l[e] / (2 * r.area[t])))
    for t, fs in sup.items():
        P = r.nodes[r.tris[t]]
        mids = 0.5 * (P + np.roll(P, -1, axis=0))
        for m, vm, am in fs:
            for n, vn, an in fs:
                G[m, n] += am * an * r.area[t] / 3 * sum(np.dot(q - vm, q - vn) for q in mids)
    return G


def test_gram_is_exact_symmetric_and_positive():
    st, sh, po, vi, kw = boards.microstrip(dmax=1.5e-3, ncross=2)
    sim = Simulation(st, sh, po, vi, verbose=False, mesh_only=True, **kw)
    r = sim.rwg
    g, q, _ = geometry(r, sim.tri_layer, sim.layers)
    rows, cols, vals = gram_entries(g, r.nodes, r.tris)
    G = np.zeros((r.Ne, r.Ne))
    G[rows, cols] = vals
    ref = midpoint_gram(r)
    assert np.abs(G - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(G, G.T) or np.abs(G - G.T).max() <= 1e-15 * np.abs(G).max()
    assert np.linalg.eigvalsh(0.5 * (G + G.T)).min() > 0


def test_surface_impedance_limits():
    m = Metal(sigma=5.8e7, t=35e-6)
    assert abs(m.zs(1.0) - 1 / (5.8e7 * 35e-6)) <= 1e-6 / (5.8e7 * 35e-6)
    f = 50e9
    depth = 1 / np.sqrt(np.pi * f * MU0 * 5.8e7)
    assert abs(m.zs(f) - (1 + 1j) / (5.8e7 * depth)) <= 1e-9 * abs(m.zs(f))
    rough = Metal(sigma=5.8e7, t=35e-6, rough=2e-6)
    for fr in (1e8, 1e9, 1e10, 1e11):
        ratio = rough.zs(fr).real / m.zs(fr).real
        assert 1.0 <= ratio <= 2.0 and rough.zs(fr).imag == m.zs(fr).imag
    same = Metal(sigma=5.8e7, t=35e-6, finish=[(5.8e7, 2e-6)], on={1: 0.5})
    assert abs(same.face_factor(5e9) - 1) < 1e-12
    nickel = Metal(sigma=5.8e7, t=35e-6, finish=[(1.45e7, 5e-6)], on={1: 0.5})
    assert nickel.zs(5e9, 1).real > m.zs(5e9).real and nickel.zs(5e9, 2) == m.zs(5e9)
