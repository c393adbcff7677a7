"""Small seeded test geometries and their tables, built with the oracle (test helper)."""
import numpy as np

from oracle import geometry as ogeo, greens as og

MU = 1e-3


def two_layer(f=4e9, dmax=0.4e-3, ncross=4, image=True):
    """A strip on interface 2 over a wider plate on interface 1 of a two-layer stack (the shape of the reference's
    test_core_matches_numpy model, as rectangles): (stack, g, q, size, tb, tri_layer, rwg)."""
    st = og.LayerStack([og.Layer(0.2e-3, 4.3, 0.02), og.Layer(0.2e-3, 4.3, 0.02)])
    n1, t1 = ogeo.make_mesh_ml([(-4e-3, -0.175e-3, 4e-3, 0.175e-3)], dmax, ncross, 0.6e-3)
    n2, t2 = ogeo.make_mesh_ml([(-4e-3, -1.5e-3, 4.5e-3, 1.5e-3)], 1.5 * dmax, ncross, 0.6e-3)
    nodes = np.vstack([n2, n1])
    tris = np.vstack([t2, t1 + len(n2)])
    tl = np.r_[np.full(len(t2), 1), np.full(len(t1), 2)]
    rwg = ogeo.RWG(nodes, tris)
    g, q, size = ogeo.geometry_dicts(rwg, tl, [1, 2])
    rho_max = 1.05 * np.hypot(*(nodes.max(0) - nodes.min(0))) + 1e-4
    tabs = og.green_tables(f, st, [1, 2], rho_max, dmax if image else 0.0)
    tb = og.core_tables(tabs, [1, 2])
    return dict(stack=st, g=g, q=q, size=size, tb=tb, tl=tl, rwg=rwg, w=2 * np.pi * f, rho_max=rho_max,
                tables=tabs)


def seeded_matrix(n, seed=3, swap=True, dtype=np.complex128):
    """The reference's LU test matrix (tests/test_gpu.py:219-223): normal complex + 4 sqrt(n) I, rows 0 and 7
    swapped."""
    rng = np.random.default_rng(seed)
    Z = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(dtype)
    Z += 4 * np.sqrt(n) * np.eye(n, dtype=dtype)
    if swap and n > 7:
        Z[[0, 7]] = Z[[7, 0]]
    return Z
