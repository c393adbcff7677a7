"""Pins of the oracle's Simulation-level glue (oracle/solve.py: via blocks placed in Z, lumped elements, port
currents -> S) by what network theory fixes, on a line the oracle meshes itself (CPU, no product code):

  * a lossless line (tan d = 0) between two via ports is reciprocal, |S12 - S21| < 1e-8, and lossless,
    |sum_i |S_i1|^2 - 1| < 2e-3 (the reference's tests/test_solver.py:353-362);
  * a lossy line (tan d = 0.02) still reciprocal, and now passive with loss: sum_i |S_i1|^2 < 1 - 1e-4;
  * a series element Z_el across the line (SeriesLumped, the lumped entries Z_el u u^T of src/mom25d/fast.py:
    628-641) equals the same line with that gap driven as a third port and the port terminated in Z_el by circuit
    algebra (the Schur complement of the 3-port's admittance matrix): the element's assembly and the port's
    excitation and current are one consistent u, to 1e-9;
  * the currents with the element satisfy the loaded system itself: (Z + Z_el u u^T) I = E, residual < 1e-10.
A plausible slip in the glue fails one of these: a via block written as B instead of B^T (reciprocity), a wrong
sign of a fan's divergence or of a port's current (losslessness), a lumped entry on the wrong edges or with l
missing (the circuit identity).
"""
import numpy as np
import pytest

from oracle import geometry as ogeo, greens as og
from oracle.solve import Board, Solver, sparams

MM = 1e-3


def _line(tand=0.0, L=10 * MM, w=0.35 * MM, h=0.2 * MM, dmax=0.3 * MM, gap=False, element=None):
    """A strip over ground with a via port to ground at each end (the via a mesh vertex at y = 0); gap: a third
    port on the line x = 0; element: a series impedance Z(f) on that line instead."""
    st = og.LayerStack([og.Layer(h, 4.3, tand)])
    nodes, tris = ogeo.make_mesh_ml([(-L / 2, -w / 2, L / 2, w / 2)], dmax, 4, 1.01 * max(w, 1.5 * dmax),
                                    xlines=[0.0], ylines=[0.0])
    rwg = ogeo.RWG(nodes, tris)
    ends = [(-L / 2, 0.0), (L / 2, 0.0)]
    chains, F, fans = [], [], []
    for i, (x, y) in enumerate(ends):
        vid = int(np.argmin(np.hypot(nodes[:, 0] - x, nodes[:, 1] - y)))
        assert np.hypot(*(nodes[vid] - (x, y))) < 1e-12
        fan = np.flatnonzero((tris == vid).any(1))
        fans.append(fan)
        chains.append((x, y, 0.1 * MM, 0, 1, i, None))
    F = np.array(sorted({t for f in fans for t in f}), int)
    pos = {t: j for j, t in enumerate(F)}
    dF = np.zeros((len(F), 2))
    for n, fan in enumerate(fans):                   # chain 0 -> 1: the fan on interface 1 is its top end
        dF[[pos[t] for t in fan], n] += -1.0 / rwg.area[fan].sum()
    on = (np.abs(rwg.ends[:, :, 0]) < 1e-12).all(1)
    idx = np.flatnonzero(on)
    sig = np.sign(rwg.nrm[idx, 0])
    assert len(idx) >= 2
    Ne = rwg.Ne
    nport = 3 if gap else 2
    E = np.zeros((Ne + 2, nport), complex)
    E[Ne, 0] = E[Ne + 1, 1] = 1.0
    if gap:
        E[idx, 2] = sig * rwg.l[idx]
    lumped = [(idx, sig, element)] if element is not None else []
    b = Board(stack=st, nodes=nodes, tris=tris, tri_layer=np.ones(len(tris), int), layers=[1], dmax=dmax,
              chains=chains, F=F, dF=dF, E=E, lumped=lumped)
    rows = [np.array([Ne]), np.array([Ne + 1])] + ([idx] if gap else [])
    signs = [np.array([1.0]), np.array([1.0])] + ([sig * rwg.l[idx]] if gap else [])
    return b, rows, signs


def _S(b, rows, signs, f, zc):
    sv = Solver(b)
    I = sv.solve(f)
    return sparams(I, b.E, zc, rows, signs), I, sv


def test_lossless_via_port_line_is_reciprocal_and_lossless():
    b, rows, signs = _line(tand=0.0)
    S, _, _ = _S(b, rows, signs, 2e9, [54.0, 54.0])
    assert abs(S[0, 1] - S[1, 0]) < 1e-8
    assert abs(np.sum(np.abs(S[:, 0]) ** 2) - 1) < 2e-3


def test_lossy_line_is_reciprocal_and_passive():
    b, rows, signs = _line(tand=0.02)
    S, _, _ = _S(b, rows, signs, 6e9, [50.0, 50.0])
    assert abs(S[0, 1] - S[1, 0]) < 1e-8
    assert np.sum(np.abs(S[:, 0]) ** 2) < 1 - 1e-4


@pytest.mark.parametrize('zel', [50.0, 10.0 + 30.0j])
def test_series_element_equals_terminated_port(zel):
    f, zc = 3e9, np.array([50.0, 50.0])
    b2, rows2, signs2 = _line(tand=0.02, element=lambda f, z=zel: z)
    S_el, I_el, sv = _S(b2, rows2, signs2, f, zc)
    b3, rows3, signs3 = _line(tand=0.02, gap=True)
    sv3 = Solver(b3)
    I3 = sv3.solve(f)
    # the 3-port admittance matrix: port j driven by a unit voltage, Y[i, j] the current of port i
    Y = np.array([[(signs3[i] * I3[rows3[i], j]).sum() for j in range(3)] for i in range(3)])
    Yl = Y[:2, :2] - np.outer(Y[:2, 2], Y[2, :2]) * zel / (1 + zel * Y[2, 2])
    # S of the two remaining ports from Yl (the port currents of unit incident-wave drives, as sparams does)
    n = 2
    Ap, Bp = np.zeros((n, n), complex), np.zeros((n, n), complex)
    for i in range(n):
        Vp = np.eye(n)[i]
        Ip = Yl[i, :]
        Ap[i], Bp[i] = (Vp + zc[i] * Ip) / (2 * np.sqrt(zc[i])), (Vp - zc[i] * Ip) / (2 * np.sqrt(zc[i]))
    S_net = Bp @ np.linalg.inv(Ap)
    assert np.abs(S_el - S_net).max() < 1e-9
    # the element's system itself: (Z + Z_el u u^T) I = E
    Z = sv.fill(f)
    r = Z @ I_el - b2.E
    assert np.abs(r).max() < 1e-10 * np.abs(Z).max() * np.abs(I_el).max() * len(r) ** 0.5
