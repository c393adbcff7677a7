"""One frequency of the direct solve on the CPU. TEST INFRASTRUCTURE ONLY.

The reference's FastOperator.solve_direct in double precision (src/mom25d/fast.py:704-850, the compiled-core
branch) with Simulation.green_tables and _via_blocks (src/mom25d/sim.py:736-745, 786-872): Green's tables, far fill,
near-field correction with its analytic parts, fan columns, via blocks, lumped elements, equilibration and LAPACK's
LU of Z^T (SciPy), solved with trans=1. Also the mixed-precision hybrid (complex64 LU + GMRES on the dense FP64 Z,
fast.py:996-1084 without factor reuse).

Input: a `Board` of plain arrays (the mesh, layers, chains, fans, excitation) so that the oracle builds its own RWG
basis, geometry dicts and near pairs from the same triangles as the product.
"""
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

from . import core as oc, geometry as ogeo, greens as og
from .gmres import gmres_batched


@dataclass
class Board:
    stack: object                   # oracle.greens.LayerStack
    nodes: np.ndarray
    tris: np.ndarray
    tri_layer: np.ndarray
    layers: list
    dmax: float
    chains: list                    # [(x, y, radius, k_bot, k_top, via id, Z(f) or None)]
    F: np.ndarray                   # fan triangles
    dF: np.ndarray                  # |F| x Nv
    E: np.ndarray                   # excitation (N x ports)
    lumped: list = field(default_factory=list)      # [(edge idx, signs, Z(f))]
    near_fac: float = 5.0
    zs: object = None               # conductor loss: f -> surface impedance (ohms), or None (lossless)


class Solver:
    """Geometry once (RWG, g / q, near pairs, analytic moments); solve(f) per frequency."""

    def __init__(self, b):
        self.b = b
        self.rwg = ogeo.RWG(b.nodes, b.tris)
        self.g, self.q, self.size = ogeo.geometry_dicts(self.rwg, b.tri_layer, b.layers)
        self.P6, self.W6 = self.q['P6'], self.q['W6']
        self.Ne, self.Nv = self.rwg.Ne, len(b.chains)
        self.rho_max = 1.05 * np.hypot(*(b.nodes.max(0) - b.nodes.min(0))) + 1e-4
        P, Q = oc.near_pairs(self.rwg.cent, self.size, b.stack.z[b.tri_layer], b.near_fac)
        self.P, self.Q = P.astype(np.int32), Q.astype(np.int32)
        self.q.update(stat=np.zeros((len(P), 12)), stat_ready=False)
        self.tri_idx = {k: np.nonzero(b.tri_layer == k)[0] for k in b.layers}

    def tables(self, f):
        self.tabs = og.green_tables(f, self.b.stack, self.b.layers, self.rho_max, self.b.dmax)
        self.tb = og.core_tables(self.tabs, self.b.layers)
        return self.tb

    def _div(self, x):
        r = self.rwg
        ap, am = r.l / r.area[r.tp], -r.l / r.area[r.tm]
        sh = (-1,) + (1,) * (x.ndim - 1)
        return ap.reshape(sh) * x[r.tp] + am.reshape(sh) * x[r.tm]

    def via_blocks(self, f, qcorr):
        """B, C (sim.py:816-872) with the fan columns (fast.py:489-504)."""
        b, r, w = self.b, self.rwg, 2 * np.pi * f
        Ne, Nv = self.Ne, self.Nv
        B, C = np.zeros((Ne, Nv), complex), np.zeros((Nv, Nv), complex)
        QF = oc.fan_q(self.g, self.tb, b.F.astype(np.int32))
        pos, sel = {q: j for j, q in enumerate(b.F)}, np.isin(self.Q, b.F)
        np.add.at(QF, (self.P[sel], np.array([pos[q] for q in self.Q[sel]], int)), qcorr[sel])
        spans = [(kb, kt) for _, _, _, kb, kt, _, _ in b.chains]
        vk = og.ViaKernels(f, b.stack, spans, b.layers, self.rho_max)
        if len(b.F):
            T = QF @ b.dF
            B += self._div(T) / (1j * w)
            C += (b.dF.T @ T[b.F]) / (1j * w)
        x = np.array([c[0] for c in b.chains])
        y = np.array([c[1] for c in b.chains])
        a = np.array([c[2] for c in b.chains])
        vid = np.array([c[5] for c in b.chains])
        for span in sorted(set(spans)):
            ns = np.array([i for i, s in enumerate(spans) if s == span])
            c = np.zeros((len(r.tris), len(ns)), complex)
            for k in b.layers:
                idx = self.tri_idx[k]
                D = vk.D[(k, span)]
                c[idx] = oc.radial_table_sum(self.P6[idx, :, 0], self.P6[idx, :, 1], self.W6[idx], x[ns], y[ns], D.v,
                                             D._dr)
            B[:, ns] += self._div(c)
        m, n = np.triu_indices(Nv)
        d = np.where(vid[m] == vid[n], a[m], np.maximum(np.hypot(x[m] - x[n], y[m] - y[n]), 0.5 * (a[m] + a[n])))
        for sm in sorted(set(spans)):
            for sn in sorted(set(spans)):
                g = np.array([spans[i] == sm and spans[j] == sn for i, j in zip(m, n)], bool)
                if not g.any():
                    continue
                mm, nn, dd = m[g], n[g], d[g]
                vals = (vk.R_table(sm, sn)(dd) if len(dd) > 64 else vk.R(sm, sn, dd)) + vk.homog(sm, sn, dd)
                np.add.at(C, (mm, nn), vals)
                off = mm != nn
                np.add.at(C, (nn[off], mm[off]), vals[off])
        seen = set()
        zs = b.zs(f) if b.zs is not None else 0.0
        for i, c in enumerate(b.chains):
            if c[6] is not None and c[5] not in seen:
                C[i, i] += c[6](f)
            seen.add(c[5])
            C[i, i] += zs * (b.stack.z[c[4]] - b.stack.z[c[3]]) / (2 * np.pi * c[2])      # the barrel's wall loss
        return B, C

    def fill(self, f):
        """Z (complex128, N x N) at f, not equilibrated."""
        b, w = self.b, 2 * np.pi * f
        self.tables(f)
        N = self.Ne + self.Nv
        Z = np.zeros((N, N), complex)
        oc.fill_far(Z, self.g, self.tb, w)
        qcorr = oc.near_into_dense(Z, self.g, self.tb, self.q, self.P, self.Q, w)
        self.q['stat_ready'] = True
        if self.Nv:
            B, C = self.via_blocks(f, qcorr)
            Z[:self.Ne, self.Ne:], Z[self.Ne:, :self.Ne], Z[self.Ne:, self.Ne:] = B, B.T, C
        for idx, sig, zf in b.lumped:
            u = sig * self.rwg.l[idx]
            Z[np.ix_(idx, idx)] += zf(f) * np.outer(u, u)
        if b.zs is not None:                      # conductor loss: Zs int f_m . f_n dS (fast.py:628-641)
            if getattr(self, '_gram', None) is None:
                self._gram = ogeo.gram_midpoint(self.rwg)
            Z[:self.Ne, :self.Ne] += b.zs(f) * self._gram
        return Z

    def solve(self, f, precision='double'):
        """Currents (N x ports) at f."""
        Z = self.fill(f)
        dsc = 1.0 / np.sqrt(np.abs(np.diagonal(Z)).astype(float) + 1e-300)
        Z *= dsc[:, None]
        Z *= dsc[None, :]
        Bx = dsc[:, None] * self.b.E
        if precision == 'double':
            lu = sla.lu_factor(Z.T, check_finite=False)
            return dsc[:, None] * sla.lu_solve(lu, Bx, trans=1, check_finite=False)
        lu = sla.lu_factor(Z.T.astype(np.complex64), check_finite=False)
        M = lambda R: sla.lu_solve(lu, R.astype(np.complex64), trans=1, check_finite=False).astype(complex)
        Y, it = gmres_batched(lambda V: Z @ V, Bx, M=M, rtol=1e-6, restart=30, maxiter=1)
        self.iters = it
        return dsc[:, None] * Y


def sparams(I, E, zc, port_rows, signs):
    """S of lumped ports from the currents (sim.py:1045-1092): port_rows / signs give each port's current as
    sum(signs * I[rows])."""
    n = E.shape[1]
    Ap, Bp = np.zeros((n, n), complex), np.zeros((n, n), complex)
    for i in range(n):
        Vp = np.eye(n)[i]
        Ip = (np.asarray(signs[i])[:, None] * I[port_rows[i]]).sum(0)
        Ap[i], Bp[i] = (Vp + zc[i] * Ip) / (2 * np.sqrt(zc[i])), (Vp - zc[i] * Ip) / (2 * np.sqrt(zc[i]))
    return Bp @ np.linalg.inv(Ap)
