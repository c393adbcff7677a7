"""The simulation object (src/mom25d/sim.py): a stack, Manhattan copper on its interfaces, vias, lumped ports and
elements -> a mesh, the RWG and via-chain unknowns, and S-parameters per frequency.

    sim = Simulation(stack, {1: [(x0, y0, x1, y1), ...]}, ports=[ViaPort(Via(x, y, 0, 1))], dmax=0.3e-3)
    S = sim.sparams(f)

Kept from the reference: the constructor's arguments that apply to rectangle boards, the unknowns (edges, then
one per via chain), the fans and their divergence dF, the Green tables per interface pair (green_tables), the via
blocks (_via_blocks), the excitation and port currents, sparams, sweep, balance and memory_plan. The solve of each
frequency is mom25d_amd.fast.FastOperator on the GPU. Out of scope here: polygon meshing (shapely/CDT), wave ports
(LinePort / DiffPort with their wave fits), metal loss, PEC planes with holes or slots, quads and line functions.
"""
from types import SimpleNamespace

import numpy as np
import torch

from .base import MU0, RWG, geometry
from .mesh import contains, find_node, tensor_mesh
from .multilayer import RunLog, device_tables, green_tables
from . import device as dv

GPU_MIN_N = 3000          # mixed precision (complex64 LU + FP64 refinement) above this many unknowns
REUSE_KEEP = 4            # kept factors a solver holds at most (mom25d_amd.fast.REUSE_KEEP)
WORK_BYTES = 0.4e9        # the LU's and the solves' workspaces, the tables, the mesh and the near plan, bounded


def _zfun(Z):
    return Z if callable(Z) else (lambda f, Z=Z: complex(Z))


def rlc(R=0.0, L=0.0, C=None, parallel=False):
    """The impedance function f -> Z of an R-L-C element, series (default) or parallel (the reference's rlc,
    src/mom25d/sim.py:47-54)."""
    def series(f):
        jw = 2j * np.pi * f
        return R + jw * L + (1 / (jw * C) if C else 0)

    def shunt(f):
        jw = 2j * np.pi * f
        y = (1 / R if R else 0) + (1 / (jw * L) if L else 0) + (jw * C if C else 0)
        return 1 / y
    return shunt if parallel else series


def available_memory():
    """Bytes of host memory available now (MemAvailable), the reference's emsdev.memory.available."""
    try:
        with open('/proc/meminfo') as fh:
            for line in fh:
                if line.startswith('MemAvailable:'):
                    return float(line.split()[1]) * 1024
    except OSError:
        pass
    return 16e9


def device_bytes(n, precision, keep=REUSE_KEEP):
    """Device bytes one n-unknown solve holds on the GPU in this build, where every matrix stays there: 'double' -
    the complex128 Z factored in place (16 n^2); 'mixed' - the FP64 Z the refinement multiplies by (16 n^2), the
    complex64 factor (8 n^2) and up to `keep` kept factors of nearby frequencies (8 n^2 each, complex128 ones 16);
    plus WORK_BYTES."""
    nn = float(n) * n
    if precision == 'double':
        return 16.0 * nn + WORK_BYTES
    return 24.0 * nn + 8.0 * nn * keep + WORK_BYTES


def memory_plan(n, has_core=True, avail=None, gpu=None, precision='auto', solver='auto', n_lu=None, has_blr=False,
                keep=REUSE_KEEP):
    """How an n-unknown solve runs and what it holds: dict(solver, precision, path, host, gpu, fits), the
    reference's plan (src/mom25d/sim.py:85-151) for this build, whose direct solver keeps everything on the GPU.

    gpu: None, (name, free bytes) or (name, free bytes, the LU's bytes, fits) as the reference's gpu_info gives.
    precision 'auto': 'mixed' above GPU_MIN_N unknowns of the LU (n_lu, default n), else 'double'; a requested
    precision is kept. path: 'hybrid' (complex64 LU and FP64 refinement, both on the GPU) or 'double' (complex128
    LU on the GPU). gpu (the result): the device bytes of device_bytes(); host: the bytes the host holds (the
    mesh, geometry and currents, O(n)); fits: whether the device bytes fit 0.9 x the free device memory. The
    solver is 'direct' (dense LU) whatever is asked: the compressed solvers (BLR, ACA) and the host paths
    (cpu-mixed, gpu-stream) are not part of this build, and a request that does not fit is reported as such
    (fits False) for the caller to queue or refuse."""
    if n < 0:
        raise ValueError("memory_plan: n must be >= 0")
    free = float(gpu[1]) if gpu else 0.0
    m = n_lu or n
    if precision == 'auto':
        precision = 'mixed' if m > GPU_MIN_N else 'double'
    if precision not in ('mixed', 'double'):
        raise ValueError(f"memory_plan: precision {precision!r}")
    dev = device_bytes(n, precision, keep)
    if gpu and len(gpu) >= 4 and precision == 'mixed':
        dev = max(dev, float(gpu[2]) + 16.0 * n * n)        # the backend's own estimate of its factor, and Z
    fits = gpu is not None and dev <= 0.9 * free
    path = 'hybrid' if precision == 'mixed' else 'double'
    host = 2e4 * n + 64 * 16.0 * n + 0.2e9
    return dict(solver='direct', precision=precision, path=path, host=host, gpu=dev, fits=fits)


class Via:
    """Plated via on axis (x, y) from interface k_bot to k_top, barrel radius; Z: an element in its lowest chain
    (sim.py:154-160)."""

    def __init__(self, x, y, k_bot, k_top, radius=0.15e-3, Z=None, name='', ring=True):
        self.x, self.y, self.k_bot, self.k_top, self.a, self.name = x, y, k_bot, k_top, radius, name
        self.ring = ring
        self.Z = None if Z is None else _zfun(Z)


class LinePort:
    """The transverse line of a gap or element across a feed (geo.py:24-60): layer, center, direction into the
    device, width; s along the direction."""
    kind = 'line'

    def __init__(self, layer, center, direction, width, ref_dist=0.0, name=''):
        d = np.asarray(direction, float)
        self.layer, self.c, self.d = layer, np.asarray(center, float), d / np.linalg.norm(d)
        self.t = np.array([-self.d[1], self.d[0]])
        self.w, self.ref, self.name = width, ref_dist, name

    @property
    def axis_aligned(self):
        return abs(self.d[0]) < 1e-12 or abs(self.d[1]) < 1e-12


class GapPort:
    """Lumped port across a strip (delta gap); reference plane at the gap (sim.py:163-175)."""
    kind = 'gap'

    def __init__(self, layer, center, direction, width, z0=50.0, name=''):
        self.lp = LinePort(layer, center, direction, width, 0.0)
        self.layer, self.zc, self.name = layer, z0, name


class ViaPort:
    """Lumped port inside a via: an EMF in its lowest chain, + toward the top (sign=-1: toward the bottom)
    (sim.py:178-184)."""
    kind = 'via'

    def __init__(self, via, z0=50.0, name='', sign=1):
        self.via, self.zc, self.name, self.layer, self.sign = via, z0, name, None, float(sign)


class SeriesLumped:
    """Series element across a strip (sim.py:221-225)."""

    def __init__(self, layer, center, direction, width, Z, name=''):
        self.lp = LinePort(layer, center, direction, width, 0.0)
        self.layer, self.Z, self.name = layer, _zfun(Z), name


BALANCE_FLOOR = 1e-2


def balance(S, exact=None):
    """S made reciprocal and the ports' scales (sim.py:1121-1141)."""
    A = np.abs(S)
    W = A * A.T
    np.fill_diagonal(W, 0.0)
    with np.errstate(divide='ignore', invalid='ignore'):
        R = np.where(W > 0, 0.5 * np.log(A / A.T), 0.0)
    L = np.diag(W.sum(1) + BALANCE_FLOOR) - W
    if exact is not None and np.any(exact):
        L = L + np.diag(np.where(exact, 1e6, 0.0))
    d = np.exp(np.linalg.solve(L, (W * R).sum(1)))
    S = S * (d[None, :] / d[:, None])
    return 0.5 * (S + S.T), d


class Simulation:
    """A board of rectangles on the stack's interfaces, meshed and set up for the direct solver on the GPU."""

    def __init__(self, stack, shapes, ports=(), vias=(), lumped=(), dmax=0.3e-3, dmax_layer=None, w_ref=None,
                 ncross=6, grade=1.0, near_fac=5.0, solver='direct', precision='auto', gpu='auto', verbose=True,
                 title='', mesh_only=False, metal=None):
        if metal is not None:
            raise ValueError("metal loss is not supported by this build (lossless conductors)")
        if solver not in ('direct', 'auto'):
            raise ValueError(f"solver {solver!r}: this build has the direct (dense LU) solver only")
        self.st, self.title, self.metal, self.log = stack, title, None, RunLog(verbose)
        self.gpu, self.reuse, self.stream = gpu, 'auto', 'off'
        self.dmax, self.ncross, self.grade, self.near_fac = dmax, ncross, grade, near_fac
        self.ports = list(ports)
        self.vias, self.lumped = list(vias), list(lumped)
        self.vias += [p.via for p in self.ports if isinstance(p, ViaPort) and p.via not in self.vias]
        self.layers = sorted(shapes)
        for k in self.layers:
            if k not in stack.metal_interfaces:
                raise ValueError(f"interface {k} cannot carry metal in this stack")
        self.shapes = {k: [tuple(map(float, r)) for r in shapes[k]] for k in self.layers}
        dmax_layer = dmax_layer or {}
        # ---- via chains: split where the via touches metal (sim.py:421-434) ----------------------------------
        self.chains, self.via_attach = [], []
        for v in self.vias:
            on = lambda k: k in self.shapes and contains(self.shapes[k], v.x, v.y)
            for k in (v.k_bot, v.k_top):
                if not (stack.is_pec(k) or on(k)):
                    raise ValueError(f"via {v.name}: ends on M{k} where there is no metal")
            att = [v.k_bot] + [k for k in range(v.k_bot + 1, v.k_top) if on(k)] + [v.k_top]
            self.via_attach.append([k for k in att if k in self.shapes])
            self.chains += [(v, a, b) for a, b in zip(att[:-1], att[1:])]
        # ---- lines the mesh must follow: gap ports, elements, via axes --------------------------------------
        lines = {k: [] for k in self.layers}
        for p in self.ports:
            if isinstance(p, GapPort):
                lines[p.layer].append(p.lp)
        for el in self.lumped:
            lines[el.layer].append(el.lp)
        with self.log.timer('mesh'):
            nodes, tris, tl, off, via_vertex = [], [], [], 0, {}
            for k in self.layers:
                xl, yl = [], []
                for lp in lines[k]:
                    if not lp.axis_aligned:
                        raise ValueError("gap ports and elements must be axis-aligned on a rectangle board")
                    (xl if abs(lp.d[0]) > 0.5 else yl).append(lp.c[0] if abs(lp.d[0]) > 0.5 else lp.c[1])
                vhere = [(i, v) for i, v in enumerate(self.vias) if k in self.via_attach[i]]
                dm = dmax_layer.get(k, dmax)
                wr = w_ref.get(k) if isinstance(w_ref, dict) else w_ref
                nd, tr = tensor_mesh(self.shapes[k], dm, ncross, 1.01 * max(wr or 0, 1.5 * dm), xl, yl, grade=grade,
                                     points=[(v.x, v.y) for _, v in vhere])
                for i, v in vhere:
                    vid = find_node(nd, v.x, v.y)
                    if vid < 0:
                        raise ValueError(f"via {v.name}: not on the M{k} mesh")
                    via_vertex[(i, k)] = vid + off
                nodes.append(nd)
                tris.append(tr + off)
                tl.append(np.full(len(tr), k))
                off += len(nd)
            self.rwg = RWG(np.vstack(nodes), np.vstack(tris))
            self.tri_layer = np.concatenate(tl)
            self.edge_layer = self.tri_layer[self.rwg.tp]
        r = self.rwg
        self.Ne, self.Nv = r.Ne, len(self.chains)
        if mesh_only:
            return
        self.tri_idx = {k: np.nonzero(self.tri_layer == k)[0] for k in self.layers}
        # ---- fans and their divergence (sim.py:525-535) ------------------------------------------------------
        self.fan = {key: np.nonzero((r.tris == vid).any(1))[0] for key, vid in via_vertex.items()}
        self.F = np.array(sorted({t for v in self.fan.values() for t in v}), int)
        fpos = {t: j for j, t in enumerate(self.F)}
        self.dF = np.zeros((len(self.F), self.Nv))
        for n, (v, kb, kt) in enumerate(self.chains):
            i = self.vias.index(v)
            for k, div in ((kb, +1.0), (kt, -1.0)):
                if (i, k) in self.fan:
                    ts = self.fan[(i, k)]
                    self.dF[[fpos[t] for t in ts], n] += div / r.area[ts].sum()
        self.log.note(f"mesh: {len(r.tris)} triangles, {r.Ne} edge + {self.Nv} via unknowns")
        self.g, self.q, self.tri_size = geometry(r, self.tri_layer, self.layers)
        self.P3, self.W3, self.P6, self.W6 = self.q['P3'], self.q['W3'], self.q['P6'], self.q['W6']
        self.rho_max = 1.05 * np.hypot(*(r.nodes.max(0) - r.nodes.min(0))) + 1e-4
        self._edge_sets()
        n = self.Ne + self.Nv
        plan = memory_plan(n, gpu=('hip', self._gpu_free()), precision=precision)
        self.plan, self.precision, self.solver = plan, plan['precision'], 'direct'
        self.log.note(f"solver: direct on the GPU (N = {n}, {self.precision} precision)")
        from .fast import FastOperator
        self.fast = FastOperator(self)

    @staticmethod
    def _gpu_free():
        if not torch.cuda.is_available():
            return 0.0
        free, _ = torch.cuda.mem_get_info()
        return float(free)

    # ---- edges on lines (sim.py:697-728) -----------------------------------------------------------------
    def _line_edges(self, lp, s=0.0, tol=1e-8):
        r = self.rwg
        e0, e1 = r.ends[:, 0] - (lp.c + lp.d * s), r.ends[:, 1] - (lp.c + lp.d * s)
        on = (np.abs(e0 @ lp.d) < tol) & (np.abs(e1 @ lp.d) < tol) & (self.edge_layer == lp.layer)
        on &= (np.abs(e0 @ lp.t) <= lp.w / 2 + tol) & (np.abs(e1 @ lp.t) <= lp.w / 2 + tol)
        idx = np.nonzero(on)[0]
        return idx, np.sign(r.nrm[idx] @ lp.d)

    def _edge_sets(self):
        self.gap = []
        for p in self.ports:
            gap = self._line_edges(p.lp) if isinstance(p, GapPort) else None
            if gap is not None and not len(gap[0]):
                raise ValueError(f"port {p.name}: no mesh edges on the gap line")
            self.gap.append(gap)
        self.lumped_edges = [self._line_edges(el.lp) for el in self.lumped]
        for el, e in zip(self.lumped, self.lumped_edges):
            if not len(e[0]):
                raise ValueError(f"lumped element {el.name}: no mesh edges on its line")

    def _port_chain(self, p):
        return next(n for n, (v, _, _) in enumerate(self.chains) if v is p.via)

    # ---- per-frequency parts ------------------------------------------------------------------------------
    def green_tables(self, f):
        """{(a, b): GreenTableML} for every pair of metal interfaces at f, on the device (sim.py:736-745)."""
        with self.log.timer('Green tables', f"{len(self.layers)} interfaces @ {f / 1e9:.3f} GHz"):
            self.tables = green_tables(f, self.st, self.layers, self.rho_max, self.dmax)
            self.dev_tables = device_tables(self.tables, self.layers)
        return self.tables

    VIA_TABLE_PAIRS = 64

    def _via_geometry(self):
        """Each chain's axis, the chains of each span, every pair (m <= n) grouped by its two spans with its
        distance (a chain with its own via: the barrel radius), geometry only (sim.py:786-814)."""
        if getattr(self, '_via_geo', None) is not None:
            return self._via_geo
        ch = self.chains
        x = np.array([v.x for v, _, _ in ch], float)
        y = np.array([v.y for v, _, _ in ch], float)
        a = np.array([v.a for v, _, _ in ch], float)
        via = {}
        vid = np.array([via.setdefault(id(v), len(via)) for v, _, _ in ch], int)
        span_of = [(kb, kt) for _, kb, kt in ch]
        spans = sorted(set(span_of))
        si = np.array([spans.index(s) for s in span_of], int)
        by_span = {s: np.flatnonzero(si == i) for i, s in enumerate(spans)}
        m, n = np.triu_indices(len(ch))
        d = np.where(vid[m] == vid[n], a[m], np.maximum(np.hypot(x[m] - x[n], y[m] - y[n]), 0.5 * (a[m] + a[n])))
        key = si[m] * len(spans) + si[n]
        order = np.argsort(key, kind='stable')
        groups = {}
        for gk in np.unique(key):
            sel = order[np.searchsorted(key[order], gk, 'left'):np.searchsorted(key[order], gk, 'right')]
            groups[(spans[gk // len(spans)], spans[gk % len(spans)])] = (m[sel], n[sel], d[sel])
        self._via_geo = dict(x=x, y=y, by_span=by_span, groups=groups, spans=spans)
        return self._via_geo

    def _via_blocks(self, f, QF):
        """B (Ne x Nv) and C (Nv x Nv), device complex128 (sim.py:816-872): the fans' end-charge potentials (QF:
        the accurate G_q moments between every triangle and the fan triangles, device Nt x |F|), the smooth
        via -> strip kernels tested with the RWG divergence, the via-via reactions with their K0 part, elements
        in vias."""
        from .via import ViaKernels
        w = 2 * np.pi * f
        r, Ne, Nv = self.rwg, self.Ne, self.Nv
        dev = dv.device()
        B = torch.zeros((Ne, Nv), dtype=dv.c128, device=dev)
        Ch = np.zeros((Nv, Nv), complex)
        vk = ViaKernels(f, self.st, [(kb, kt) for _, kb, kt in self.chains], self.layers, self.rho_max)
        geo = self._via_geometry()
        ops = self._div_ops()
        if len(self.F):
            dF = torch.from_numpy(self.dF).to(dev, dv.c128)
            T = QF @ dF                                                     # (Nt, Nv) fan-charge potentials
            B += ops(T) / (1j * w)
            Ch += (self.dF.T @ T[torch.from_numpy(self.F).to(dev)].cpu().numpy()) / (1j * w)
        for span, ns in geo['by_span'].items():
            c = torch.zeros((len(r.tris), len(ns)), dtype=dv.c128, device=dev)
            for k in self.layers:
                idx = self.tri_idx[k]
                D = vk.D[(k, span)]
                c[torch.from_numpy(idx).to(dev)] = dv.radial_table_sum(self.P6[idx, :, 0], self.P6[idx, :, 1],
                                                                       self.W6[idx], geo['x'][ns], geo['y'][ns],
                                                                       D.v, D._dr)
            B[:, torch.from_numpy(ns).to(dev)] += ops(c)
        for (sm, sn), (m, n, dd) in geo['groups'].items():
            vals = (vk.R_table(sm, sn)(dd) if len(dd) > self.VIA_TABLE_PAIRS else vk.R(sm, sn, dd)) + vk.homog(sm, sn, dd)
            np.add.at(Ch, (m, n), vals)
            off = m != n
            np.add.at(Ch, (n[off], m[off]), vals[off])
        seen = set()
        for n, (v, kb, kt) in enumerate(self.chains):
            if v.Z is not None and id(v) not in seen:
                Ch[n, n] += v.Z(f)
            seen.add(id(v))
        return B, torch.from_numpy(Ch).to(dev)

    def _div_ops(self):
        """x (Nt, k) -> D^T x (Ne, k) on the device: +l/A on T+, -l/A on T- (base.py:47)."""
        if getattr(self, '_div', None) is None:
            r, dev = self.rwg, dv.device()
            tp, tm = torch.from_numpy(r.tp).to(dev), torch.from_numpy(r.tm).to(dev)
            ap = torch.from_numpy(r.l / r.area[r.tp]).to(dev, dv.c128)[:, None]
            am = torch.from_numpy(-r.l / r.area[r.tm]).to(dev, dv.c128)[:, None]
            self._div = lambda x: ap * x[tp] + am * x[tm]
        return self._div

    def lumped_entries(self, f):
        """(rows, cols, values) of the lumped elements' Z_el u u^T (fast.py:628-641)."""
        rows, cols, vals = [], [], []
        for el, (idx, sig) in zip(self.lumped, self.lumped_edges):
            u = sig * self.rwg.l[idx]
            R, Cc = np.meshgrid(idx, idx, indexing='ij')
            rows.append(R.ravel()); cols.append(Cc.ravel()); vals.append((el.Z(f) * np.outer(u, u)).ravel())
        if not rows:
            return None
        return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)

    # ---- excitation and port quantities (sim.py:988-1003) -------------------------------------------------
    def excitation(self):
        E = np.zeros((self.Ne + self.Nv, len(self.ports)), complex)
        for j, p in enumerate(self.ports):
            if isinstance(p, ViaPort):
                E[self.Ne + self._port_chain(p), j] = p.sign
            else:
                idx, sig = self.gap[j]
                E[idx, j] = sig * self.rwg.l[idx]
        return E

    def port_current(self, I, j):
        p = self.ports[j]
        if isinstance(p, ViaPort):
            return p.sign * I[self.Ne + self._port_chain(p)]
        idx, sig = self.gap[j]
        return ((sig * self.rwg.l[idx])[:, None] * I[idx]).sum(0)

    def currents(self, f, E):
        """Basis currents for excitation columns E (sim.py:1014-1031): a result that is not finite raises
        FloatingPointError (no CPU solver to fall back on)."""
        I = self.fast.solve_direct(f, E, precision=self.precision)
        if not np.isfinite(I).all():
            raise FloatingPointError(f"mom25d_amd: the direct solve at {f / 1e9:.4g} GHz is not finite "
                                     f"(N = {len(I)}): a degenerate mesh, or a port with no return path")
        return I

    def sparams(self, f, keep=False):
        """S at f for the lumped ports (sim.py:1045-1092)."""
        return self.sparams_from(f, self.currents(f, self.excitation()), keep)

    def sparams_from(self, f, I, keep=False):
        """S at f from the basis currents I of the excitation() columns."""
        n = len(self.ports)
        Ap, Bp = np.zeros((n, n), complex), np.zeros((n, n), complex)
        for i, p in enumerate(self.ports):
            Vp, Ip = np.eye(n)[i], self.port_current(I, i)
            Ap[i], Bp[i] = (Vp + p.zc * Ip) / (2 * np.sqrt(p.zc)), (Vp - p.zc * Ip) / (2 * np.sqrt(p.zc))
        S = Bp @ np.linalg.inv(Ap)
        self.last = dict(f=f, I=I if keep else None, A=Ap, B=Bp)
        return S

    def sweep(self, freqs):
        """S, reciprocity and passivity per frequency (sim.py:1094-1111)."""
        out = SimpleNamespace(f=np.asarray(freqs), S=[], recip=[], passiv=[])
        for f in freqs:
            S = self.sparams(f)
            out.S.append(S)
            out.recip.append(np.max(np.abs(S - S.T)))
            out.passiv.append(np.max(np.linalg.svd(S, compute_uv=False)))
        for k in ('S', 'recip', 'passiv'):
            setattr(out, k, np.array(getattr(out, k)))
        return out
