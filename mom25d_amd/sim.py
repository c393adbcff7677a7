"""The simulation object: a layered stack, Manhattan copper on its interfaces, vias, lumped ports and elements ->
a mesh, the RWG and via-chain unknowns, and S-parameters per frequency (the reference's src/mom25d/sim.py, for
boards given as rectangles).

    sim = Simulation(stack, {1: [(x0, y0, x1, y1), ...]}, ports=[ViaPort(Via(x, y, 0, 1))], dmax=0.3e-3)
    S = sim.sparams(f)

The unknowns are the mesh's RWG edges, then one current per via chain (a via split where it touches metal). Every
frequency is solved on the GPU by mom25d_amd.fast.FastOperator. The constructor takes the reference's arguments;
those that ask for what a rectangle board does not have (polygon meshers, quad cells, line functions, zones, PEC
planes with holes or slots) or for a CPU solver raise ValueError.
"""
from types import SimpleNamespace

import numpy as np
import torch

from .base import MU0, RWG, geometry, gram_entries
from .mesh import contains, find_node, tensor_mesh
from .multilayer import RunLog, device_tables, green_tables
from . import device as dv

GPU_MIN_N = 3000          # mixed precision (complex64 LU + FP64 refinement) above this many unknowns
REUSE_KEEP = 4            # kept factors a solver holds at most (mom25d_amd.fast.REUSE_KEEP)
WORK_BYTES = 0.4e9        # the LU's and the solves' workspaces, the tables, the mesh and the near plan, bounded


def _as_function(Z):
    return Z if callable(Z) else (lambda f, Z=Z: complex(Z))


def rlc(R=0.0, L=0.0, C=None, parallel=False):
    """f -> Z of a resistor, inductor and capacitor in series (default) or in parallel (sim.py:47-54)."""
    def series(f):
        s = 2j * np.pi * f
        return R + s * L + (1 / (s * C) if C else 0)

    def shunt(f):
        s = 2j * np.pi * f
        return 1 / ((1 / R if R else 0) + (1 / (s * L) if L else 0) + (s * C if C else 0))
    return shunt if parallel else series


def available_memory():
    """Bytes of host memory available now (the kernel's MemAvailable)."""
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
    complex64 factor (8 n^2) and up to `keep` kept factors of nearby frequencies (8 n^2 each); 'lean' - the complex64
    Z factored in place and the kept factors (the refinement without Z); plus WORK_BYTES."""
    nn = float(n) * n
    if precision == 'double':
        return 16.0 * nn + WORK_BYTES
    if precision == 'lean':
        return 8.0 * nn * (1 + keep) + WORK_BYTES
    return 24.0 * nn + 8.0 * nn * keep + WORK_BYTES


def memory_plan(n, has_core=True, avail=None, gpu=None, precision='auto', solver='auto', n_lu=None, has_blr=False,
                keep=REUSE_KEEP):
    """How an n-unknown solve runs and what it holds: dict(solver, precision, path, host, gpu, fits), the
    reference's plan (src/mom25d/sim.py:85-151) for this build, whose direct solver keeps everything on the GPU.

    gpu: None, (name, free bytes) or (name, free bytes, the LU's bytes, fits) as the reference's gpu_info gives.
    precision 'auto': 'mixed' above GPU_MIN_N unknowns of the LU (n_lu, default n), else 'double'; a requested
    precision is kept. path: 'hybrid' (complex64 LU and FP64 refinement with the FP64 Z on the GPU), 'gpu-stream'
    (mixed, only the complex64 Z stored: the refinement multiplies without Z - where the hybrid path does not fit
    and this one does) or 'double' (complex128 LU). gpu (the result): the device bytes; host: the bytes the host
    holds (the mesh, geometry and currents, O(n)); fits: whether the device bytes fit 0.9 x the free device memory.
    The solver is 'direct' (dense LU) whatever is asked: the compressed solvers (BLR, ACA) are not part of this
    build, and a request that does not fit is reported as such (fits False) for the caller to queue or refuse."""
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
    path = 'hybrid' if precision == 'mixed' else 'double'
    if precision == 'mixed' and gpu is not None and dev > 0.9 * free:
        lean = device_bytes(n, 'lean', keep)
        if lean <= 0.9 * free:
            path, dev = 'gpu-stream', lean
    fits = gpu is not None and dev <= 0.9 * free
    host = 2e4 * n + 64 * 16.0 * n + 0.2e9
    return dict(solver='direct', precision=precision, path=path, host=host, gpu=dev, fits=fits)


# ---- board objects ------------------------------------------------------------------------------------------
class Via:
    """A plated via on axis (x, y) from interface k_bot up to k_top, barrel radius; Z: an element in its lowest
    chain (sim.py:154-160)."""

    def __init__(self, x, y, k_bot, k_top, radius=0.15e-3, Z=None, name='', ring=True):
        self.x, self.y, self.k_bot, self.k_top, self.a, self.name = x, y, k_bot, k_top, radius, name
        self.ring = ring
        self.Z = None if Z is None else _as_function(Z)


class LinePort:
    """A wave port on a straight feed (geo.py:24-60): a delta-gap source across the feed at `center` (interface
    `layer`, `width` wide), `direction` pointing into the device along the feed, the reference plane ref_dist
    beyond the gap. The feed's current is sampled on cuts across it every `step` (dmax if None) from s0 (2 step if
    None) to s1 (ref_dist if None) from the gap, and fitted by a forward and a backward wave of one propagation
    constant (fitted, or gamma_fn(f) when set); zc is the reference impedance. It is also the geometry of a gap
    port's or a series element's line (t = d turned by 90 degrees)."""
    kind = 'wave'

    def __init__(self, layer, center, direction, width, ref_dist=0.0, s0=None, step=None, zc=50.0, name='', s1=None):
        d = np.asarray(direction, float)
        self.d = d / np.linalg.norm(d)
        self.t = np.array([-self.d[1], self.d[0]])
        self.layer, self.c, self.w, self.ref, self.name = layer, np.asarray(center, float), width, ref_dist, name
        self.s0, self.step, self.s1, self.zc = s0, step, s1, float(zc)
        self.zc_fn = self.gamma_fn = None

    @property
    def axis_aligned(self):
        return bool(np.isclose(self.d, 0.0, atol=1e-12).any())

    def line(self, s=0.0):
        """End points of the cut across the feed at distance s from the gap."""
        p = self.c + self.d * s
        return p - self.t * self.w / 2, p + self.t * self.w / 2

    def cuts(self, dmax):
        """The sample positions along the feed: every step back from s1 to no closer than s0."""
        step = self.step or dmax
        s0 = 2 * step if self.s0 is None else self.s0
        s1 = self.ref if self.s1 is None else self.s1
        n = int(np.floor((s1 - s0) / step + 1e-9))
        return s1 - step * np.arange(n, -1, -1)


FIT_SPAN = 0.4          # rad: the phase a wave port's cuts must span for a fitted propagation constant to be trusted


def wave_gamma(Ik, step):
    """The propagation constant of samples Ik (ncut x ncol) of I(s) = a e^(-gamma s) + b e^(gamma s) at uniform
    spacing step: any such sequence obeys I[k-1] + I[k+1] = 2 cosh(gamma step) I[k], so cosh(gamma step) is the
    least-squares solution of that relation over every interior cut and column; gamma with Re, Im >= 0."""
    Ik = np.asarray(Ik, complex).reshape(len(Ik), -1)
    lhs, rhs = 2 * Ik[1:-1].ravel(), (Ik[:-2] + Ik[2:]).ravel()
    ch = np.vdot(lhs, rhs) / np.vdot(lhs, lhs)
    g = np.arccosh(complex(ch)) / step
    return complex(abs(g.real), abs(g.imag))


def wave_split(Ik, s, gamma):
    """(a, b) per column with Ik[:, j] ~ a_j e^(-gamma s) + b_j e^(gamma s), least squares over the cuts s."""
    s = np.asarray(s, float)
    M = np.stack([np.exp(-gamma * s), np.exp(gamma * s)], 1)
    coef = np.linalg.lstsq(M, np.asarray(Ik, complex).reshape(len(s), -1), rcond=None)[0]
    return coef[0], coef[1]


class GapPort:
    """A lumped port across a strip (a delta gap at its line) (sim.py:163-175)."""
    kind = 'gap'

    def __init__(self, layer, center, direction, width, z0=50.0, name=''):
        self.lp, self.layer, self.zc, self.name = LinePort(layer, center, direction, width), layer, z0, name


class ViaPort:
    """A lumped port in a via: an EMF in its lowest chain, driving current up (sign -1: down) (sim.py:178-184)."""
    kind = 'via'

    def __init__(self, via, z0=50.0, name='', sign=1):
        self.via, self.zc, self.name, self.layer, self.sign = via, z0, name, None, float(sign)


class SeriesLumped:
    """An element in series across a strip, at its line (sim.py:221-225)."""

    def __init__(self, layer, center, direction, width, Z, name=''):
        self.lp, self.layer, self.Z, self.name = LinePort(layer, center, direction, width), layer, _as_function(Z), name


class Metal:
    """Conductor loss through a surface impedance (sim.py:228-270): a sheet of conductivity sigma and thickness t
    has Zs = (k / sigma) coth(k t), k = (1 + j) / delta (1 / (sigma t) at DC, (1 + j) / (sigma delta) when thick).
    rough (m RMS): the resistance scaled by Hammerstad and Jensen's 1 + (2 / pi) atan(1.4 (rough / delta)^2).
    finish [(sigma, t), ...] outermost first, on the interfaces of `on` ({interface: share}): there Zs is scaled by
    1 + share (F - 1), F the finished face's impedance over the bare face's (each finish layer a line section of
    impedance k_i / sigma_i loading the one beneath it)."""

    def __init__(self, sigma=5.8e7, t=35e-6, rough=0.0, finish=(), on=None):
        self.sigma, self.t, self.rough = sigma, t, rough
        self.finish, self.on = [tuple(x) for x in finish], dict(on or {})

    @staticmethod
    def _k(f, sigma):
        return (1 + 1j) * np.sqrt(np.pi * f * MU0 * sigma)

    def _bare(self, f):
        k = self._k(f, self.sigma)
        return k / self.sigma / np.tanh(k * self.t)

    def zs(self, f, k=None):
        """Zs (ohms) at f; k: the interface (its finish, if any)."""
        z = self._bare(f)
        if self.rough:
            depth = 1 / np.sqrt(np.pi * f * MU0 * self.sigma)
            z = z.real * (1 + 2 / np.pi * np.arctan(1.4 * (self.rough / depth) ** 2)) + 1j * z.imag
        share = self.on.get(k, 0.0) if self.finish else 0.0
        return z * (1 + share * (self.face_factor(f) - 1)) if share else z

    def face_factor(self, f):
        """The finished face's impedance over the bare copper face's at f."""
        bare = z = self._bare(f)
        for sigma, t in reversed(self.finish):           # inner finish layers first
            zc, th = self._k(f, sigma) / sigma, np.tanh(self._k(f, sigma) * t)
            z = zc * (z + zc * th) / (zc + z * th)
        return z / bare

    def zs_edges(self, f, edge_layer):
        """Zs of every RWG edge, by its interface: one number unless a finish covers some interfaces."""
        if not (self.finish and self.on):
            return self.zs(f)
        return np.array([self.zs(f, k) if k in self.on else self.zs(f) for k in edge_layer], complex)


BALANCE_FLOOR = 1e-2


def balance(S, exact=None):
    """(S made reciprocal, port scales d): S_ij d_j / d_i with the log-scales d chosen by weighted least squares so
    that the scaled S_ij and S_ji agree in magnitude (weights |S_ij S_ji|, a floor that keeps d near 1 where a port
    couples weakly; `exact` ports keep d = 1), then the symmetric part (sim.py:1121-1141)."""
    mag = np.abs(S)
    wgt = mag * mag.T
    np.fill_diagonal(wgt, 0.0)
    with np.errstate(divide='ignore', invalid='ignore'):
        half_log = np.where(wgt > 0, 0.5 * np.log(mag / mag.T), 0.0)
    lap = np.diag(wgt.sum(1) + BALANCE_FLOOR) - wgt
    if exact is not None and np.any(exact):
        lap = lap + np.diag(np.where(exact, 1e6, 0.0))
    d = np.exp(np.linalg.solve(lap, (wgt * half_log).sum(1)))
    Sb = S * d[None, :] / d[:, None]
    return 0.5 * (Sb + Sb.T), d


# ---- the simulation -----------------------------------------------------------------------------------------
_UNSUPPORTED = dict(quads=False, zone=None, dmax_out=None, line_basis=None, coarsen=None, frills=(), slots=(),
                    dmax_slot=None)


class Simulation:
    """A board of rectangles on the stack's interfaces, meshed and set up for the direct solver on the GPU."""

    def __init__(self, stack, shapes, ports=(), vias=(), lumped=(), dmax=0.3e-3, dmax_layer=None, w_ref=None,
                 ncross=6, grade=1.0, near_fac=5.0, mesher='auto', metal=None, solver='dense', aca_tol=1e-4,
                 refine=1.0e-3, precision='auto', gpu='auto', verbose=True, title='', mesh_only=False, core='auto',
                 quads=False, quad_tol=35.0, zone=None, dmax_out=None, signal_vias='round', stitching_vias='auto',
                 signal=None, via_barrel=True, line_basis=None, line_w_max=2.0e-3, line_step=2.0, coarsen=None,
                 grow=0.5, structured='auto', gap_rings=(1 / 4, 1 / 8, 1 / 16), frills=(), slots=(), dmax_slot=None):
        asked = dict(quads=quads, zone=zone, dmax_out=dmax_out, line_basis=line_basis, coarsen=coarsen,
                     frills=tuple(frills), slots=tuple(slots), dmax_slot=dmax_slot)
        for key, value in asked.items():
            if value != _UNSUPPORTED[key] and not (isinstance(value, tuple) and not value):
                raise ValueError(f"Simulation: {key}={value!r} is not supported by this build (rectangle boards, "
                                 "RWG edges and via chains only)")
        if mesher not in ('auto', 'tensor'):
            raise ValueError(f"Simulation: mesher {mesher!r}: this build meshes rectangles with a tensor grid")
        if solver not in ('dense', 'direct', 'auto'):
            raise ValueError(f"Simulation: solver {solver!r}: this build has the direct (dense LU) solver only")
        if gpu in ('off', False):
            raise ValueError("Simulation: gpu='off' - this build solves on the GPU only")
        if precision not in ('auto', 'mixed', 'double'):
            raise ValueError(f"Simulation: precision {precision!r}")
        self.st, self.title, self.metal, self.log = stack, title, metal, RunLog(verbose)
        self.gpu, self.reuse, self.stream = gpu, 'auto', 'auto'
        self.dmax, self.ncross, self.grade, self.near_fac = dmax, ncross, grade, near_fac
        self.ports, self.lumped, self.vias = list(ports), list(lumped), list(vias)
        self.samples = [p.cuts(dmax) if isinstance(p, LinePort) else None for p in self.ports]
        for p in self.ports:
            if isinstance(p, ViaPort) and p.via not in self.vias:
                self.vias.append(p.via)
        self.layers = sorted(shapes)
        bad = [k for k in self.layers if k not in stack.metal_interfaces]
        if bad:
            raise ValueError(f"interface {bad[0]} cannot carry metal in this stack")
        self.shapes = {k: [tuple(map(float, r)) for r in shapes[k]] for k in self.layers}
        self._split_vias()
        with self.log.timer('mesh'):
            self._mesh(dmax_layer or {}, w_ref)
        self.Ne, self.Nv = self.rwg.Ne, len(self.chains)
        if mesh_only:
            return
        self._fans()
        self.log.note(f"mesh: {len(self.rwg.tris)} triangles, {self.Ne} edge + {self.Nv} via unknowns")
        self.g, self.q, self.tri_size = geometry(self.rwg, self.tri_layer, self.layers)
        self.P3, self.W3, self.P6, self.W6 = self.q['P3'], self.q['W3'], self.q['P6'], self.q['W6']
        span = self.rwg.nodes.max(0) - self.rwg.nodes.min(0)
        self.rho_max = 1.05 * np.hypot(*span) + 1e-4
        self.gap = [self._edges_on(p.lp, f"port {p.name}") if isinstance(p, GapPort) else
                    self._edges_on(p, f"port {p.name}") if isinstance(p, LinePort) else None for p in self.ports]
        self.cuts = [None if s is None else [self._edges_on(p, f"port {p.name}: the sample cut at {x * 1e3:.3f} mm", x)
                                             for x in s] for p, s in zip(self.ports, self.samples)]
        self.lumped_edges = [self._edges_on(el.lp, f"lumped element {el.name}") for el in self.lumped]
        self._gram = None
        n = self.Ne + self.Nv
        self.plan = memory_plan(n, gpu=('hip', self._gpu_free()), precision=precision)
        self.precision, self.solver = self.plan['precision'], 'direct'
        self.log.note(f"solver: direct on the GPU (N = {n}, {self.precision} precision, {self.plan['path']})")
        from .fast import FastOperator
        self.fast = FastOperator(self)

    # ---- set-up -----------------------------------------------------------------------------------------------
    def _split_vias(self):
        """Each via into chains between the interfaces where it touches metal (its ends must: on metal or on a PEC
        plane of the stack); via_attach: per via, the meshed interfaces it attaches to (sim.py:421-434)."""
        self.chains, self.via_attach = [], []
        for v in self.vias:
            def on_metal(k):
                return k in self.shapes and contains(self.shapes[k], v.x, v.y)
            for k in (v.k_bot, v.k_top):
                if not (self.st.is_pec(k) or on_metal(k)):
                    raise ValueError(f"via {v.name}: ends on M{k} where there is no metal")
            stops = [v.k_bot] + [k for k in range(v.k_bot + 1, v.k_top) if on_metal(k)] + [v.k_top]
            self.via_attach.append([k for k in stops if k in self.shapes])
            self.chains += [(v, lo, hi) for lo, hi in zip(stops[:-1], stops[1:])]

    def _mesh(self, dmax_layer, w_ref):
        """Each interface's copper meshed with the lines of its gap ports and elements and its via axes as grid
        lines; the triangles of all interfaces in one RWG basis; the via vertex on each interface a via attaches to."""
        nodes, tris, layer_of, self._via_vertex = [], [], [], {}
        base = 0
        for k in self.layers:
            across = {0: [], 1: []}                 # ports and elements: a line across x (a constant x) or y
            feeds = [(p.lp, [0.0]) for p in self.ports if isinstance(p, GapPort) and p.layer == k] + \
                [(p, [0.0] + list(sp)) for p, sp in zip(self.ports, self.samples) if isinstance(p, LinePort)
                 and p.layer == k] + [(el.lp, [0.0]) for el in self.lumped if el.layer == k]
            for lp, at in feeds:
                if not lp.axis_aligned:
                    raise ValueError("ports and elements must be axis-aligned on a rectangle board")
                axis = 0 if abs(lp.d[0]) > 0.5 else 1
                across[axis] += [lp.c[axis] + lp.d[axis] * x for x in at]
            here = [(i, v) for i, v in enumerate(self.vias) if k in self.via_attach[i]]
            dm = dmax_layer.get(k, self.dmax)
            ref = w_ref.get(k) if isinstance(w_ref, dict) else w_ref
            nd, tr = tensor_mesh(self.shapes[k], dm, self.ncross, 1.01 * max(ref or 0, 1.5 * dm), across[0],
                                 across[1], grade=self.grade, points=[(v.x, v.y) for _, v in here])
            for i, v in here:
                node = find_node(nd, v.x, v.y)
                if node < 0:
                    raise ValueError(f"via {v.name}: not on the M{k} mesh")
                self._via_vertex[(i, k)] = node + base
            nodes.append(nd)
            tris.append(tr + base)
            layer_of.append(np.full(len(tr), k))
            base += len(nd)
        self.rwg = RWG(np.vstack(nodes), np.vstack(tris))
        self.tri_layer = np.concatenate(layer_of)
        self.edge_layer = self.tri_layer[self.rwg.tp]
        self.tri_idx = {k: np.flatnonzero(self.tri_layer == k) for k in self.layers}

    def _fans(self):
        """The fan of a via on an interface: the triangles around its vertex there. F: every fan triangle once; dF
        (|F| x Nv): a chain's end charge spread uniformly over its fans, + at its bottom, - at its top (sim.py:
        525-535)."""
        r = self.rwg
        self.fan = {key: np.flatnonzero((r.tris == vtx).any(1)) for key, vtx in self._via_vertex.items()}
        self.F = np.unique(np.concatenate(list(self.fan.values()))) if self.fan else np.zeros(0, int)
        row = {t: j for j, t in enumerate(self.F)}
        self.dF = np.zeros((len(self.F), len(self.chains)))
        via_index = {id(v): i for i, v in enumerate(self.vias)}
        for n, (v, lo, hi) in enumerate(self.chains):
            for k, sign in ((lo, 1.0), (hi, -1.0)):
                tri = self.fan.get((via_index[id(v)], k))
                if tri is not None:
                    self.dF[[row[t] for t in tri], n] += sign / r.area[tri].sum()

    def _edges_on(self, lp, what, s=0.0, tol=1e-8):
        """(edges, signs) of the RWG edges lying on a feed's cut at distance s from its gap, within its width, on
        its interface: the sign is that of the edge's current direction along the feed (sim.py:697-728)."""
        r = self.rwg
        rel = r.ends - (lp.c + lp.d * s)[None, None, :]       # (Ne, 2 ends, 2)
        along, across = rel @ lp.d, rel @ lp.t
        on = (np.abs(along) < tol).all(1) & (np.abs(across) <= lp.w / 2 + tol).all(1) & (self.edge_layer == lp.layer)
        idx = np.flatnonzero(on)
        if not len(idx):
            raise ValueError(f"{what}: no mesh edges on its line")
        return idx, np.sign(r.nrm[idx] @ lp.d)

    @staticmethod
    def _gpu_free():
        if not torch.cuda.is_available():
            return 0.0
        return float(torch.cuda.mem_get_info()[0])

    def _port_chain(self, p):
        return next(n for n, (v, _, _) in enumerate(self.chains) if v is p.via)

    # ---- per-frequency parts -----------------------------------------------------------------------------------
    def green_tables(self, f):
        """{(a, b): GreenTableML} for every pair of metal interfaces at f, on the device (sim.py:736-745)."""
        with self.log.timer('Green tables', f"{len(self.layers)} interfaces @ {f / 1e9:.3f} GHz"):
            self.tables = green_tables(f, self.st, self.layers, self.rho_max, self.dmax)
            self.dev_tables = device_tables(self.tables, self.layers)
        return self.tables

    VIA_TABLE_PAIRS = 64

    def _via_geometry(self):
        """Geometry of the via blocks, once: each chain's axis, the chains of each span, and every chain pair
        (m <= n) grouped by its two spans with its distance (the barrel radius for two chains of one via, at least
        the mean radius otherwise) (sim.py:786-814)."""
        if getattr(self, '_via_geo', None) is not None:
            return self._via_geo
        ch = self.chains
        x, y, a = (np.array([getattr(v, k) for v, _, _ in ch], float) for k in ('x', 'y', 'a'))
        owner = np.unique([id(v) for v, _, _ in ch], return_inverse=True)[1].ravel()
        spans = sorted({(lo, hi) for _, lo, hi in ch})
        sp = np.array([spans.index((lo, hi)) for _, lo, hi in ch], int)
        m, n = np.triu_indices(len(ch))
        dist = np.where(owner[m] == owner[n], a[m], np.maximum(np.hypot(x[m] - x[n], y[m] - y[n]), 0.5 * (a[m] + a[n])))
        groups = {}
        for i in range(len(spans)):
            for j in range(len(spans)):
                sel = np.flatnonzero((sp[m] == i) & (sp[n] == j))
                if len(sel):
                    groups[(spans[i], spans[j])] = (m[sel], n[sel], dist[sel])
        self._via_geo = dict(x=x, y=y, by_span={s: np.flatnonzero(sp == i) for i, s in enumerate(spans)},
                             groups=groups, spans=spans)
        return self._via_geo

    def _via_blocks(self, f, QF):
        """B (Ne x Nv) and C (Nv x Nv), device complex128 (sim.py:816-872): the fans' end-charge potentials (QF:
        the accurate G_q moments between every triangle and the fan triangles, device Nt x |F|), the smooth
        via -> strip kernels tested with the RWG divergence, the via-via reactions with their K0 part, elements
        in vias and, with metal loss, the barrel's wall loss."""
        from .via import ViaKernels
        jw = 2j * np.pi * f
        Nt, dev = len(self.rwg.tris), dv.device()
        B = torch.zeros((self.Ne, self.Nv), dtype=dv.c128, device=dev)
        C = np.zeros((self.Nv, self.Nv), complex)
        vk = ViaKernels(f, self.st, [(lo, hi) for _, lo, hi in self.chains], self.layers, self.rho_max)
        geo, div, vd = self._via_geometry(), self._div_ops(), self._via_device()
        if len(self.F):
            pot = QF @ vd['dF']                                               # (Nt, Nv) the end charges' potential
            B += div(pot) / jw
            C += (self.dF.T @ pot[vd['F']].cpu().numpy()) / jw
        for span, ns in geo['by_span'].items():
            col = torch.zeros((Nt, len(ns)), dtype=dv.c128, device=dev)
            for k in self.layers:
                idx, px, py, w6 = vd['tri'][k]
                D = vk.D[(k, span)]
                col[idx] = dv.radial_table_sum(px, py, w6, vd['x'][span], vd['y'][span], D.v, D._dr)
            B[:, vd['ns'][span]] += div(col)
        for (sm, sn), (m, n, dd) in geo['groups'].items():
            remainder = vk.R_table(sm, sn)(dd) if len(dd) > self.VIA_TABLE_PAIRS else vk.R(sm, sn, dd)
            vals = remainder + vk.homog(sm, sn, dd)
            np.add.at(C, (m, n), vals)
            off = m != n
            np.add.at(C, (n[off], m[off]), vals[off])
        zs = self.metal.zs(f) if self.metal is not None else 0.0
        first = set()
        for n, (v, lo, hi) in enumerate(self.chains):
            if v.Z is not None and id(v) not in first:                       # an element: in its lowest chain
                C[n, n] += v.Z(f)
            first.add(id(v))
            C[n, n] += zs * (self.st.z[hi] - self.st.z[lo]) / (2 * np.pi * v.a)
        return B, torch.from_numpy(C).to(dev)

    def _via_device(self):
        """The via blocks' fixed arrays on the device, once: the fans' charge operator and triangles, each
        interface's triangles with their 6-point quadrature, each span's chains and their axes."""
        if getattr(self, '_via_dev', None) is None:
            dev, geo = dv.device(), self._via_geometry()
            on = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
            self._via_dev = dict(
                dF=on(self.dF, dv.c128), F=on(self.F, torch.int64),
                tri={k: (on(idx, torch.int64), on(self.P6[idx, :, 0], dv.f64), on(self.P6[idx, :, 1], dv.f64),
                         on(self.W6[idx], dv.f64)) for k, idx in ((k, self.tri_idx[k]) for k in self.layers)},
                x={s: on(geo['x'][ns], dv.f64) for s, ns in geo['by_span'].items()},
                y={s: on(geo['y'][ns], dv.f64) for s, ns in geo['by_span'].items()},
                ns={s: on(ns, torch.int64) for s, ns in geo['by_span'].items()})
        return self._via_dev

    def _div_ops(self):
        """x (Nt, k) -> D^T x (Ne, k) on the device: each edge's +l / A_+ on T+ and -l / A_- on T- (base.py:47)."""
        if getattr(self, '_div', None) is None:
            r, dev = self.rwg, dv.device()
            tp, tm = torch.from_numpy(r.tp).to(dev), torch.from_numpy(r.tm).to(dev)
            wp = torch.from_numpy(r.l / r.area[r.tp]).to(dev, dv.c128)[:, None]
            wm = torch.from_numpy(-r.l / r.area[r.tm]).to(dev, dv.c128)[:, None]
            self._div = lambda x: wp * x[tp] + wm * x[tm]
        return self._div

    def lumped_entries(self, f):
        """(rows, cols, values) added to Z at f: the lumped elements' Z_el u u^T (u: the element's edges, +-l) and,
        with metal loss, Zs (per edge's interface) times the RWG Gram matrix (fast.py:628-641); None if none."""
        parts = []
        for el, (idx, sig) in zip(self.lumped, self.lumped_edges):
            u = sig * self.rwg.l[idx]
            R, Cc = np.meshgrid(idx, idx, indexing='ij')
            parts.append((R.ravel(), Cc.ravel(), (el.Z(f) * np.outer(u, u)).ravel()))
        if self.metal is not None:
            if self._gram is None:
                self._gram = gram_entries(self.g, self.rwg.nodes, self.rwg.tris)
            r, c, g = self._gram
            zs = self.metal.zs_edges(f, self.edge_layer)
            parts.append((r, c, (zs[r] if np.ndim(zs) else zs) * g))
        if not parts:
            return None
        return tuple(np.concatenate([p[i] for p in parts]) for i in range(3))

    # ---- excitation, currents and S (sim.py:988-1111) ---------------------------------------------------------
    def excitation(self):
        """One column per port: a via port's EMF in its chain's row (its sign), a gap port's +-l on its edges."""
        E = np.zeros((self.Ne + self.Nv, len(self.ports)), complex)
        for j, p in enumerate(self.ports):
            if isinstance(p, ViaPort):
                E[self.Ne + self._port_chain(p), j] = p.sign
            else:
                idx, sig = self.gap[j]
                E[idx, j] = sig * self.rwg.l[idx]
        return E

    def port_current(self, I, j):
        """Port j's current for each excitation column of I: its chain's current (signed) or the current through
        its gap line (a wave port's: at its gap)."""
        p = self.ports[j]
        if isinstance(p, ViaPort):
            return p.sign * I[self.Ne + self._port_chain(p)]
        idx, sig = self.gap[j]
        return (sig * self.rwg.l[idx]) @ I[idx]

    def currents(self, f, E):
        """Basis currents for excitation columns E (sim.py:1014-1031); a result that is not finite raises
        FloatingPointError (there is no CPU solver to fall back on)."""
        I = self.fast.solve_direct(f, E, precision=self.precision)
        if not np.isfinite(I).all():
            raise FloatingPointError(f"mom25d_amd: the direct solve at {f / 1e9:.4g} GHz is not finite "
                                     f"(N = {len(I)}): a degenerate mesh, or a port with no return path")
        return I

    def cut_current(self, I, j):
        """The current through each sample cut of wave port j (ncut x ncol) (sim.py:1005-1007)."""
        return np.array([(sig * self.rwg.l[idx]) @ I[idx] for idx, sig in self.cuts[j]])

    def sparams(self, f, keep=False):
        """S at f (sim.py:1045-1092)."""
        return self.sparams_from(f, self.currents(f, self.excitation()), keep)

    def sparams_from(self, f, I, keep=False):
        """S at f from the basis currents I of the excitation() columns (each port driven in turn): a lumped
        port's incident and reflected waves a = (V + Zc I) / (2 sqrt Zc), b = (V - Zc I) / (2 sqrt Zc) with V the
        1 V EMF; a wave port's from its feed's forward and backward current waves (a e^(-gamma s), b e^(gamma s),
        s from its reference plane), sqrt(Zc) a and -sqrt(Zc) b; S = B A^-1, made reciprocal by balance() when it
        has wave ports. self.last keeps the fit: gamma and its residual per wave port."""
        for p in self.ports:
            if getattr(p, 'zc_fn', None) is not None:
                p.zc = float(p.zc_fn(f))
        n = len(self.ports)
        A, B = np.zeros((n, n), complex), np.zeros((n, n), complex)
        gam, resid = np.full(n, np.nan + 0j), np.zeros(n)
        V = np.eye(n)
        for i, p in enumerate(self.ports):
            rz = np.sqrt(p.zc)
            if isinstance(p, LinePort):
                s, Ik = self.samples[i], self.cut_current(I, i)
                known = p.gamma_fn is not None
                g = gam[i] = complex(p.gamma_fn(f)) if known else wave_gamma(Ik, abs(s[1] - s[0]))
                a, b = wave_split(Ik, s - p.ref, g)
                fit = np.exp(-g * (s - p.ref))[:, None] * a[None, :] + np.exp(g * (s - p.ref))[:, None] * b[None, :]
                resid[i] = np.linalg.norm(Ik - fit, axis=0).max() / np.linalg.norm(Ik, axis=0).max()
                A[i], B[i] = rz * a, -rz * b
                span = abs(g.imag) * abs(s[-1] - s[0])
                if not known and span < FIT_SPAN:
                    self.log.note(f"WARNING {f / 1e9:.2f} GHz port {p.name}: its cuts span {span:.2f} rad - the "
                                  "wave separation is unreliable (a longer straight feed, or a gap or via port)")
                if resid[i] > 1e-2:
                    self.log.note(f"WARNING {f / 1e9:.2f} GHz port {p.name}: wave-fit residual {resid[i]:.1e} - "
                                  "cuts too close to a discontinuity, or too short a feed at this frequency")
            else:
                Ip = self.port_current(I, i)
                A[i], B[i] = (V[i] + p.zc * Ip) / (2 * rz), (V[i] - p.zc * Ip) / (2 * rz)
        S = B @ np.linalg.inv(A)
        wave = np.array([isinstance(p, LinePort) for p in self.ports])
        scale = np.ones(n)
        if wave.any() and n > 1:
            S, scale = balance(S, ~wave)
        self.last = dict(f=f, I=I if keep else None, A=A, B=B, gamma=gam, resid=resid, scale=scale)
        return S

    def sweep(self, freqs):
        """S per frequency with its reciprocity (max |S - S^T|) and passivity (largest singular value)
        (sim.py:1094-1111)."""
        S = np.array([self.sparams(f) for f in freqs])
        return SimpleNamespace(f=np.asarray(freqs), S=S, recip=np.abs(S - S.transpose(0, 2, 1)).max((1, 2)),
                               passiv=np.array([np.linalg.svd(s, compute_uv=False).max() for s in S]))
