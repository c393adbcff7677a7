"""Seeded synthetic boards of the benchmark configurations (rectangles on a layered stack, vias, lumped via ports).

  microstrip()          C1: a 1.5 mm x 30 mm strip on FR4 (er 4.4, tand 0.02, h 0.8 mm) over ground, 50-ohm via
                        ports at both ends, dmax 0.3 mm
  seeded_board(seed)    C2: 4 layers (er U(3.5, 4.6), h U(0.1, 0.4) mm, tand 0.02, ground at the bottom), 12
                        Manhattan traces (width U(0.15, 0.5) mm) on interfaces 2 and 4, 6 vias (r 0.15 mm) joining
                        them, 4 via ports to ground
  coupled_lines(seed)   C3: a 2-layer board with two parallel 0.3 mm strips 0.2 mm apart, 40 mm long, 4 ports
  trace_board(seed)     a one-layer board of rng.integers(2, 17) traces from the C1 generator
  uneven_board(seed)    C4 and C5: the same kind of board (narrower, shorter traces) meshed to a drawn size N ~
                        U(n_range) (the mesh density, and where needed the cells across a trace, solved for): C4
                        U(500, 6000), C5 U(2000, 20000)
Each returns (stack, shapes, ports, vias, kwargs) for Simulation(stack, shapes, ports, vias, **kwargs); the seeds
and the drawn parameters are deterministic (numpy default_rng).
"""
import numpy as np

from .multilayer import Layer, LayerStack
from .sim import Via, ViaPort

MM = 1e-3


def microstrip(dmax=0.3 * MM, ncross=4):
    st = LayerStack([Layer(0.8 * MM, 4.4, 0.02)])
    L, w = 30 * MM, 1.5 * MM
    shapes = {1: [(-L / 2, -w / 2, L / 2, w / 2)]}
    ports = [ViaPort(Via(-L / 2, 0.0, 0, 1, 0.15 * MM, name='P1'), 50.0, 'P1'),
             ViaPort(Via(L / 2, 0.0, 0, 1, 0.15 * MM, name='P2'), 50.0, 'P2')]
    return st, shapes, ports, [], dict(dmax=dmax, ncross=ncross, w_ref={1: w})


def _trace(rng, start, w, nseg=2, seg=(4 * MM, 10 * MM), box=9 * MM):
    """A Manhattan polyline of nseg segments from start, alternating x / y, as overlapping rectangles: (rects, end)."""
    rects, p = [], np.asarray(start, float)
    axis = int(rng.integers(0, 2))
    for _ in range(nseg):
        length = rng.uniform(*seg) * rng.choice([-1.0, 1.0])
        q = p.copy()
        q[axis] = np.clip(p[axis] + length, -box, box)
        if abs(q[axis] - p[axis]) < 1 * MM:
            q[axis] = p[axis] - np.sign(length) * 2 * MM
        lo, hi = np.minimum(p, q), np.maximum(p, q)
        if axis == 0:
            rects.append((lo[0] - w / 2, p[1] - w / 2, hi[0] + w / 2, p[1] + w / 2))
        else:
            rects.append((p[0] - w / 2, lo[1] - w / 2, p[0] + w / 2, hi[1] + w / 2))
        p, axis = q, 1 - axis
    return rects, p


def _snap(x, step=0.05 * MM):
    return np.round(np.asarray(x) / step) * step


def seeded_board(seed=2, dmax=0.25 * MM, ncross=3, ntrace=12, nvia=6, nport=4):
    rng = np.random.default_rng(seed)
    st = LayerStack([Layer(rng.uniform(0.1, 0.4) * MM, rng.uniform(3.5, 4.6), 0.02) for _ in range(4)])
    ka, kb = 2, 4
    shapes, vias, ports = {ka: [], kb: []}, [], []
    per = ntrace // 2
    for i in range(per):
        w1, w2 = _snap(rng.uniform(0.15, 0.5) * MM, 0.01 * MM), _snap(rng.uniform(0.15, 0.5) * MM, 0.01 * MM)
        start = _snap(rng.uniform(-8, 8, 2) * MM)
        r1, mid = _trace(rng, start, w1)
        r2, _ = _trace(rng, mid, w2)
        shapes[ka] += [tuple(_snap(r, 0.01 * MM)) for r in r1]
        shapes[kb] += [tuple(_snap(r, 0.01 * MM)) for r in r2]
        if i < nvia:
            vias.append(Via(float(mid[0]), float(mid[1]), ka, kb, 0.15 * MM, name=f'V{i + 1}'))
        if i < nport:
            ports.append(ViaPort(Via(float(start[0]), float(start[1]), 0, ka, 0.15 * MM, name=f'P{i + 1}'), 50.0,
                                 f'P{i + 1}'))
    return st, shapes, ports, vias, dict(dmax=dmax, ncross=ncross, w_ref={ka: 0.5 * MM, kb: 0.5 * MM})


def coupled_lines(seed=3, dmax=0.25 * MM, ncross=4):
    rng = np.random.default_rng(seed)
    st = LayerStack([Layer(rng.uniform(0.1, 0.4) * MM, rng.uniform(3.5, 4.6), 0.02) for _ in range(2)])
    L, w, gap = 40 * MM, 0.3 * MM, 0.2 * MM
    yc = (w + gap) / 2
    shapes = {2: [(-L / 2, yc - w / 2, L / 2, yc + w / 2), (-L / 2, -yc - w / 2, L / 2, -yc + w / 2)]}
    ports = [ViaPort(Via(x, y, 0, 2, 0.1 * MM, name=n), 50.0, n)
             for n, (x, y) in zip(('P1', 'P2', 'P3', 'P4'), ((-L / 2, yc), (L / 2, yc), (-L / 2, -yc), (L / 2, -yc)))]
    return st, shapes, ports, [], dict(dmax=dmax, ncross=ncross, w_ref={2: w})


def trace_board(seed, dmax=0.3 * MM, ncross=4, ntrace=None):
    """The C1 generator with rng.integers(2, 17) traces of width U(0.3, 1.5) mm on its one interface; a via port
    to ground at each trace's start."""
    rng = np.random.default_rng(seed)
    st = LayerStack([Layer(0.8 * MM, 4.4, 0.02)])
    n = int(ntrace or rng.integers(2, 17))
    shapes, ports = {1: []}, []
    for i in range(n):
        w = _snap(rng.uniform(0.3, 1.5) * MM, 0.01 * MM)
        start = _snap(rng.uniform(-12, 12, 2) * MM)
        r, _ = _trace(rng, start, w, nseg=int(rng.integers(1, 4)), seg=(5 * MM, 15 * MM), box=15 * MM)
        shapes[1] += [tuple(_snap(x, 0.01 * MM)) for x in r]
        if i < 4:
            ports.append(ViaPort(Via(float(start[0]), float(start[1]), 0, 1, 0.15 * MM, name=f'P{i + 1}'), 50.0,
                                 f'P{i + 1}'))
    return st, shapes, ports, [], dict(dmax=dmax, ncross=ncross, w_ref={1: 1.5 * MM})


def uneven_board(seed, n_range=(500, 6000), ncross=3):
    """A one-layer FR4 board (the C1 stack) of rng.integers(2, 17) traces of width U(0.3, 0.6) mm and 1-3 segments
    of U(1, 2.5) mm, a via port to ground at the first four traces' starts, meshed to a size drawn from
    U(n_range): the mesh density dmax is solved for (the unknowns grow as 1 / dmax along the traces), between 0.02
    and 2 mm, with fewer cells across the traces where even the coarsest mesh has too many unknowns, and coarsened
    until the size is within the range's top. Returns the usual (stack, shapes, ports, vias, kwargs) with
    kwargs['target_n'] the drawn size."""
    from .sim import Simulation
    rng = np.random.default_rng(seed)
    st = LayerStack([Layer(0.8 * MM, 4.4, 0.02)])
    n = int(rng.integers(2, 17))
    shapes, ports = {1: []}, []
    for i in range(n):
        w = _snap(rng.uniform(0.3, 0.6) * MM, 0.01 * MM)
        start = _snap(rng.uniform(-12, 12, 2) * MM)
        r, _ = _trace(rng, start, w, nseg=int(rng.integers(1, 4)), seg=(1.0 * MM, 2.5 * MM), box=15 * MM)
        shapes[1] += [tuple(_snap(x, 0.01 * MM)) for x in r]
        if i < 4:
            ports.append(ViaPort(Via(float(start[0]), float(start[1]), 0, 1, 0.15 * MM, name=f'P{i + 1}'), 50.0,
                                 f'P{i + 1}'))
    target = float(rng.uniform(*n_range))

    def size(dmax, nc):
        sim = Simulation(st, shapes, ports, [], verbose=False, mesh_only=True, dmax=dmax, ncross=nc,
                         w_ref={1: 0.6 * MM})
        return sim.Ne + sim.Nv
    for nc in range(ncross, 0, -1):             # fewer cells across the traces where even coarse ones are too many
        dmax = 0.3 * MM
        for _ in range(5):
            dmax = float(np.clip(dmax * size(dmax, nc) / target, 0.02 * MM, 2.0 * MM))
        if size(dmax, nc) <= 1.25 * target:
            break
    for _ in range(60):                         # never above the range's top
        if dmax >= 2.0 * MM or size(dmax, nc) <= n_range[1]:
            break
        dmax = min(2.0 * MM, 1.03 * dmax)
    return st, shapes, ports, [], dict(dmax=dmax, ncross=nc, w_ref={1: 0.6 * MM}, target_n=int(target))
