"""One call for a program that has a board and wants S-parameters back (the reference's src/mom25d/api.py:22-100).

    from mom25d_amd.api import solve_board
    res = solve_board(board, np.linspace(1e9, 6e9, 401), dmax=0.3e-3)
    res.S, res.names, res.zref, res.solves, res.seconds

The reference reads kicad2ems's board files (polygons, a constrained-Delaunay mesher, shapely); this build meshes
rectangles, so `board` is a rectangle board: the (stack, shapes, ports, vias, kwargs) tuple of mom25d_amd.boards, or
a dict (or the path of a JSON file) in this build's rectangle format:

    {"layers": [[h, er, tand], ...],                  bottom to top (metres)
     "top": "open" | "pec", "bottom": "pec" | "open", "top_er": 1.0, "bottom_er": 1.0,
     "shapes": {"1": [[x0, y0, x1, y1], ...], ...},   rectangles per interface
     "vias": [[x, y, k_bot, k_top, radius], ...],
     "ports": [{"via": [x, y, k_bot, k_top, radius], "z0": 50, "name": "P1"}, ...],
     "dmax": 3e-4, "ncross": 4}                        optional mesh settings

A kicad2ems file (its "nets", "pads" or polygon outlines) raises ValueError naming what is missing. The sweep is
the adaptive one of run.solve_band on the requested frequencies.
"""
import json
import os
import time
from dataclasses import dataclass, field

import numpy as np

CELLS = 30.0


@dataclass
class Result:
    f: np.ndarray                   # Hz: the requested frequencies
    S: np.ndarray                   # (nf, n, n) at zref
    names: list                     # port names
    zref: list                      # output reference impedance per port, ohms
    ports: list                     # [dict(name, kind='lumped', zc)]
    unknowns: int                   # RWG edges + via chains
    triangles: int
    solver: str                     # 'direct' (dense LU on the GPU)
    precision: str                  # 'double' or 'mixed'
    mesher: str                     # 'tensor'
    solves: int                     # EM solves of the sweep
    solved: list                    # the frequencies solved (Hz)
    seconds: float                  # model, mesh and sweep
    core: bool                      # the GPU kernels (always)
    notes: list = field(default_factory=list)
    afs: list = field(default_factory=list)
    afs_tol: float = None
    model: object = None            # the rectangle board as read
    sim: object = None              # the Simulation


def _rect_board(obj):
    """(stack, shapes, ports, vias, kwargs) from the rectangle format (a dict)."""
    from .multilayer import Layer, LayerStack
    from .sim import Via, ViaPort
    if any(k in obj for k in ('nets', 'pads', 'outline', 'polygons', 'zones')) or 'layers' not in obj:
        raise ValueError("solve_board: this looks like a kicad2ems board file; its polygon front end (shapely and a "
                         "constrained-Delaunay mesher) is not part of this build, which takes rectangle boards "
                         "(mom25d_amd.api's format)")
    st = LayerStack([Layer(float(h), float(er), float(td)) for h, er, td in obj['layers']],
                    top=obj.get('top', 'open'), top_er=obj.get('top_er', 1.0), bottom=obj.get('bottom', 'pec'),
                    bottom_er=obj.get('bottom_er', 1.0))
    shapes = {int(k): [tuple(map(float, r)) for r in v] for k, v in obj['shapes'].items()}
    via = lambda v, name='': Via(float(v[0]), float(v[1]), int(v[2]), int(v[3]), float(v[4]) if len(v) > 4 else 0.15e-3,
                                 name=name)
    vias = [via(v, f'V{i + 1}') for i, v in enumerate(obj.get('vias', []))]
    ports = [ViaPort(via(p['via'], p.get('name', f'P{i + 1}')), float(p.get('z0', 50.0)), p.get('name', f'P{i + 1}'))
             for i, p in enumerate(obj.get('ports', []))]
    kw = {k: obj[k] for k in ('dmax', 'ncross') if k in obj}
    return st, shapes, ports, vias, kw


def build_board(board, dmax=0.3e-3, metal=None, wideband=None, log=print, title=None, mesh_only=False, **ignored):
    """(model, Simulation) of a rectangle board (see the module's note)."""
    from .sim import Simulation
    if isinstance(board, str):
        with open(board) as fh:
            board = json.load(fh)
    if isinstance(board, dict):
        model = board
        st, shapes, ports, vias, kw = _rect_board(board)
    else:
        st, shapes, ports, vias, kw = board
        model = None
    kw = dict(kw)
    kw.pop('target_n', None)
    kw.setdefault('dmax', dmax)
    if wideband:
        for l in st.layers:
            l.fref = float(wideband)
    sim = Simulation(st, shapes, ports, vias, metal=metal, verbose=False, title=title or '', mesh_only=mesh_only, **kw)
    return model, sim


def solve_board(board, freq, dmax=0.3e-3, plane_factor=3.0, port='auto', pairs=False, mesher='auto', metal=None,
                ground=None, max_solves=60, afs_tol=5e-3, log=print, progress=None, on_step=None, core='auto',
                quads=False, zone=None, signal_vias='round', stitching_vias='auto', lines=None, keep_at=(),
                quad_tol=35.0, coarsen=None, planes='pec', stitching_pads='drop', wideband=None, cells=CELLS):
    """S-parameters of a rectangle board on `freq` (Hz) (api.py:50-100): the adaptive sweep through at most
    max_solves solves to afs_tol, evaluated on freq. port: 'auto' or 'via' (lumped via ports; wave ports are not
    part of this build); metal: a sim.Metal or None; wideband: the frequency the layers' er and tand hold at
    (Djordjevic-Sarkar elsewhere). Options that ask for the polygon front end's features (pairs, quads, zone, lines,
    coarsen, meshed planes, another mesher) raise ValueError."""
    from .run import solve_band
    wanted = dict(port=(port, ('auto', 'via')), pairs=(pairs, (False,)), mesher=(mesher, ('auto', 'tensor')),
                  quads=(quads, (False,)), zone=(zone, (None,)), lines=(lines, (None,)), coarsen=(coarsen, (None,)),
                  planes=(planes, ('pec',)), ground=(ground, (None,)))
    for name, (value, ok) in wanted.items():
        if value not in ok:
            raise ValueError(f"solve_board: {name}={value!r} is not supported by this build (rectangle boards, "
                             "lumped via ports)")
    freq = np.asarray(freq, float)
    t0 = time.time()
    model, sim = build_board(board, dmax=dmax, metal=metal, wideband=wideband, log=log,
                             title=os.path.basename(board) if isinstance(board, str) else 'board')
    zref = [p.zc for p in sim.ports]
    f, S, info = solve_band(sim, float(freq[0]), float(freq[-1]), points=len(freq), zref=zref,
                            max_solves=max_solves, afs_tol=afs_tol, log=log, progress=progress, on_step=on_step,
                            keep_at=keep_at, freq=freq)
    if len(f) != len(freq) or not np.allclose(f, freq):
        S = np.apply_along_axis(lambda s: np.interp(freq, f, s.real) + 1j * np.interp(freq, f, s.imag), 0, S)
    return Result(f=freq, S=S, names=list(info['names']), zref=list(info['zref']),
                  ports=[dict(name=p.name, kind='lumped', zc=float(p.zc)) for p in sim.ports],
                  unknowns=int(sim.Ne + sim.Nv), triangles=int(len(sim.rwg.tris)), solver='direct',
                  precision=sim.precision, mesher='tensor', solves=int(info['solves']), solved=list(info['solved']),
                  seconds=time.time() - t0, core=True, notes=list(sim.log.notes), model=model, sim=sim,
                  afs=list(info['afs']), afs_tol=info['afs_tol'])
