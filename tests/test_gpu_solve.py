"""The whole per-frequency solve on the GPU against the oracle's CPU restatement (tables, Z, S) on the
configuration boards that the oracle finishes in seconds, and size-independent properties on the 8k-unknown
board.

Tolerances: Z - row-relative 1e-12, and per-entry 1e-8 above 1e-4 of the row's largest entry. The fill kernels are
bit-exact given the same tables (test_gpu_kernels.py); here the tables are the GPU's own (the Hankel sums' J0 and
the spline's order of operations round differently from glibc / SciPy: the tables agree to 1e-12 of their largest
value), and far entries whose four half-pair terms cancel by up to 1e4 carry that difference to 2.5e-9 relative.
S - |dS| <= 1e-5 for the FP64 path, <= 1e-4 for the mixed path after refinement (the north star's bounds).
"""
import numpy as np
import pytest

from mom25d_amd import boards
from mom25d_amd.sim import Simulation
from oracle import greens as og
from oracle.solve import Solver, sparams as osparams
from tests.cases import oracle_board, port_rows

pytestmark = pytest.mark.gpu


def make(board, precision, **kw):
    st, sh, po, vi, bkw = board()
    bkw.update(kw)
    return Simulation(st, sh, po, vi, precision=precision, verbose=False, **bkw)


@pytest.fixture(scope='module')
def c1(gpu):
    sim = make(boards.microstrip, 'double')
    return sim, Solver(oracle_board(sim))


def rel_row(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b).max(axis=-1, keepdims=True)))


def rel_entry(a, b, floor):
    fl = floor * np.abs(b).max(axis=-1, keepdims=True)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), fl)))


def test_tables_match_oracle(c1):
    sim, osv = c1
    f = 3.3e9
    tg = sim.green_tables(f)
    to = og.green_tables(f, osv.b.stack, osv.b.layers, osv.rho_max, osv.b.dmax)
    for key in to:
        a, b = tg[key], to[key]
        for x, y in ((a.A_rest.cpu().numpy(), b.A_rest), (a.Q_rest.cpu().numpy(), b.Q_rest)):
            assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()
        for name in ('cA', 'cq', 'cAi', 'cqi', 'dimg', 'd'):
            assert abs(getattr(a, name) - getattr(b, name)) <= 1e-13 * max(abs(getattr(b, name)), 1e-300), name


def test_fill_matches_oracle(c1):
    sim, osv = c1
    f = 3.3e9
    Zg = sim.fast.fill(f).cpu().numpy()
    Zo = osv.fill(f)
    assert rel_row(Zg, Zo) < 1e-12
    assert rel_entry(Zg, Zo, 1e-4) < 1e-8


def test_microstrip_s_double_and_mixed_match_oracle(c1):
    sim, osv = c1
    rows, signs = port_rows(sim)
    zc = [p.zc for p in sim.ports]
    for f in (1e9, 3.5e9, 6e9):
        So = osparams(osv.solve(f), osv.b.E, zc, rows, signs)
        Sd = sim.sparams(f)
        assert np.abs(Sd - So).max() <= 1e-5, (f, np.abs(Sd - So).max())
        assert np.abs(Sd - Sd.T).max() < 1e-8                           # reciprocal
    sm = make(boards.microstrip, 'mixed')
    for f in (1e9, 3.5e9):
        So = osparams(osv.solve(f), osv.b.E, zc, rows, signs)
        Sm = sm.sparams(f)
        assert np.abs(Sm - So).max() <= 1e-4, (f, np.abs(Sm - So).max())
    assert any('mixed precision' in n or 'reused' in n for n in sm.log.notes)


def test_coupled_lines_match_oracle(gpu):
    sim = make(boards.coupled_lines, 'double')
    osv = Solver(oracle_board(sim))
    rows, signs = port_rows(sim)
    zc = [p.zc for p in sim.ports]
    f = 5e9
    So = osparams(osv.solve(f), osv.b.E, zc, rows, signs)
    Sd = sim.sparams(f)
    assert np.abs(Sd - So).max() <= 1e-5
    Sm = make(boards.coupled_lines, 'mixed').sparams(f)
    assert np.abs(Sm - So).max() <= 1e-4


def test_seeded_8k_board_double_vs_mixed(gpu):
    """C2 (N ~ 7.8k, vias between two interfaces, 4 via ports): mixed agrees with double to 1e-4, reciprocal and
    passive; a reused factor at a nearby frequency gives the fresh answer."""
    sd = make(boards.seeded_board, 'double')
    sm = make(boards.seeded_board, 'mixed')
    f = 4e9
    Sd, Sm = sd.sparams(f), sm.sparams(f)
    assert np.abs(Sm - Sd).max() <= 1e-4
    assert np.abs(Sd - Sd.T).max() < 1e-6
    assert np.linalg.svd(Sd, compute_uv=False).max() < 1 + 1e-6
    f2 = 4.2e9
    Sr = sm.sparams(f2)
    assert any('reused' in n for n in sm.log.notes)
    assert np.abs(Sr - sd.sparams(f2)).max() <= 1e-4


def test_rounded_double_factor_preconditions_as_the_double_one(gpu):
    """C2's low-frequency end: the complex64 factor does not converge at 1 GHz, so the point is solved with a
    complex128 factor, which is kept (by default rounded to complex64) and preconditions 1.15 GHz. Kept rounded and
    kept as complex128, the two runs give S within the mixed bound of the double-precision answer and within 2e-5 of
    each other, in GMRES iterations that differ by at most two a port."""
    from mom25d_amd import device as dv, fast
    sd = make(boards.seeded_board, 'double')
    runs = {}
    for keep in ('c64', 'c128'):
        fast.KEEP_DOUBLE = keep
        try:
            sm = make(boards.seeded_board, 'mixed')
            S = [sm.sparams(f) for f in (1.0e9, 1.15e9)]
            runs[keep] = (S, list(sm.fast.iters), [F.dtype for _, F, _, _ in sm.fast._kept])
        finally:
            fast.KEEP_DOUBLE = 'c64'
    assert runs['c64'][2] == [dv.c64] and runs['c128'][2] == [dv.c128]
    for f, a, b in zip((1.0e9, 1.15e9), runs['c64'][0], runs['c128'][0]):
        Sd = sd.sparams(f)
        assert np.abs(a - Sd).max() <= 1e-4 and np.abs(b - Sd).max() <= 1e-4, f
        assert np.abs(a - b).max() <= 2e-5, f
    assert max(abs(x - y) for x, y in zip(runs['c64'][1], runs['c128'][1])) <= 2


def _series_line(Z):
    """The reference's series-resistor board (tests/test_solver.py:339-350): a 0.35 mm strip, 12 mm long, on 0.2 mm
    of er 4.3, wave ports 0.5 mm in from both ends looking inwards (cuts from 1.5 mm every 0.7 mm to 4 mm), a series
    element Z at the middle; here the reference planes are at the element (5.5 mm in) instead of 4 mm in, so that
    circuit theory applies without a line section on either side."""
    from mom25d_amd.multilayer import Layer, LayerStack
    from mom25d_amd.sim import LinePort, SeriesLumped
    w, h, L = 0.35e-3, 0.2e-3, 12e-3
    st = LayerStack([Layer(h, 4.3)])
    p1 = LinePort(1, (-L / 2 + 0.5e-3, 0), (1, 0), w, 5.5e-3, s0=1.5e-3, step=0.7e-3, s1=4e-3, name='P1')
    p2 = LinePort(1, (L / 2 - 0.5e-3, 0), (-1, 0), w, 5.5e-3, s0=1.5e-3, step=0.7e-3, s1=4e-3, name='P2')
    return Simulation(st, {1: [(-L / 2, -w / 2, L / 2, w / 2)]}, [p1, p2],
                      lumped=[SeriesLumped(1, (0, 0), (1, 0), w, Z)], dmax=0.25e-3, w_ref=w, verbose=False)


def test_wave_ports_see_a_series_element_as_circuit_theory_does(gpu):
    """Wave ports on the reference's series-resistor line (its own test: |S21| = 2 Z0 / (2 Z0 + 50) within 0.01).
    The ports' S is in the feed's own impedance Z0 whatever zc both share, so without the reference's line
    characterisation the circuit identities that hold for any Z0 are checked, with the reference planes at the
    element: a series element gives S21 = 1 - S11 (50 ohms, within 0.01), the Z0 it implies, 50 (1 - S11) / (2 S11), is the 0.35 mm / 0.2 mm strip's (40-60 ohms,
    nearly real), and S21 then matches 2 Z0 / (2 Z0 + 50); a lossless element (j 50 ohms) on the lossless line
    conserves power (|S11|^2 + |S21|^2 within 0.01 of 1); S is symmetric; the fitted gamma's phase constant lies
    between the free-space and the substrate's wavenumbers and its residual is small."""
    f = 5e9
    sim = _series_line(50.0)
    S = sim.sparams(f)
    assert abs(S[1, 0] - (1 - S[0, 0])) < 0.01 and abs(S[0, 0] - S[1, 1]) < 0.01 and abs(S[0, 1] - S[1, 0]) < 1e-9
    z0 = 50.0 * (1 - S[0, 0]) / (2 * S[0, 0])
    assert 40 < z0.real < 60 and abs(z0.imag) < 0.1 * z0.real
    assert abs(abs(S[1, 0]) - 2 * z0.real / (2 * z0.real + 50)) < 0.01
    k0 = 2 * np.pi * f / 299792458.0
    g = sim.last['gamma']
    assert np.all((g.imag > k0) & (g.imag < k0 * np.sqrt(4.3))) and sim.last['resid'].max() < 1e-2
    Sx = _series_line(50j).sparams(f)
    assert abs(np.sum(np.abs(Sx[:, 0]) ** 2) - 1) < 0.01


def test_matrix_free_operator_and_lean_path(gpu):
    """matvec_exact (far field triangle-aggregated, near correction as a sparse matrix, via rows and columns) equals
    the dense FP64 Z times X to 1e-12 of each column's largest value on the coupled-line board (N ~ 3k, 4 via
    ports); the lean mixed path (complex64 Z only, the refinement through matvec_exact) gives the double-precision
    S within the mixed bound 1e-4, fresh and on a reused factor."""
    import torch
    sd = make(boards.coupled_lines, 'double')
    f = 5e9
    Z = sd.fast.fill(f)
    rng = np.random.default_rng(3)
    X = torch.from_numpy(rng.standard_normal((Z.shape[0], 5)) + 1j * rng.standard_normal((Z.shape[0], 5))).cuda()
    Y0 = (Z @ X).cpu().numpy()
    Y1 = sd.fast.matvec_exact(X).cpu().numpy()
    assert np.abs(Y1 - Y0).max() <= 1e-12 * np.abs(Y0).max(axis=0).max()
    del Z
    sl = make(boards.coupled_lines, 'mixed')
    sl.stream = 'on'
    for fr in (f, 5.3e9):
        Sd, Sl = sd.sparams(fr), sl.sparams(fr)
        assert np.abs(Sl - Sd).max() <= 1e-4, (fr, np.abs(Sl - Sd).max())
    assert any('reused' in n for n in sl.log.notes)


def test_metal_loss_matches_oracle_and_is_lossy(gpu):
    """The microstrip with copper loss (5.8e7 S/m, 35 um, 1 um roughness): S against the oracle (its Gram matrix by
    the midpoint rule, its wall loss its own) to 1e-5; the loss shows: |S21| below the lossless one, the ports'
    power sum below 1; reciprocal."""
    from mom25d_amd.sim import Metal
    metal = Metal(5.8e7, 35e-6, rough=1e-6)
    st, sh, po, vi, kw = boards.microstrip()
    lossy = Simulation(st, sh, po, vi, precision='double', verbose=False, metal=metal, **kw)
    osv = Solver(oracle_board(lossy))
    rows, signs = port_rows(lossy)
    zc = [p.zc for p in lossy.ports]
    plain = make(boards.microstrip, 'double')
    for f in (2e9, 5e9):
        So = osparams(osv.solve(f), osv.b.E, zc, rows, signs)
        Sl, Sp = lossy.sparams(f), plain.sparams(f)
        assert np.abs(Sl - So).max() <= 1e-5, (f, np.abs(Sl - So).max())
        assert abs(Sl[1, 0]) < abs(Sp[1, 0]) and np.sum(np.abs(Sl[:, 0]) ** 2) < 1 - 1e-4
        assert np.abs(Sl - Sl.T).max() < 1e-8


def test_solve_board_on_a_rectangle_board(gpu):
    """solve_board with a rectangle board in the build's JSON form (the microstrip): the adaptive sweep's S equals
    the simulation's own at the frequencies it solved and stays within its tolerance of a direct solve elsewhere; a
    kicad2ems-like file is refused."""
    from mom25d_amd.api import solve_board
    board = dict(layers=[[0.8e-3, 4.4, 0.02]], shapes={'1': [[-15e-3, -0.75e-3, 15e-3, 0.75e-3]]},
                 ports=[dict(via=[-15e-3, 0.0, 0, 1, 0.15e-3], name='P1'), dict(via=[15e-3, 0.0, 0, 1, 0.15e-3],
                                                                              name='P2')],
                 dmax=0.3e-3, ncross=4)
    freq = np.linspace(1e9, 6e9, 51)
    res = solve_board(board, freq, max_solves=16, afs_tol=2e-3, log=lambda *a: None)
    assert res.S.shape == (51, 2, 2) and res.names == ['P1', 'P2'] and res.solves <= 16
    sim = res.sim
    for f in res.solved[:3]:
        k = int(np.argmin(np.abs(freq - f)))
        assert np.abs(res.S[k] - sim.sparams(freq[k])).max() <= 1e-9
    k = 17
    assert np.abs(res.S[k] - sim.sparams(freq[k])).max() <= 1e-2
    with pytest.raises(ValueError, match='kicad2ems'):
        solve_board(dict(nets=[], layers=[]), freq, log=lambda *a: None)


def test_series_element_is_a_terminated_port(gpu):
    """A series resistor across a line (SeriesLumped: Z_el u u^T added to Z) equals the same line with its gap
    driven as a third port and that port terminated in Z_el by circuit algebra (the Schur complement of the
    3-port's impedance matrix), to 1e-9: the element's entries, the gap port's excitation and its current are one
    consistent u on the GPU path as in the oracle's (test_oracle_network)."""
    from mom25d_amd.multilayer import Layer, LayerStack
    from mom25d_amd.sim import GapPort, SeriesLumped, Via, ViaPort
    MM = 1e-3
    L, w = 10 * MM, 0.35 * MM
    st = LayerStack([Layer(0.2 * MM, 4.3, 0.02)])
    shapes = {1: [(-L / 2, -w / 2, L / 2, w / 2)]}
    vp = lambda: [ViaPort(Via(-L / 2, 0.0, 0, 1, 0.1 * MM), 50.0, 'P1'), ViaPort(Via(L / 2, 0.0, 0, 1, 0.1 * MM), 50.0, 'P2')]
    kw = dict(dmax=0.3 * MM, ncross=4, w_ref={1: w}, precision='double', verbose=False)
    Zel = 37.0
    with_el = Simulation(st, shapes, vp(), lumped=[SeriesLumped(1, (0.0, 0.0), (1.0, 0.0), w, Zel)], **kw)
    three = Simulation(st, shapes, vp() + [GapPort(1, (0.0, 0.0), (1.0, 0.0), w, 50.0, 'G')], **kw)
    I3 = np.eye(3)
    for f in (2e9, 7e9):
        S3 = three.sparams(f)
        Z3 = 50.0 * (I3 + S3) @ np.linalg.inv(I3 - S3)
        Z2 = Z3[:2, :2] - np.outer(Z3[:2, 2], Z3[2, :2]) / (Z3[2, 2] + Zel)
        S2 = (Z2 - 50.0 * np.eye(2)) @ np.linalg.inv(Z2 + 50.0 * np.eye(2))
        Se = with_el.sparams(f)
        assert np.abs(Se - S2).max() <= 1e-9, (f, np.abs(Se - S2).max())
