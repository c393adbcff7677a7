"""The mesher, RWG basis and geometry dicts (host, CPU): the product's basis numbers the edges as the reference's
(first seen, T+ the first triangle: src/mom25d/base.py:16-47), so Z compares entry by entry with the oracle's;
boards of the configurations mesh with their vias on grid nodes."""
import numpy as np
import pytest

from mom25d_amd import boards
from mom25d_amd.base import RWG, geometry
from mom25d_amd.sim import Simulation
from oracle import geometry as ogeo


@pytest.mark.parametrize('make', [boards.microstrip, boards.seeded_board, boards.coupled_lines])
def test_board_meshes_match_the_oracles_basis(make):
    st, sh, po, vi, kw = make()
    sim = Simulation(st, sh, po, vi, verbose=False, mesh_only=True, **kw)
    r = sim.rwg
    o = ogeo.RWG(r.nodes, r.tris)
    for k in ('tp', 'tm', 'vp', 'vm', 'l', 'nrm', 'area', 'cent'):
        assert np.array_equal(getattr(r, k), getattr(o, k)), k
    g, q, size = geometry(r, sim.tri_layer, sim.layers)
    go, qo, so = ogeo.geometry_dicts(o, sim.tri_layer, sim.layers)
    for k in go:
        assert np.array_equal(g[k], go[k]), k
    for k in qo:
        assert np.array_equal(q[k], qo[k]), k
    assert np.array_equal(size, so)


def test_config_sizes():
    """The benchmark boards' unknowns (recorded with the results): C1 ~ 1.1k, C2 ~ 7.8k, C3 ~ 3.5k."""
    for make, lo, hi in ((boards.microstrip, 900, 1300), (boards.seeded_board, 7000, 9000),
                         (boards.coupled_lines, 3000, 4000)):
        st, sh, po, vi, kw = make()
        sim = Simulation(st, sh, po, vi, verbose=False, mesh_only=True, **kw)
        assert lo <= sim.Ne + sim.Nv <= hi


def test_via_off_metal_is_rejected():
    from mom25d_amd.sim import Via, ViaPort
    st, sh, po, vi, kw = boards.microstrip()
    with pytest.raises(ValueError):
        Simulation(st, sh, [ViaPort(Via(0.0, 5e-3, 0, 1))], verbose=False, mesh_only=True, **kw)


def test_rwg_divergence_sums_to_zero_per_edge():
    """D^T applied to the areas: +l on T+, -l on T- (each basis function carries no net charge)."""
    st, sh, po, vi, kw = boards.microstrip()
    sim = Simulation(st, sh, po, vi, verbose=False, mesh_only=True, **kw)
    r = sim.rwg
    assert np.allclose(r.div_apply(r.area), 0.0, atol=1e-15)
