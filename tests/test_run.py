"""solve_band's host logic (src/mom25d/run.py): renormalisation against circuit theory, and the sweep plumbing
with a stand-in Simulation (the GPU solve itself is covered by test_gpu_solve)."""
import numpy as np
import pytest

from mom25d_amd.run import renorm, out_refs, solve_band


def s_of_z(Z, z0):
    I = np.eye(len(Z))
    zn = Z / z0
    return (zn - I) @ np.linalg.inv(zn + I)


def test_renorm_matches_circuit_theory():
    rng = np.random.default_rng(0)
    Z = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    Z = Z + Z.T + 60 * np.eye(3)
    S50, S75 = s_of_z(Z, 50.0), s_of_z(Z, 75.0)
    assert np.allclose(renorm(S50, 50.0, 75.0), S75, atol=1e-12)
    assert np.allclose(renorm(S75, [75, 75, 75], 50.0), S50, atol=1e-12)
    assert renorm(S50, 50.0, 50.0) is S50


class _Port:
    def __init__(self, name, zc):
        self.name, self.zc = name, zc


class _Sim:
    def __init__(self):
        self.ports = [_Port('a', 50.0), _Port('b', 50.0)]
        self.calls = []

    def sparams(self, f, keep=False):
        self.calls.append(f)
        self.last = dict(I=np.full(3, f))
        x = 0.1 * f / 1e9
        return np.array([[x, 1 - x], [1 - x, x]], complex)


def test_solve_band_linear_and_keep():
    sim = _Sim()
    f, S, info = solve_band(sim, 1e9, 3e9, linear=3, zref={'a': 50, 'b': 50}, log=lambda *_: None, keep_at=[2e9])
    assert np.allclose(f, [1e9, 2e9, 3e9]) and S.shape == (3, 2, 2)
    assert info['solves'] == 3 and info['names'] == ['a', 'b'] and info['zref'] == [50.0, 50.0]
    assert list(sim.kept) == [2e9]
    assert np.allclose(S[1], sim.sparams(2e9))
    assert list(out_refs(sim, 25.0)) == [25.0, 25.0]


class _Log:
    def __init__(self):
        self.notes = []

    def note(self, msg):
        self.notes.append(msg)


def test_solve_band_adaptive():
    """The default sweep (adaptive): the stand-in's S is linear in f, so the model is exact after a few solves;
    it converges well before the cap, its model equals the truth on the requested grid, keep_at is among the
    solved frequencies, and the history reaches on_step."""
    sim = _Sim()
    sim.log = _Log()
    steps = []
    f, S, info = solve_band(sim, 1e9, 3e9, points=41, log=lambda *_: None, keep_at=[1.7e9], on_step=steps.append,
                            max_solves=12)
    assert len(f) == 41 and S.shape == (41, 2, 2)
    truth = np.array([_Sim().sparams(fi) for fi in f])
    assert np.abs(S - truth).max() < 1e-9
    assert info['solves'] < 12 and info['afs'][-1]['converged'] and info['afs_tol'] == 2e-3
    assert 1.7e9 in info['solved'] and list(sim.kept) == [1.7e9]
    assert steps and steps[-1] is info['afs'][-1] and not sim.log.notes
