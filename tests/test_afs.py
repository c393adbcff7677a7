"""The adaptive sweep's rational fit and its stopping rule (CPU): AAA interpolates the reference's two-resonance
test function from 25 samples (test_solver.py:365-372's bound, 1e-6 on 301 points), reproduces a rational function
of known degree to rounding with one more support point than its degree and finds its poles; the sweep converges on
a smooth two-port with few solves and its model is within a few tol of the truth; a model that ignores new samples
never passes for converged (the reference's regression, test_solver.py:401-419: it runs to the cap, its history
entry per model, f_new None at the end)."""
import numpy as np

from mom25d_amd import afs


def two_resonances(f):
    g = f / 1e9
    return np.array([sum(1 / (1j * (g - p) + 0.01) for p in (2.4, 3.6))])


def test_aaa_interpolates_resonances():
    fs = np.linspace(1e9, 6e9, 25)
    r = afs.aaa(fs / 1e9, np.array([two_resonances(f) for f in fs]))
    fg = np.linspace(1e9, 6e9, 301)
    truth = np.array([two_resonances(f)[0] for f in fg])
    assert np.max(np.abs(r(fg / 1e9)[:, 0] - truth)) < 1e-6


def test_aaa_recovers_a_rational_function_and_its_poles():
    poles = np.array([0.3 + 0.05j, -0.2 + 0.1j, 0.7 - 0.02j])
    f = lambda z: np.stack([sum(1.0 / (z - p) for p in poles), 2.0 + sum((k + 1) / (z - p) for k, p in enumerate(poles))], 1)
    z = np.linspace(-1, 1, 40)
    r = afs.aaa(z, f(z))
    zz = np.linspace(-1, 1, 500) + 0.01j
    assert np.abs(r(zz) - f(zz)).max() <= 1e-9 * np.abs(f(zz)).max()
    assert len(r.zk) <= 5
    found = r.poles()
    for p in poles:
        assert np.min(np.abs(found - p)) < 1e-8


def test_sweep_converges_on_a_smooth_two_port():
    def S(f):
        g = f / 1e9
        s11 = 0.3 * np.exp(-1j * g) / (1 + 0.1j * (g - 3.0))
        s21 = 0.8 * np.exp(-2j * g)
        return np.array([[s11, s21], [s21, s11]])
    fg, Sg, fs, Ss, hist = afs.adaptive_sweep(S, 1e9, 6e9, tol=1e-4, ngrid=201, max_points=40, log=lambda *a: None)
    assert hist[-1]['converged'] and len(fs) < 25
    truth = np.array([S(f) for f in fg])
    assert np.abs(Sg - truth).max() < 1e-3
    assert all(np.array_equal(Ss[i], S(f)) for i, f in enumerate(fs))


def test_a_model_that_ignores_new_samples_never_converges(monkeypatch):
    real, first = afs.aaa, []

    def stale(z, F, **kw):
        first.append(first[0] if first else (z, F))
        return real(*first[-1], **kw)
    monkeypatch.setattr(afs, 'aaa', stale)
    truth = lambda f: np.array([[0.3 + 0.2 * np.sin(f / 4e8)]], complex)
    seen = []
    _, _, fs, _, hist = afs.adaptive_sweep(lambda f: seen.append(f) or truth(f), 1e9, 6e9, n0=5, tol=2e-3,
                                           max_points=12, ngrid=401, log=lambda *a: None)
    assert len(fs) == 12
    assert all(h['change'] < 2e-3 for h in hist[2:])
    assert [h['solves'] for h in hist] == list(range(6, 13)) and hist[-1]['f_new'] is None
    assert all(h['f_new'] in seen for h in hist[:-1])
