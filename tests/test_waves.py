"""The wave-port fit (src/mom25d/base.py:161-172 and the wave-port branch of src/mom25d/sim.py:1045-1092): the
propagation constant and the two waves recovered from cut currents of an exact two-wave line, the sample cuts of a
port, and the fit's tolerance of noise. Parity: the fit is pinned by the closed form it inverts (a sequence
a e^(-g s) + b e^(g s) is recovered exactly), not by the reference's values."""
import numpy as np

from mom25d_amd.sim import LinePort, wave_gamma, wave_split


def two_waves(g, s, a, b):
    return np.exp(-g * s)[:, None] * a[None, :] + np.exp(g * s)[:, None] * b[None, :]


def test_gamma_and_waves_are_recovered_exactly():
    g = 35.0 + 240.0j                                  # 1/m: a lossy quasi-TEM line near 10 GHz
    s = 1.5e-3 + 0.7e-3 * np.arange(6) - 4e-3          # cuts, from a reference plane 4 mm out
    a = np.array([1.0 + 0.2j, -0.3 + 0.1j, 0.02j])
    b = np.array([0.4 - 0.1j, 1.1 + 0.0j, 0.0])        # the last column: a pure forward wave
    Ik = two_waves(g, s, a, b)
    gf = wave_gamma(Ik, 0.7e-3)
    assert abs(gf - g) < 1e-9 * abs(g)
    af, bf = wave_split(Ik, s, gf)
    assert np.abs(af - a).max() < 1e-9 and np.abs(bf - b).max() < 1e-9
    assert abs(wave_gamma(Ik[:, 2:], 0.7e-3) - g) < 1e-9 * abs(g)   # one column alone


def test_gamma_is_taken_with_positive_parts_and_tolerates_noise():
    g = 10.0 + 180.0j
    s = 0.5e-3 * np.arange(8)
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal(2) + 1j * rng.standard_normal(2), rng.standard_normal(2) + 1j * rng.standard_normal(2)
    Ik = two_waves(g, s, a, b)
    gf = wave_gamma(Ik, 0.5e-3)
    assert gf.real >= 0 and gf.imag >= 0 and abs(gf - g) < 1e-9 * abs(g)
    noisy = Ik * (1 + 1e-6 * rng.standard_normal(Ik.shape))
    assert abs(wave_gamma(noisy, 0.5e-3) - g) < 1e-3 * abs(g)


def test_sample_cuts():
    p = LinePort(1, (0, 0), (1, 0), 0.35e-3, 4e-3, s0=1.5e-3, step=0.7e-3)
    s = p.cuts(0.25e-3)
    assert np.allclose(s, [1.9e-3, 2.6e-3, 3.3e-3, 4.0e-3])
    q = LinePort(1, (0, 0), (0, -1), 0.35e-3, 3e-3)      # defaults: step = dmax, s0 = 2 step, s1 = the reference
    s = q.cuts(0.5e-3)
    assert np.isclose(s[-1], 3e-3) and s[0] >= 1e-3 - 1e-12 and np.allclose(np.diff(s), 0.5e-3)
