"""Pins of the oracle's Green's tables and via kernels (CPU) by closed forms and independent integrals:

  * an all-air stack: the kernels are free space plus one ground image, on and between interfaces (the reference's
    tests/test_green.py:27-39, 2e-4);
  * a grounded slab: the same kernels from the slab's closed-form spectral functions (transmission-line model
    written out, no recursion), transformed with quasi-static images taken out analytically and a finer contour
    rule (1e-4, the bound of the reference's tests/test_green.py:16-24);
  * a substrate split in two identical layers is the one layer (test_green.py:52-58);
  * vias in air: D == 0, and R + the K0 part against a nested adaptive quadrature of the free-space and image
    potentials (tests/test_solver.py:18-32, 1e-4); the tabulated R against the direct transform (35-42, 1e-4).
"""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import jv

from oracle import greens as og

C0, MU0, EPS0 = og.C0, og.MU0, og.EPS0
RHO = np.array([1e-5, 1e-4, 5e-4, 2e-3, 8e-3, 15e-3])


@pytest.mark.parametrize("o,s", [(1, 1), (1, 2), (3, 2), (1, 3)])
def test_air_stack_is_free_space_and_one_image(o, s):
    f = 10e9
    k0 = 2 * np.pi * f / C0
    st = og.LayerStack([og.Layer(0.1e-3, 1.0), og.Layer(0.15e-3, 1.0), og.Layer(0.2e-3, 1.0)])
    g = og.GreenTableML(f, st, o, s, 20e-3)
    zo, zs = st.z[o], st.z[s]
    R, Ri = np.sqrt(RHO ** 2 + (zo - zs) ** 2), np.sqrt(RHO ** 2 + (zo + zs) ** 2)
    img = np.exp(-1j * k0 * R) / R - np.exp(-1j * k0 * Ri) / Ri
    A, Q = g.total(RHO)
    assert np.max(abs(A - MU0 / (4 * np.pi) * img) / abs(MU0 / (4 * np.pi) * img)) < 2e-4
    assert np.max(abs(Q - img / (4 * np.pi * EPS0)) / abs(img / (4 * np.pi * EPS0))) < 2e-4


def _slab_kernels(f, h, er, rho, nq=12):
    """G_A, G_q on the interface of a grounded slab (source and observation there), written out from the slab's
    transmission-line model: TE  G~_A = mu0 / (j kz0 + kz1 cot(kz1 h)),
    TM  V_e = j kz0 kz1 / (w eps0 (er kz0 cot(kz1 h) + j kz1)),  G~_q = j w (V_e - j w G~_A) / kr^2,
    Im kz <= 0. The quasi-static parts (kr -> inf) are taken out in the spectral domain and added back as their
    images: mu0/(4 pi) (1/rho - 1/sqrt(rho^2 + 4h^2)) and the charge image series of K = (er - 1)/(er + 1)."""
    w = 2 * np.pi * f
    k0 = w / C0
    kd = k0 * np.sqrt(er)
    t1, alpha = 1.5 * kd.real, min(0.5 * k0, 4.0 / rho.max())
    wid1, wid2 = min(0.3 * k0, 0.8 * alpha), 2 * np.pi / rho.max()
    K = 1e5 + 40 * k0 * np.sqrt(er.real)
    br = np.concatenate([np.linspace(0, t1, int(np.ceil(t1 / wid1)) + 1), np.append(np.arange(t1, K, wid2), K)[1:]])
    xg, wg = np.polynomial.legendre.leggauss(nq)
    a, b = br[:-1, None], br[1:, None]
    t = (0.5 * (b - a) * xg + 0.5 * (a + b)).ravel()
    wt = (0.5 * (b - a) * wg).ravel()
    lifted = t < t1
    kr = t + 1j * np.where(lifted, alpha * np.sin(np.pi * t / t1), 0)
    dk = wt * (1 + 1j * np.where(lifted, alpha * np.pi / t1 * np.cos(np.pi * t / t1), 0))
    kz0, kz1 = -1j * np.sqrt(kr ** 2 - k0 ** 2), -1j * np.sqrt(kr ** 2 - kd ** 2)
    cot = 1 / np.tan(kz1 * h)
    GA = MU0 / (1j * kz0 + kz1 * cot)
    Ve = 1j * kz0 * kz1 / (w * EPS0 * (er * kz0 * cot + 1j * kz1))
    Gq = 1j * w * (Ve - 1j * w * GA) / kr ** 2
    GA_qs = MU0 * (1 - np.exp(-2 * kr * h)) / (2 * kr)
    Gq_qs = 1 / (EPS0 * kr * (1 + er / np.tanh(kr * h)))
    J = jv(0, rho[:, None] * kr[None, :])
    dA = J @ ((GA - GA_qs) * kr * dk) / (2 * np.pi)
    dQ = J @ ((Gq - Gq_qs) * kr * dk) / (2 * np.pi)
    Kc = (er - 1) / (er + 1)
    n = np.arange(1, 401)
    img = ((-Kc) ** (n - 1))[None, :] / np.sqrt(rho[:, None] ** 2 + (2 * n * h)[None, :] ** 2)
    cq = 1 / (2 * np.pi * EPS0 * (1 + er))
    A = MU0 / (4 * np.pi) * (1 / rho - 1 / np.sqrt(rho ** 2 + 4 * h * h)) + dA
    Q = cq * (1 / rho - (1 + Kc) * img.sum(1)) + dQ
    return A, Q


def test_grounded_slab_matches_its_closed_form_spectral_functions():
    f, h, er = 5e9, 0.2e-3, 4.3 * (1 - 0.02j)
    A1, Q1 = _slab_kernels(f, h, er, RHO)
    g = og.GreenTableML(f, og.LayerStack([og.Layer(h, 4.3, 0.02)]), 1, 1, 20e-3)
    A2, Q2 = g.total(RHO)
    assert np.max(abs(A2 - A1) / abs(A1)) < 1e-4
    assert np.max(abs(Q2 - Q1) / abs(Q1)) < 1e-4
    # with the image extracted (close ground: image_sing) the totals are the same kernels
    gi = og.GreenTableML(f, og.LayerStack([og.Layer(h, 4.3, 0.02)]), 1, 1, 20e-3, image_sing=True)
    assert gi.dimg == pytest.approx(2 * h)
    A3, Q3 = gi.total(RHO)
    assert np.max(abs(A3 - A1) / abs(A1)) < 1e-4 and np.max(abs(Q3 - Q1) / abs(Q1)) < 1e-4


def test_split_substrate_is_one_layer():
    f, h = 5e9, 0.2e-3
    one = og.GreenTableML(f, og.LayerStack([og.Layer(h, 4.3, 0.02)]), 1, 1, 20e-3)
    two = og.GreenTableML(f, og.LayerStack([og.Layer(h / 2, 4.3, 0.02), og.Layer(h / 2, 4.3, 0.02)]), 2, 2, 20e-3)
    A1, Q1 = one.total(RHO)
    A2, Q2 = two.total(RHO)
    assert np.max(abs(A2 - A1) / abs(A1)) < 2e-5 and np.max(abs(Q2 - Q1) / abs(Q1)) < 2e-5


def test_via_kernels_in_air_are_images():
    h, a, f = 0.5e-3, 0.15e-3, 10e9
    w = 2 * np.pi * f
    k = w / C0
    vk = og.ViaKernels(f, og.LayerStack([og.Layer(h, 1.0)]), [(0, 1)], [1], 20e-3)
    assert np.max(np.abs(vk.D[(1, (0, 1))](np.array([1e-4, 1e-3])))) < 1e-10
    for d in (a, 1e-3):
        tot = vk.R((0, 1), (0, 1), d)[0] + vk.homog((0, 1), (0, 1), d)[0]
        Rp = np.hypot(d, 2 * h)
        phi = (np.exp(-1j * k * d) / d - np.exp(-1j * k * Rp) / Rp) / (1j * w * 4 * np.pi * EPS0)
        gfn = lambda z, zp: np.exp(-1j * k * np.hypot(d, z - zp)) / np.hypot(d, z - zp)
        I = sum((1j ** p) * quad(lambda z: quad(lambda zp: (gfn(z, zp).real, gfn(z, zp).imag)[p], -h, h,
                                                points=[z], limit=200)[0], 0, h, limit=200)[0] for p in (0, 1))
        ref = 1j * w * MU0 / (4 * np.pi) * I + phi
        assert abs(tot + phi - ref) / abs(ref) < 1e-4


def test_via_table_is_the_direct_transform():
    st = og.LayerStack([og.Layer(0.2e-3, 4.3), og.Layer(1.0e-3, 4.3), og.Layer(0.2e-3, 4.3)])
    vk = og.ViaKernels(5e9, st, [(0, 3), (1, 3)], [1, 2, 3], 40e-3)
    d = np.concatenate([[0.15e-3], np.geomspace(0.3e-3, 38e-3, 60)])
    for sm, sn in (((0, 3), (0, 3)), ((0, 3), (1, 3))):
        exact, tab = vk.R(sm, sn, d), vk.R_table(sm, sn)(d)
        assert np.abs(tab - exact).max() < 1e-4 * np.abs(exact).max()
