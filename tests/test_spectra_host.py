"""The stack's spectral functions of the C ABI (m25x_spectral_point: the host build of the network that the
m25x_spectra kernel evaluates on the contour) against the oracle's restatement of the reference's network
(oracle/greens.py, itself pinned by closed forms in test_oracle_greens and by circuit theory in
test_oracle_network), and against a closed form: a source over a ground in a uniform medium, whose TE voltage is
(1 - exp(-2j kz h)) / (2 Y0) - free space plus the ground's image.

CPU only: m25x_spectral_point runs on the host. The two formulations differ (reflection coefficients here, tangents
and secants in the reference), so they agree to rounding: 1e-11 of each value."""
import numpy as np
import pytest

from oracle import greens as og

MU0 = 4e-7 * np.pi


def stacks():
    rng = np.random.default_rng(2)
    four = [(rng.uniform(0.1, 0.4) * 1e-3, rng.uniform(3.5, 4.6), 0.02) for _ in range(4)]
    return {
        'microstrip': (dict(), [(0.8e-3, 4.4, 0.02)]),
        'four_layer': (dict(), four),
        'open_bottom': (dict(bottom='open', bottom_er=2.2), [(0.3e-3, 3.0, 0.01), (0.5e-3, 4.0, 0.02)]),
        'stripline': (dict(top='pec'), [(0.2e-3, 3.4, 0.01), (0.25e-3, 3.4, 0.01), (0.2e-3, 4.1, 0.02)]),
        'covered_air': (dict(top_er=2.0), [(0.4e-3, 1.0, 0.0), (0.1e-3, 9.8, 0.001)]),
    }


def product_stack(kw, lay):
    from mom25d_amd.multilayer import Layer, LayerStack
    return LayerStack([Layer(h, er, td) for h, er, td in lay], **kw)


def oracle_stack(kw, lay):
    return og.LayerStack([og.Layer(h, er, td) for h, er, td in lay], **kw)


@pytest.mark.parametrize('name', list(stacks()))
def test_spectral_point_matches_oracle(name):
    kw, lay = stacks()[name]
    ps, os_ = product_stack(kw, lay), oracle_stack(kw, lay)
    f = 3.7e9
    w = 2 * np.pi * f
    k0 = w / 299792458.0
    kmax = k0 * np.sqrt(os_.max_er())
    points = [0.3 * k0 + 0.2j * k0, 1.1 * kmax + 0.05j * k0, 0.9 * k0 + 0.01j * k0, 3 * kmax, 40 * kmax, 2e4 + 0j]
    ifs = os_.metal_interfaces
    for kr in points:
        for o in ifs:
            for s in ifs:
                ga, gq = ps.spectral_point(kr, w, o, s)
                ra, rq = os_.spectral(np.array([kr]), w, o, s)
                for mine, ref in ((ga, ra[0]), (gq, rq[0])):
                    assert abs(mine - ref) <= 1e-11 * abs(ref) + 1e-300, (name, kr, o, s, mine, ref)


def test_source_over_ground_is_free_space_plus_image():
    """Air over a ground, source and observation at height h: G~_A = mu0 (1 - exp(-2j kz h)) / (2j kz) exactly."""
    h = 0.7e-3
    ps = product_stack({}, [(h, 1.0, 0.0)])
    w = 2 * np.pi * 5e9
    k0 = w / 299792458.0
    for kr in (0.4 * k0 + 0.3j * k0, 2.5 * k0, 300.0 + 0j):
        kz = -1j * np.sqrt(kr * kr - k0 * k0 + 0j)
        expect = MU0 * (1 - np.exp(-2j * kz * h)) / (2j * kz)
        ga, _ = ps.spectral_point(kr, w, 1, 1)
        assert abs(ga - expect) <= 1e-12 * abs(expect)


def test_bad_stacks_are_refused():
    from mom25d_amd.multilayer import Layer, LayerStack
    with pytest.raises(ValueError):
        LayerStack([Layer(1e-3, 4.0)] * 17)
    st = LayerStack([Layer(1e-3, 4.0)])
    with pytest.raises(ValueError):
        st.spectral_point(100.0, 2 * np.pi * 1e9, 0, 5)


def test_requests_are_checked_by_type():
    """A one-layer table request that extracts two parts (its count is not an interface) is accepted; a request
    that names an interface beyond the stack is refused before any launch (no GPU needed: the check runs first)."""
    import ctypes as C
    from mom25d_amd import _lib as L
    from mom25d_amd.multilayer import Layer, LayerStack
    st = LayerStack([Layer(1e-3, 4.0)])
    w = 2 * np.pi * 1e9
    ri = np.array([[L.SPEC_TABLE, 0, 1, 1, 2, 0, 0, 0]], np.int32)
    bad = np.array([[L.SPEC_VIA_D, 0, 1, 0, 2, 0, 0, 0]], np.int32)
    z = C.c_void_p(8)                          # never dereferenced: the call stops at the check or at npts = 0
    L.call('m25x_spectra', C.byref(st.struct(w)), w, z, z, 0, z, z, 1, ri.ctypes.data_as(C.c_void_p), z, 2, None)
    with pytest.raises(ValueError, match='interface 2'):
        L.call('m25x_spectra', C.byref(st.struct(w)), w, z, z, 1, z, z, 1, bad.ctypes.data_as(C.c_void_p), z, 2,
               None)
