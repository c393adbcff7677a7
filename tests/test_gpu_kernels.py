"""The fill kernels on the GPU against the oracle (the reference's algorithm restated in C, itself pinned to the
reference's compiled kernels in test_oracle_pins.py), on the same seeded inputs.

Parity metric on Z (FP64):
  far fill   bit-identical to the oracle (both round as the reference's CPU build: no fused multiply-adds, the
             same operation order), hence to the reference's compiled fill_far;
  near field per-entry relative 1e-10 wherever |Z_ij| >= 1e-4 max_k |Z_ik|, and |dZ_ij| <= 1e-13 max_k |Z_ik|
             everywhere. Below that floor an entry is the difference of terms ~1e4 times larger (far and near
             parts of distant near pairs cancel), so a one-ulp difference between the GPU's and glibc's asinh /
             atan2 in the analytic moments shows there at 1e-7 relative; the reference's own near field (atomic
             adds in no fixed order) is not reproducible bit for bit either.
"""
import numpy as np
import pytest

from oracle import core as oc
from tests.cases import two_layer

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel_entry(a, b, floor=1e-12):
    """max per-entry |a - b| / max(|b|, floor max|b_row|)."""
    fl = floor * np.abs(b).max(axis=-1, keepdims=True)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), fl)))


def rel_row(a, b):
    """max |a - b| / max|b_row|."""
    return float(np.max(np.abs(a - b) / np.abs(b).max(axis=-1, keepdims=True)))


@pytest.fixture(scope='module')
def case():
    return two_layer()


def test_fill_far_matches_oracle(gpu, case):
    from mom25d_amd import core
    g, tb, w = case['g'], case['tb'], case['w']
    Ne = len(g['ep'])
    N = Ne + 3                                      # an edge block inside a larger Z (via unknowns after it)
    Zo = np.zeros((N, N), complex)
    oc.fill_far(Zo, g, tb, w)
    Zg = np.full((N, N), 7 + 7j)                    # the kernel writes the edge block only
    core.fill_far(Zg, g, tb, w)
    assert np.array_equal(Zg[:Ne, :Ne], Zo[:Ne, :Ne])                     # bit for bit
    assert np.all(Zg[Ne:, :] == 7 + 7j) and np.all(Zg[:, Ne:] == 7 + 7j)
    assert np.array_equal(Zg[:Ne, :Ne], Zg[:Ne, :Ne].T)                   # symmetric bit for bit
    from oracle import ref
    if ref.core() is not None:                                            # the reference's own compiled fill
        Zr = np.zeros((N, N), complex)
        ref.core().fill_far(Zr, g, tb, w)
        assert np.array_equal(Zg[:Ne, :Ne], Zr[:Ne, :Ne])
    Z64 = np.zeros((Ne, Ne), np.complex64)
    core.fill_far_c64(Z64, g, tb, w)
    assert np.abs(Z64 - Zo[:Ne, :Ne]).max() < 1e-6 * np.abs(Zo).max()


def test_far_block_matches_oracle(gpu, case):
    from mom25d_amd import core
    g, tb, w = case['g'], case['tb'], case['w']
    Ne = len(g['ep'])
    rows, cols = np.arange(0, Ne, 7), np.arange(3, Ne, 5)
    assert np.array_equal(core.far_block(rows, cols, g, tb, w), oc.far_block(rows, cols, g, tb, w))


def test_near_field_matches_oracle(gpu, case):
    from mom25d_amd import core
    g, q, tb, w, st = case['g'], case['q'], case['tb'], case['w'], case['stack']
    P, Q = oc.near_pairs(case['rwg'].cent, case['size'], st.z[case['tl']], 5.0)
    P32, Q32 = P.astype(np.int32), Q.astype(np.int32)
    Ne = len(g['ep'])
    Zo = np.zeros((Ne, Ne), complex)
    oc.fill_far(Zo, g, tb, w)
    Zg = Zo.copy()
    qo = dict(q, stat=np.zeros((len(P), 12)), stat_ready=False)
    qg = dict(q, stat=np.zeros((len(P), 12)), stat_ready=False)
    co = oc.near_into_dense(Zo, g, tb, qo, P32, Q32, w)
    cg = core.near_into_dense(Zg, g, tb, qg, P32, Q32, w)
    assert rel_row(Zg, Zo) < 1e-13 and rel_entry(Zg, Zo, floor=1e-4) < TOL
    assert np.abs(cg - co).max() <= 1e-13 * np.abs(co).max()
    assert np.abs(qg['stat'] - qo['stat']).max() <= 1e-12 * np.abs(qo['stat']).max()
    # a second frequency from the kept analytic parts equals a fresh computation
    qg['stat_ready'] = True
    w2 = 1.1 * w
    Z1, Z2 = np.zeros((Ne, Ne), complex), np.zeros((Ne, Ne), complex)
    core.near_into_dense(Z1, g, tb, qg, P32, Q32, w2)
    core.near_into_dense(Z2, g, tb, dict(q), P32, Q32, w2)
    assert np.abs(Z1 - Z2).max() <= 1e-13 * np.abs(Z2).max()
    # COO triplets: the same values per slot as the oracle, -1 rows alike
    ro, cio, vo, _ = oc.near_coo(g, tb, q, P32, Q32, w)
    rg, cig, vg, _ = core.near_coo(g, tb, q, P32, Q32, w)
    assert np.array_equal(ro, rg)
    used = ro >= 0
    assert np.array_equal(cio[used], cig[used])
    assert np.abs(vg[used] - vo[used]).max() <= 1e-13 * np.abs(vo).max()
    assert np.abs(core.near_q(g, tb, q, P32, Q32, w) - co).max() <= 1e-12 * np.abs(co).max()


def test_matrix_free_products_match_oracle(gpu, case):
    """matvec_far / matvec_far_multi / near_matvec through the drop-in module against the oracle's (which equal the
    reference's compiled kernels bit for bit, test_oracle_pins): relative 1e-12 of each column's largest value
    (the triangle-pair sums run in slices on the GPU, in one sequence on the CPU), for k = 1, 3 and 11 columns
    (two calls of at most 8); and the far product equal to the dense far fill times X."""
    from mom25d_amd import core
    g, q, tb, w, st = case['g'], case['q'], case['tb'], case['w'], case['stack']
    Ne = len(g['ep'])
    rng = np.random.default_rng(11)
    for k in (1, 3, 11):
        X = rng.standard_normal((Ne, k)) + 1j * rng.standard_normal((Ne, k))
        Yg, Yo = core.matvec_far_multi(X, g, tb, w), oc.matvec_far_multi(X, g, tb, w)
        assert np.abs(Yg - Yo).max() <= 1e-12 * np.abs(Yo).max(), k
    x = X[:, 4]
    yg, yo = core.matvec_far(x, g, tb, w), oc.matvec_far(x, g, tb, w)
    assert np.abs(yg - yo).max() <= 1e-12 * np.abs(yo).max()
    Z = np.zeros((Ne, Ne), complex)
    core.fill_far(Z, g, tb, w)
    assert np.abs(yg - Z @ x).max() <= 1e-12 * np.abs(Z @ x).max()
    P, Q = oc.near_pairs(case['rwg'].cent, case['size'], st.z[case['tl']], 5.0)
    P32, Q32 = P.astype(np.int32), Q.astype(np.int32)
    yn, yno = core.near_matvec(x, g, tb, q, P32, Q32, w), oc.near_matvec(x, g, tb, q, P32, Q32, w)
    assert np.abs(yn - yno).max() <= 1e-12 * np.abs(yno).max()


def test_near_pairs_match_oracle(gpu, case):
    from mom25d_amd import core
    st = case['stack']
    args = (case['rwg'].cent, case['size'], st.z[case['tl']], 5.0)
    Po, Qo = oc.near_pairs(*args)
    Pg, Qg = core.near_pairs(*args)
    assert np.array_equal(Po, Pg) and np.array_equal(Qo, Qg)


def test_fan_q_and_radial_sum(gpu, case):
    from mom25d_amd import core
    g, tb, q = case['g'], case['tb'], case['q']
    F = np.arange(0, len(g['lt']), 11).astype(np.int32)
    a, b = core.fan_q(g, tb, F), oc.fan_q(g, tb, F)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    rng = np.random.default_rng(5)
    table = rng.standard_normal(200) + 1j * rng.standard_normal(200)
    P6, W6 = q['P6'], q['W6']
    vx, vy = rng.uniform(-4e-3, 4e-3, 5), rng.uniform(-1e-3, 1e-3, 5)
    args = (P6[..., 0], P6[..., 1], W6, vx, vy, table, case['rho_max'] / 199)
    a, b = core.radial_table_sum(*args), oc.radial_table_sum(*args)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_bessel_and_hankel(gpu):
    from mom25d_amd import core
    rng = np.random.default_rng(7)
    z = np.concatenate([rng.uniform(0, 30, 300) + 1j * rng.uniform(-2, 2, 300), [0j, 14 + 0j, 14.5 + 0.1j]])
    for n in (0, 1, 2):
        # the series' largest term is about 3e4 at |z| = 14: both evaluations carry ~5e-12 of rounding
        # (the reference's own bound, mom25d_core.cpp:1506-1508)
        a, b = core.bessel_jn_complex(n, z), oc.bessel_jn_complex(n, z)
        assert np.abs(a - b).max() <= 2e-11 * max(1.0, np.abs(b).max())
    t = np.linspace(0, 2e4, 3000)
    rho = np.linspace(0, 5e-3, 41)
    W = rng.normal(size=(2, len(t))) * (1 + 0.3j)
    a, b = core.hankel_real(rho, t, W), oc.hankel_real(rho, t, W)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    # a contour lifted at the start: real nodes by J0, lifted ones by the complex series, summed apart
    from mom25d_amd import device as dv
    kr = t + 1j * np.where(t < 500, 50 * np.sin(np.pi * t / 500), 0.0)
    real = np.abs(kr.imag) < 1e-300
    ref = oc.hankel_real(rho, kr.real[real], W[:, real]) + (oc.bessel_jn_complex(0, rho[:, None] * kr[None, ~real])
                                                            @ W[:, ~real].T).T
    got = dv.hankel_sum(rho, kr, W).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_scale_c64(gpu):
    from mom25d_amd import core
    rng = np.random.default_rng(1)
    Z = rng.normal(size=(50, 50)) + 1j * rng.normal(size=(50, 50))
    d = rng.uniform(0.5, 2, 50)
    Zr = (Z * d[:, None]) * d[None, :]
    Z64 = core.scale_c64(Z, d)
    assert np.abs(Z - Zr).max() == 0 and Z64.dtype == np.complex64
    assert np.allclose(Z64, Zr, rtol=1e-6)


def test_kernels_reject_bad_arguments(gpu):
    from mom25d_amd import core
    with pytest.raises(ValueError):
        core.bessel_jn_complex(3, np.zeros(3, complex))
    with pytest.raises(ValueError):
        core.radial_table_sum(np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)), np.zeros(1), np.zeros(1),
                              np.zeros(1, complex), 1.0)
    with pytest.raises(ValueError):
        core.scale_c64(np.zeros((3, 4), complex), np.ones(3))
