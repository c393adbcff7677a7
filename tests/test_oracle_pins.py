"""Pins of the oracle (CPU, no GPU): the oracle's C kernels against the reference's own compiled kernels
(oracle/_ref, built from the reference sources by oracle/build_ref.sh) on the same inputs, and against what the
mathematics fixes where no reference run is needed (symmetries, limits, subdivision convergence).

A plausible slip in the restatement (a dropped term, a swapped index, a wrong sign) changes the bits of the fill
and fails the bitwise comparisons; the invariant checks catch what the reference comparison cannot see when
oracle/_ref is absent.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import core as oc, geometry as ogeo, ref
from tests.cases import two_layer


@pytest.fixture(scope='module')
def case():
    return two_layer()


def refcore():
    m = ref.core()
    if m is None:
        pytest.skip("oracle/_ref not built (no reference tree where the repo was built)")
    return m


def test_fill_far_is_the_references_bit_for_bit(case):
    R = refcore()
    g, tb, w = case['g'], case['tb'], case['w']
    Ne = len(g['ep'])
    for dt, fill_o, fill_r in ((complex, oc.fill_far, R.fill_far), (np.complex64, oc.fill_far_c64, R.fill_far_c64)):
        Zo, Zr = np.zeros((Ne + 2, Ne + 2), dt), np.zeros((Ne + 2, Ne + 2), dt)
        fill_o(Zo, g, tb, w)
        fill_r(Zr, g, tb, w)
        assert np.array_equal(Zo, Zr)
    rows, cols = np.arange(0, Ne, 3).astype(np.int32), np.arange(1, Ne, 4).astype(np.int32)
    assert np.array_equal(oc.far_block(rows, cols, g, tb, w), R.far_block(rows, cols, g, tb, w))


def test_near_field_is_the_references(case):
    R = refcore()
    g, q, tb, w, st = case['g'], case['q'], case['tb'], case['w'], case['stack']
    args = (case['rwg'].cent, case['size'], st.z[case['tl']], 5.0)
    P, Q = oc.near_pairs(*args)
    Pr, Qr = R.near_pairs(*args)
    assert np.array_equal(P, Pr) and np.array_equal(Q, Qr)
    P32, Q32 = P.astype(np.int32), Q.astype(np.int32)
    Ne = len(g['ep'])
    Zo, Zr = np.zeros((Ne, Ne), complex), np.zeros((Ne, Ne), complex)
    qo = dict(q, stat=np.zeros((len(P), 12)), stat_ready=False)
    qr = dict(q, stat=np.zeros((len(P), 12)), stat_ready=False)
    co, cr = oc.near_into_dense(Zo, g, tb, qo, P32, Q32, w), R.near_into_dense(Zr, g, tb, qr, P32, Q32, w)
    assert np.array_equal(co, cr) and np.array_equal(qo['stat'], qr['stat'])
    # the adds are atomic in no fixed order on both sides: equal up to that order
    assert np.abs(Zo - Zr).max() <= 1e-14 * np.abs(Zr).max()
    a, b = oc.near_coo(g, tb, q, P32, Q32, w), R.near_coo(g, tb, q, P32, Q32, w)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[3], b[3])
    used = a[0] >= 0                       # the reference leaves an unused slot's column and value unwritten
    assert np.array_equal(a[1][used], b[1][used]) and np.array_equal(a[2][used], b[2][used])
    assert np.array_equal(oc.near_q(g, tb, q, P32, Q32, w), R.near_q(g, tb, q, P32, Q32, w))
    F = np.arange(0, len(g['lt']), 9).astype(np.int32)
    assert np.array_equal(oc.fan_q(g, tb, F), R.fan_q(g, tb, F))


def test_numerics_are_the_references():
    R = refcore()
    rng = np.random.default_rng(0)
    t = np.linspace(0, 2e4, 500)
    rho = np.linspace(0, 5e-3, 21)
    W = rng.normal(size=(3, len(t))) * (1 + 0.3j)
    assert np.array_equal(oc.hankel_real(rho, t, W), R.hankel_real(rho, t, W))
    z = rng.uniform(0, 30, 200) + 1j * rng.uniform(-3, 3, 200)
    for n in (0, 1, 2):
        assert np.array_equal(oc.bessel_jn_complex(n, z), R.bessel_jn_complex(n, z))
    px, py, Wq = rng.uniform(0, 1e-2, (40, 6)), rng.uniform(0, 1e-2, (40, 6)), rng.uniform(0, 1, (40, 6))
    vx, vy = rng.uniform(0, 1e-2, 4), rng.uniform(0, 1e-2, 4)
    tab = rng.normal(size=150) + 1j * rng.normal(size=150)
    assert np.array_equal(oc.radial_table_sum(px, py, Wq, vx, vy, tab, 1e-4),
                          R.radial_table_sum(px, py, Wq, vx, vy, tab, 1e-4))
    Z = rng.normal(size=(40, 40)) + 1j * rng.normal(size=(40, 40))
    d = rng.uniform(0.5, 2, 40)
    Z1, Z2 = Z.copy(), Z.copy()
    assert np.array_equal(oc.scale_c64(Z1, d), R.scale_c64(Z2, d)) and np.array_equal(Z1, Z2)


# ---- what the mathematics fixes (no reference run) ----------------------------------------------------------
def _offplane_one(V, P6, W6, d):
    lib = oc.lib()
    m0, mrr = C.c_double(), C.c_double()
    mro, mrs = (C.c_double * 2)(), (C.c_double * 2)()
    V, P6, W6 = (np.ascontiguousarray(a, float) for a in (V, P6, W6))
    lib.orc_offplane_one(V.ctypes.data_as(C.c_void_p), P6.ctypes.data_as(C.c_void_p),
                         W6.ctypes.data_as(C.c_void_p), C.c_double(d), C.byref(m0), mro, mrs, C.byref(mrr))
    return m0.value, np.array(mro[:]), np.array(mrs[:]), mrr.value


def _split4(T):
    """4 sub-triangles per triangle (midpoint subdivision), (n, 3, 2) -> (4n, 3, 2)."""
    a, b, c = T[:, 0], T[:, 1], T[:, 2]
    ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
    return np.concatenate([np.stack(s, 1) for s in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))])


def _outer_subdivided(Vp, Vq, d, levels):
    """The analytic inner integral over Vq, the 6-point rule over 4^levels sub-triangles of Vp."""
    T = Vp[None]
    for _ in range(levels):
        T = _split4(T)
    tot = np.zeros(6)
    for t in T:
        A = 0.5 * abs((t[1, 0] - t[0, 0]) * (t[2, 1] - t[0, 1]) - (t[1, 1] - t[0, 1]) * (t[2, 0] - t[0, 0]))
        P6 = ogeo._BAR6 @ t
        m0, mro, mrs, mrr = _offplane_one(Vq, P6, A * ogeo._W6, d)
        tot += np.r_[m0, mro, mrs, mrr]
    return tot


def test_offplane_moments_converge_and_reduce_to_the_far_limit():
    tri = lambda c, s, rot: np.c_[c[0] + s * np.cos(rot + np.array([0, 2.1, 4.2])),
                                  c[1] + s * np.sin(rot + np.array([0, 2.1, 4.2]))]
    Vp, Vq = tri((0.3, 0.03), 0.18, 1.1) * 1e-3, tri((0, 0), 0.54, 0.3) * 1e-3
    for d in (0.0, 0.2e-3):
        coarse, fine = _outer_subdivided(Vp, Vq, d, 0), _outer_subdivided(Vp, Vq, d, 4)
        assert np.abs(coarse - fine).max() < 2e-2 * np.abs(fine).max()           # the rule converges to it
        assert np.abs(_outer_subdivided(Vp, Vq, d, 3) - fine).max() < 2e-4 * np.abs(fine).max()
    # far apart: m0 -> A_p A_q / R
    Vfar = Vq + np.array([8e-3, 5e-3])
    m0 = _outer_subdivided(Vp, Vfar, 0.0, 0)[0]
    area = lambda V: 0.5 * abs((V[1, 0] - V[0, 0]) * (V[2, 1] - V[0, 1]) - (V[1, 1] - V[0, 1]) * (V[2, 0] - V[0, 0]))
    Ap = area(Vp)
    Aq = area(Vfar)
    R = np.linalg.norm(Vp.mean(0) - Vfar.mean(0))
    assert abs(m0 / (Ap * Aq / R) - 1) < 2e-3


def test_offplane_pair_is_one_value_either_way(case):
    """The rule over the smaller triangle (ties: the lower index) whichever is the observation: (p, q) and
    (q, p) give the same m0, mrr and swapped observation / source moments, bit for bit."""
    q = case['q']
    nodes, tris, P6, W6 = q['nodes'], q['tris'], q['P6'], q['W6']
    rng = np.random.default_rng(4)
    p = rng.integers(0, len(tris), 50).astype(np.int32)
    r = rng.integers(0, len(tris), 50).astype(np.int32)
    for d in (0.0, 0.2e-3):
        a = oc.offplane_pairs(nodes, tris, P6, W6, p, r, d)
        b = oc.offplane_pairs(nodes, tris, P6, W6, r, p, d)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[3], b[3])
        assert np.array_equal(a[1], b[2]) and np.array_equal(a[2], b[1])


def test_near_charge_correction_is_small_for_a_distant_pair(case):
    """The near correction is accurate minus centroid: for a well separated pair the two agree, so its charge
    moment (qcorr) is a small part of the centroid value G_q A_p A_q (fan_q); a wrong sign or a dropped part
    would make it of that value's size. Near the pair it is not small (the singular part is the point)."""
    g, q, tb, w = case['g'], case['q'], case['tb'], case['w']
    cent, lt = g['cent'], g['lt']
    same = np.flatnonzero(lt == lt[0])
    p = same[0]
    dist = np.linalg.norm(cent[same] - cent[p], axis=1)
    far, adj = same[np.argmax(dist)], same[np.argsort(dist)[1]]
    for qq, small in ((far, True), (adj, False)):
        qc = oc.near_q(g, tb, q, np.array([p], np.int32), np.array([qq], np.int32), w)[0]
        centroid = oc.fan_q(g, tb, np.array([qq], np.int32))[p, 0]
        assert (abs(qc) < 1e-2 * abs(centroid)) == small
    # the pair with itself: the correction is the whole (finite) self term, the centroid rule's being 0 at r = 0
    # for the singular part
    qself = oc.near_q(g, tb, q, np.array([p], np.int32), np.array([p], np.int32), w)[0]
    assert np.isfinite(qself) and abs(qself) > 0


def test_gmres_reaches_its_tolerance_and_stalls_as_told():
    """The reference's seeded complex system (tests/test_solver.py:375-398): residual < 1e-9; a broken
    preconditioner given up after 4 iterations with stall (4, 0.1), all 30 without."""
    from oracle.gmres import gmres_batched
    rng = np.random.default_rng(1)
    n = 200
    ev = (1 + rng.random(n)) * np.exp(1j * rng.uniform(-1.2, 1.2, n))
    Qm = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    A = Qm @ np.diag(ev) @ Qm.conj().T + 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    X, _ = gmres_batched(lambda V: A @ V, B, rtol=1e-10, restart=60, maxiter=5)
    assert np.max(np.linalg.norm(A @ X - B, axis=0) / np.linalg.norm(B, axis=0)) < 1e-9
    good = np.linalg.inv(A + 1e-3 * rng.standard_normal((n, n)))
    X, it = gmres_batched(lambda V: A @ V, B, M=lambda R: good @ R, rtol=1e-8, restart=30, maxiter=1, stall=(4, 0.1))
    assert max(it) <= 4 and np.max(np.linalg.norm(A @ X - B, axis=0) / np.linalg.norm(B, axis=0)) < 1e-7
    calls = []
    bad = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    X, it = gmres_batched(lambda V: A @ V, B, M=lambda R: calls.append(1) or bad @ R, rtol=1e-8, restart=30,
                          maxiter=3, stall=(4, 0.1))
    assert it == [4, 4] and len(calls) == 4 and np.isfinite(X).all()
    assert gmres_batched(lambda V: A @ V, B, M=lambda R: bad @ R, rtol=1e-8, restart=30, maxiter=1)[1] == [30, 30]


def test_matrix_free_products_are_the_references(case):
    """matvec_far / matvec_far_multi (triangle-aggregated far product) bit for bit with the reference's kernels, and
    equal to the dense far fill times X (a different summation: the far block is the same operator); near_matvec
    equal to the COO triplets applied to x (the reference adds atomically: equal up to the order)."""
    g, q, tb, w, st = case['g'], case['q'], case['tb'], case['w'], case['stack']
    Ne = len(g['ep'])
    rng = np.random.default_rng(5)
    X = rng.standard_normal((Ne, 3)) + 1j * rng.standard_normal((Ne, 3))
    Y = oc.matvec_far_multi(X, g, tb, w)
    Zf = np.zeros((Ne, Ne), complex)
    oc.fill_far(Zf, g, tb, w)
    assert np.abs(Y - Zf @ X).max() <= 1e-12 * np.abs(Zf @ X).max()
    assert np.array_equal(oc.matvec_far(X[:, 1], g, tb, w), Y[:, 1])
    P, Q = oc.near_pairs(case['rwg'].cent, case['size'], st.z[case['tl']], 5.0)
    P32, Q32 = P.astype(np.int32), Q.astype(np.int32)
    ri, ci, vv, _ = oc.near_coo(g, tb, q, P32, Q32, w)
    ok = ri >= 0
    yc = np.zeros(Ne, complex)
    np.add.at(yc, ri[ok], vv[ok] * X[ci[ok], 0])
    yn = oc.near_matvec(X[:, 0], g, tb, q, P32, Q32, w)
    assert np.abs(yn - yc).max() <= 1e-14 * np.abs(yc).max()
    R = refcore()
    assert np.array_equal(Y, R.matvec_far_multi(np.ascontiguousarray(X), g, tb, w))
    assert np.array_equal(oc.matvec_far(X[:, 2], g, tb, w), R.matvec_far(np.ascontiguousarray(X[:, 2]), g, tb, w))
    yr = R.near_matvec(np.ascontiguousarray(X[:, 0]), g, tb, q, P32, Q32, w)
    assert np.abs(yn - yr).max() <= 1e-14 * np.abs(yr).max()
