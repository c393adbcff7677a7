"""The oracle's C kernels (m25_oracle.c) behind the reference's compiled-module names and argument layouts
(src/mom25d/core/mom25d_core.cpp:1658-1702). TEST INFRASTRUCTURE ONLY.

Dict layouts (src/mom25d/fast.py:389-400, 556-568):
  g  : ep, em, lt, te (int32), ta, tv, ap, am, vp, vm, cent, area (float64)
  tb : A, Q (ntab x nfine complex128), dr, d, dimg (float64), cA, cq, cAi, cqi (complex128), tab (L x L int32),
       kind (ntab int32)
  q  : nodes, P3, W3, P6, W6 (float64), tris (int32), optional stat (n x 12 float64) and stat_ready
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    """liboracle.so (built by `make -C oracle`), loaded once."""
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError(f"the oracle is not built ({path}): make -C oracle")
        _lib = C.CDLL(path)
        _lib.orc_last_error.restype = C.c_char_p
        _lib.orc_near_pairs.restype = C.c_long
    return _lib


class _Tables(C.Structure):
    _fields_ = [('ntab', C.c_int), ('nfine', C.c_int), ('L', C.c_int),
                ('A', C.c_void_p), ('Q', C.c_void_p), ('cA', C.c_void_p), ('cq', C.c_void_p),
                ('cAi', C.c_void_p), ('cqi', C.c_void_p),
                ('dr', C.c_void_p), ('d', C.c_void_p), ('dimg', C.c_void_p),
                ('tab', C.c_void_p), ('kind', C.c_void_p)]


class _Mesh(C.Structure):
    _fields_ = [('Ne', C.c_int), ('Nt', C.c_int),
                ('ep', C.c_void_p), ('em', C.c_void_p), ('lt', C.c_void_p), ('te', C.c_void_p),
                ('ap', C.c_void_p), ('am', C.c_void_p), ('vp', C.c_void_p), ('vm', C.c_void_p),
                ('cent', C.c_void_p), ('area', C.c_void_p), ('ta', C.c_void_p), ('tv', C.c_void_p)]


class _NearOut(C.Structure):
    _fields_ = [('mode', C.c_int), ('Z', C.c_void_p), ('ld', C.c_long), ('single', C.c_int),
                ('ri', C.c_void_p), ('ci', C.c_void_p), ('vv', C.c_void_p), ('qcorr', C.c_void_p),
                ('stat', C.c_void_p), ('stat_ready', C.c_int), ('x', C.c_void_p), ('y', C.c_void_p)]


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class _Hold:
    """Keeps the contiguous copies alive while a C view of them is used."""
    def __init__(self):
        self.keep = []

    def __call__(self, a, dt):
        a = _c(a, dt)
        self.keep.append(a)
        return _p(a)


def _tables(tb, h):
    A = _c(tb['A'], complex)
    t = _Tables()
    t.ntab, t.nfine, t.L = A.shape[0], A.shape[1], np.asarray(tb['tab']).shape[0]
    t.A, t.Q = h(A, complex), h(tb['Q'], complex)
    for k in ('cA', 'cq', 'cAi', 'cqi'):
        setattr(t, k, h(tb[k], complex))
    for k in ('dr', 'd', 'dimg'):
        setattr(t, k, h(tb[k], float))
    t.tab, t.kind = h(tb['tab'], np.int32), h(tb['kind'], np.int32)
    return t


def _mesh(g, h):
    m = _Mesh()
    m.Ne, m.Nt = len(g['ep']), len(g['lt'])
    for k in ('ep', 'em', 'lt', 'te'):
        setattr(m, k, h(g[k], np.int32))
    for k in ('ap', 'am', 'vp', 'vm', 'cent', 'area', 'ta', 'tv'):
        setattr(m, k, h(g[k], float))
    return m


def _check(rc):
    if rc != 0:
        raise ValueError(lib().orc_last_error().decode())


def fill_far(Z, g, tb, w):
    """Far-field entries of every edge pair into Z (C order, complex128) in place (mom25d_core.cpp:132-173)."""
    _fill(Z, g, tb, w, 0)


def fill_far_c64(Z, g, tb, w):
    """The same into a complex64 Z (mom25d_core.cpp:1662)."""
    _fill(Z, g, tb, w, 1)


def _fill(Z, g, tb, w, single):
    if not (Z.flags.c_contiguous and Z.dtype == (np.complex64 if single else np.complex128)):
        raise ValueError("fill_far: Z must be a C-ordered complex array of its kind")
    h = _Hold()
    m, t = _mesh(g, h), _tables(tb, h)
    _check(lib().orc_fill_far(_p(Z), C.c_int(single), C.c_long(Z.shape[1]), C.byref(m), C.byref(t),
                              C.c_double(w)))


def far_block(rows, cols, g, tb, w):
    """Centroid-rule entries of row x column edge lists (mom25d_core.cpp:176-211)."""
    h = _Hold()
    rows, cols = _c(rows, np.int32), _c(cols, np.int32)
    out = np.zeros((len(rows), len(cols)), complex)
    m, t = _mesh(g, h), _tables(tb, h)
    _check(lib().orc_far_block(_p(rows), C.c_int(len(rows)), _p(cols), C.c_int(len(cols)), C.byref(m),
                               C.byref(t), C.c_double(w), _p(out)))
    return out


def offplane_pairs(nodes, tris, P6, W6, P, Q, d):
    """Analytic 1/sqrt(rho^2 + d^2) moments of pairs (mom25d_core.cpp:367-387): (m0, mro, mrs, mrr)."""
    h = _Hold()
    P, Q = _c(P, np.int32), _c(Q, np.int32)
    n = len(P)
    m0, mro, mrs, mrr = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 2)), np.zeros(n)
    _check(lib().orc_offplane_pairs(h(nodes, float), h(tris, np.int32), h(P6, float), h(W6, float), _p(P), _p(Q),
                                    C.c_int(n), C.c_double(d), _p(m0), _p(mro), _p(mrs), _p(mrr)))
    return m0, mro, mrs, mrr


def _near(mode, g, tb, q, P, Q, w, Z=None):
    h = _Hold()
    P, Q = _c(P, np.int32), _c(Q, np.int32)
    n = len(P)
    o = _NearOut()
    o.mode = mode
    qc = np.zeros(n, complex)
    o.qcorr = _p(qc)
    out = {}
    if mode == 0:
        if not (Z.flags.c_contiguous and Z.dtype in (np.complex64, np.complex128)):
            raise ValueError("near_into_dense: Z must be a C-ordered complex64 or complex128 array")
        o.Z, o.ld, o.single = _p(Z), Z.shape[1], int(Z.dtype == np.complex64)
    elif mode == 1:
        x = _c(Z, complex)
        out['x'], out['y'] = x, np.zeros(len(g['ep']), complex)
        o.x, o.y = _p(x), _p(out['y'])
    elif mode == 2:
        out['ri'], out['ci'] = np.zeros(9 * n, np.int32), np.zeros(9 * n, np.int32)
        out['vv'] = np.zeros(9 * n, complex)
        o.ri, o.ci, o.vv = _p(out['ri']), _p(out['ci']), _p(out['vv'])
    if 'stat' in q:
        st = q['stat']
        if not (st.dtype == np.float64 and st.flags.c_contiguous and st.size == 12 * n):
            raise ValueError("quad['stat']: a C-ordered float64 array of n x 12")
        o.stat, o.stat_ready = _p(st), int(bool(q['stat_ready']))
    m, t = _mesh(g, h), _tables(tb, h)
    _check(lib().orc_near(C.byref(m), C.byref(t), h(q['nodes'], float), h(q['tris'], np.int32), h(q['P3'], float),
                          h(q['W3'], float), h(q['P6'], float), h(q['W6'], float), _p(P), _p(Q), C.c_int(n),
                          C.c_double(w), C.byref(o)))
    return qc, out


def near_into_dense(Z, g, tb, q, P, Q, w):
    """Z += the near-field correction (atomic adds); returns qcorr (mom25d_core.cpp:520-537)."""
    return _near(0, g, tb, q, P, Q, w, Z)[0]


def near_matvec(x, g, tb, q, P, Q, w):
    """y = dZ x, the near-field correction applied to x (Ne,) without storing it (atomic adds; mom25d_core.cpp:
    539-556)."""
    return _near(1, g, tb, q, P, Q, w, x)[1]['y']


def matvec_far_multi(X, g, tb, w):
    """Y = Z_far X for X (Ne, k), triangle-aggregated, without Z (mom25d_core.cpp:257-314)."""
    h = _Hold()
    X = _c(X, complex)
    if X.ndim != 2 or X.shape[0] != len(g['ep']):
        raise RuntimeError("matvec_far_multi: X must be (Ne, k)")
    Y = np.zeros(X.shape, complex)
    m, t = _mesh(g, h), _tables(tb, h)
    _check(lib().orc_matvec_far_multi(_p(X), C.c_int(X.shape[1]), C.byref(m), C.byref(t), C.c_double(w), _p(Y)))
    return Y


def matvec_far(x, g, tb, w):
    """y = Z_far x for one column (mom25d_core.cpp:214-252)."""
    return matvec_far_multi(np.asarray(x, complex).reshape(-1, 1), g, tb, w)[:, 0]


def near_coo(g, tb, q, P, Q, w):
    """(ri, ci, vv, qcorr): 9 triplets a pair, row -1 for unused slots (mom25d_core.cpp:559-575)."""
    qc, o = _near(2, g, tb, q, P, Q, w)
    return o['ri'], o['ci'], o['vv'], qc


def near_q(g, tb, q, P, Q, w):
    """The near pairs' G_q moment corrections alone (mom25d_core.cpp:579-594)."""
    return _near(3, g, tb, q, P, Q, w)[0]


def near_pairs(cent, size, z, fac):
    """Near triangle pairs, sorted by (p, q) (mom25d_core.cpp:1216-1263): (P, Q) int64."""
    h = _Hold()
    Nt = len(size)
    cp, sp, zp = h(cent, float), h(size, float), h(z, float)
    n = lib().orc_near_pairs(cp, sp, zp, C.c_int(Nt), C.c_double(fac), None, None)
    P, Q = np.zeros(n, np.int64), np.zeros(n, np.int64)
    lib().orc_near_pairs(cp, sp, zp, C.c_int(Nt), C.c_double(fac), _p(P), _p(Q))
    return P, Q


def radial_table_sum(px, py, W, vx, vy, table, dr):
    """out[t, v] = sum_q D(|P[t, q] - V[v]|) W[t, q] (mom25d_core.cpp:1269-1298)."""
    h = _Hold()
    px, py, W = _c(px, float), _c(py, float), _c(W, float)
    vx, vy = _c(vx, float), _c(vy, float)
    if py.shape != px.shape or W.shape != px.shape or vy.shape != vx.shape:
        raise ValueError("radial_table_sum: px, py, W must be (nt, nq), vx and vy (nv,)")
    nt, nq = px.shape
    out = np.zeros((nt, len(vx)), complex)
    tab = _c(table, complex)
    _check(lib().orc_radial_table_sum(_p(px), _p(py), _p(W), C.c_long(nt), C.c_long(nq), _p(vx), _p(vy),
                                      C.c_long(len(vx)), _p(tab), C.c_long(len(tab)), C.c_double(dr), _p(out)))
    return out


def hankel_real(rho, kr, W):
    """out[s, i] = sum_k J0(rho_i kr_k) W[s, k], real kr (mom25d_core.cpp:1430-1454)."""
    rho, kr, W = _c(rho, float), _c(kr, float), np.atleast_2d(_c(W, complex))
    if W.shape[1] != len(kr):
        raise ValueError("hankel_real: W must be (nspec, len(kr))")
    out = np.zeros((W.shape[0], len(rho)), complex)
    _check(lib().orc_hankel_real(_p(rho), C.c_int(len(rho)), _p(kr), C.c_int(len(kr)), _p(W), C.c_int(W.shape[0]),
                                 _p(out)))
    return out


def bessel_jn_complex(n, z):
    """J_n(z) for complex z, n = 0, 1, 2, elementwise (mom25d_core.cpp:1510-1558)."""
    z = _c(z, complex)
    out = np.zeros(z.shape, complex)
    _check(lib().orc_bessel_jn_complex(C.c_int(int(n)), _p(z), C.c_long(z.size), _p(out)))
    return out


def scale_c64(Z, d):
    """Z <- D Z D in place, and its complex64 copy (mom25d_core.cpp:1563-1580)."""
    d = _c(d, float)
    if Z.ndim != 2 or Z.shape[0] != Z.shape[1] or len(d) != Z.shape[0] or not Z.flags.c_contiguous \
            or Z.dtype != np.complex128:
        raise ValueError("scale_c64: Z square, d its size")
    out = np.zeros(Z.shape, np.complex64)
    _check(lib().orc_scale_c64(_p(Z), C.c_long(Z.shape[0]), _p(d), _p(out)))
    return out


def fan_q(g, tb, F):
    """G_q x area_t x area_f between every triangle and the fan triangles F (mom25d_core.cpp:1585-1603)."""
    h = _Hold()
    F = _c(F, np.int32)
    out = np.zeros((len(g['lt']), len(F)), complex)
    m, t = _mesh(g, h), _tables(tb, h)
    _check(lib().orc_fan_q(C.byref(m), C.byref(t), _p(F), C.c_int(len(F)), _p(out)))
    return out


def num_threads():
    return lib().orc_num_threads()
