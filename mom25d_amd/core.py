"""The reference's compiled-kernel module, `mom25d._core` (src/mom25d/core/mom25d_core.cpp:1658-1702), served by
the gfx950 kernels: the same function names, argument meanings, dict layouts, in-place semantics and exception
types, so that a caller of the reference (src/mom25d/fast.py:27-51, `fast._core`) can be pointed here.

Host NumPy arrays in and out, as the reference takes them: each call uploads its inputs, runs on the current GPU
stream and copies the results back (synchronously). The device-resident path (tables, geometry and the near
pairs' analytic moments kept on the GPU across frequencies) is mom25d_amd.fast; this module is the drop-in seam.
"""
import numpy as np
import torch

from . import device as dv

static_openmp = False


def _tabs(tb):
    return dv.DevTables.from_dict(tb)


def _mesh(g, q=None):
    return dv.DevMesh(g, q)


def _check_z(Z, what):
    if not (isinstance(Z, np.ndarray) and Z.ndim == 2 and Z.flags.c_contiguous and
            Z.dtype in (np.complex64, np.complex128)):
        raise ValueError(f"{what}: Z must be a C-ordered complex64 or complex128 array")


def fill_far(Z, g, tb, w):
    """Far-field entries of every edge pair into Z[:Ne, :Ne] (complex128), in place (mom25d_core.cpp:132-173)."""
    _fill(Z, g, tb, w, np.complex128)


def fill_far_c64(Z, g, tb, w):
    """The same into a complex64 Z (mom25d_core.cpp:1662)."""
    _fill(Z, g, tb, w, np.complex64)


def _fill(Z, g, tb, w, dt):
    _check_z(Z, 'fill_far')
    if Z.dtype != dt:
        raise TypeError(f"fill_far: Z must be {np.dtype(dt).name}")
    m = _mesh(g)
    Ne = m.Ne
    if Z.shape[0] < Ne or Z.shape[1] < Ne:
        raise ValueError("fill_far: Z smaller than the edge block")
    Zd = torch.empty((Ne, Ne), dtype=dv.c128 if dt == np.complex128 else dv.c64, device=dv.device())
    dv.fill_far(Zd, m, _tabs(tb), w)
    Z[:Ne, :Ne] = Zd.cpu().numpy()


def far_block(rows, cols, g, tb, w):
    """Centroid-rule entries of row x column edge lists (mom25d_core.cpp:176-211)."""
    return dv.far_block(np.asarray(rows, np.int32), np.asarray(cols, np.int32), _mesh(g), _tabs(tb), w).cpu().numpy()


def _stat_plan(quad, P, Q, m):
    plan = dv.NearPlan(m, np.asarray(P, np.int32), np.asarray(Q, np.int32), keep_stat='stat' in quad)
    if 'stat' in quad:
        st = quad['stat']
        if not (isinstance(st, np.ndarray) and st.dtype == np.float64 and st.flags.c_contiguous and
                st.size == 12 * len(P)):
            raise ValueError("quad['stat']: a C-ordered float64 array of n x 12")
        plan.stat.copy_(torch.from_numpy(st.reshape(-1, 12)))
    return plan


def _near(g, tb, quad, P, Q, w):
    m, t = _mesh(g, quad), _tabs(tb)
    plan = _stat_plan(quad, P, Q, m)
    if 'stat' in quad:
        plan.stat_ready, plan.stat_kind = bool(quad['stat_ready']), t.sig
    M, qc = dv.near_moments(m, t, plan, w, stat='stat' in quad)
    if 'stat' in quad and not quad['stat_ready']:
        quad['stat'][...] = plan.stat.cpu().numpy().reshape(quad['stat'].shape)
    return m, plan, M, qc


def near_into_dense(Z, g, tb, quad, P, Q, w):
    """Z += the near-field correction (complex128 or complex64 by itemsize); returns qcorr (mom25d_core.cpp:
    520-537). quad['stat'] is written in place when quad['stat_ready'] is false (the caller then sets it)."""
    _check_z(Z, 'near_into_dense')
    m, plan, M, qc = _near(g, tb, quad, P, Q, w)
    Zd = torch.from_numpy(Z).to(dv.device())
    dv.near_scatter(Zd, m, plan, M, w)
    Z[...] = Zd.cpu().numpy()
    return qc.cpu().numpy()


def near_coo(g, tb, quad, P, Q, w):
    """(ri, ci, vv, qcorr): 9 triplets a pair, row -1 for unused slots (mom25d_core.cpp:559-575)."""
    m, plan, M, qc = _near(g, tb, quad, P, Q, w)
    ri, ci, vv = dv.near_coo(m, plan, M, w)
    return ri.cpu().numpy(), ci.cpu().numpy(), vv.cpu().numpy(), qc.cpu().numpy()


def near_q(g, tb, quad, P, Q, w):
    """The near pairs' G_q moment corrections alone (mom25d_core.cpp:579-594)."""
    return _near(g, tb, quad, P, Q, w)[3].cpu().numpy()


def near_matvec(x, g, tb, quad, P, Q, w):
    """y = dZ x: the near-field correction applied to x (Ne,) without storing it in Z (mom25d_core.cpp:539-556)."""
    x = np.asarray(x, complex)
    if x.ndim != 1 or len(x) != len(g['ep']):
        raise ValueError("near_matvec: x must be (Ne,)")
    m, plan, M, _ = _near(g, tb, quad, P, Q, w)
    vals = dv.near_values(m, plan, M, w)
    return dv.csr_matvec(plan, vals, dv.on_dev(x, dv.c128)).cpu().numpy()


def matvec_far(x, g, tb, w):
    """y = Z_far x for one column, without forming Z (mom25d_core.cpp:214-252)."""
    x = np.asarray(x, complex)
    if x.ndim != 1 or len(x) != len(g['ep']):
        raise ValueError("matvec_far: x must be (Ne,)")
    return dv.matvec_far(dv.on_dev(x, dv.c128), _mesh(g), _tabs(tb), w).cpu().numpy()


def matvec_far_multi(X, g, tb, w):
    """Y = Z_far X for X (Ne, k), one kernel evaluation per triangle pair for all columns (mom25d_core.cpp:
    257-314)."""
    X = np.asarray(X, complex)
    if X.ndim != 2 or X.shape[0] != len(g['ep']):
        raise RuntimeError("matvec_far_multi: X must be (Ne, k)")
    return dv.matvec_far(dv.on_dev(X, dv.c128), _mesh(g), _tabs(tb), w).cpu().numpy()


def near_pairs(cent, size, z, fac):
    """(P, Q) int64 near triangle pairs, sorted by (p, q) (mom25d_core.cpp:1216-1263)."""
    P, Q = dv.near_pairs(np.asarray(cent, float), np.asarray(size, float), np.asarray(z, float), float(fac))
    return P.cpu().numpy().astype(np.int64), Q.cpu().numpy().astype(np.int64)


def fan_q(g, tb, F):
    """G_q x area_t x area_f between every triangle and the fan triangles F (mom25d_core.cpp:1585-1603)."""
    return dv.fan_q(_mesh(g), _tabs(tb), np.asarray(F, np.int32)).cpu().numpy()


def radial_table_sum(px, py, W, vx, vy, table, dr):
    """out[t, v] = sum_q D(|P[t, q] - V[v]|) W[t, q] (mom25d_core.cpp:1269-1298)."""
    px, py, W = (np.asarray(a, float) for a in (px, py, W))
    vx, vy = np.asarray(vx, float), np.asarray(vy, float)
    if px.ndim != 2 or py.shape != px.shape or W.shape != px.shape or vy.shape != vx.shape:
        raise ValueError("radial_table_sum: px, py, W must be (nt, nq), vx and vy (nv,)")
    if len(table) < 2 or not dr > 0:
        raise ValueError("radial_table_sum: a table of two points or more, dr > 0")
    return dv.radial_table_sum(px, py, W, vx, vy, np.asarray(table, complex), dr).cpu().numpy()


def hankel_real(rho, kr, W):
    """out[s, i] = sum_k J0(rho_i kr_k) W[s, k] for real kr (mom25d_core.cpp:1430-1454)."""
    kr = np.asarray(kr, float)
    W = np.atleast_2d(np.asarray(W, complex))
    if W.shape[1] != len(kr):
        raise ValueError("hankel_real: W must be (nspec, len(kr))")
    return dv.hankel_sum(np.asarray(rho, float), kr.astype(complex), W).cpu().numpy()


def bessel_jn_complex(n, z):
    """J_n(z) for complex z, n = 0, 1, 2, elementwise (mom25d_core.cpp:1510-1558)."""
    if n not in (0, 1, 2):
        raise ValueError("bessel_jn_complex: order 0, 1 or 2")
    z = np.asarray(z, complex)
    return dv.bessel_jn_complex(int(n), z.ravel()).cpu().numpy().reshape(z.shape)


def scale_c64(Z, d):
    """Z <- D Z D in place, and its complex64 copy (mom25d_core.cpp:1563-1580)."""
    d = np.asarray(d, float)
    if not (isinstance(Z, np.ndarray) and Z.ndim == 2 and Z.shape[0] == Z.shape[1] == len(d) and
            Z.flags.c_contiguous and Z.dtype == np.complex128):
        raise ValueError("scale_c64: Z square, d its size")
    Zd = torch.from_numpy(Z).to(dv.device())
    Z64 = torch.empty(Z.shape, dtype=dv.c64, device=dv.device())
    dv.scale_c64(Zd, dv.on_dev(d, dv.f64), Z64)
    Z[...] = Zd.cpu().numpy()
    return Z64.cpu().numpy()


def num_threads():
    """The GPU's compute units (the reference: OpenMP threads)."""
    dv.device()
    return torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
