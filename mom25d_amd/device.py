"""Device-resident data and thin wrappers over the C ABI (libm25x.so): the Green's tables, the mesh and
quadrature geometry, the near-pair plan, and one Python function per kernel entry point.

Memory comes from PyTorch's caching allocator on the current device; every call runs on torch's current stream
(its handle is passed to the library). Nothing here falls back to the CPU: without a GPU or without the built
library the calls raise.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L

f64, i32, c128, c64 = torch.float64, torch.int32, torch.complex128, torch.complex64


def device():
    """The current HIP device (RuntimeError without one), initialised for the library."""
    if not torch.cuda.is_available():
        raise RuntimeError("mom25d_amd: no GPU - the kernels are HIP (gfx950) only")
    idx = torch.cuda.current_device()
    L.init(idx)
    return torch.device('cuda', idx)


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def on_dev(a, dtype):
    """A contiguous device tensor of dtype from a NumPy array (or tensor)."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device(), dtype=dtype).contiguous()
    np_dt = {f64: np.float64, i32: np.int32, c128: np.complex128, c64: np.complex64,
             torch.int64: np.int64}[dtype]
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np_dt)).to(device())


def _cx_pairs(z):
    """complex128 tensor (n,) -> float64 (n, 2) view-compatible contiguous copy."""
    return torch.view_as_real(z.to(c128).contiguous()).contiguous()


class DevTables:
    """The kernels' tables on the device (the reference's tb dict, src/mom25d/fast.py:556-568). AQ interleaves
    (Re A, Im A, Re Q, Im Q) per point so one lookup reads 64 contiguous bytes."""

    def __init__(self, A, Q, dr, d, dimg, cA, cq, cAi, cqi, tab, kind):
        dv = device()
        A, Q = A.to(dv, c128), Q.to(dv, c128)
        self.ntab, self.nfine = A.shape
        self.AQ = torch.stack([A.real, A.imag, Q.real, Q.imag], -1).contiguous()
        self.dr, self.d, self.dimg = (x.to(dv, f64).contiguous() for x in (dr, d, dimg))
        self.cA, self.cq, self.cAi, self.cqi = (_cx_pairs(x.to(dv)) for x in (cA, cq, cAi, cqi))
        self.tab, self.kind = tab.to(dv, i32).contiguous(), kind.to(dv, i32).contiguous()
        self.L = self.tab.shape[0]
        # which parts are extracted (the near pairs' kept analytic moments depend on them)
        self.sig = self.kind.cpu().numpy().tobytes() + self.dimg.cpu().numpy().tobytes() + \
            self.d.cpu().numpy().tobytes()
        self._struct()

    @classmethod
    def packed(cls, AQ, scal, tab, kind):
        """From AQ (ntab, nfine, 4) float64 on the device and the per-table scalars on the host (dict of arrays
        dr, d, dimg, cA, cq, cAi, cqi; tab, kind int32): one upload of the scalars, no read-back."""
        t = cls.__new__(cls)
        dv = device()
        t.AQ = AQ.to(dv, f64).contiguous()
        t.ntab, t.nfine = int(t.AQ.shape[0]), int(t.AQ.shape[1])
        n = t.ntab
        real = np.concatenate([scal['dr'], scal['d'], scal['dimg']]).astype(np.float64)
        cplx = np.concatenate([scal[k] for k in ('cA', 'cq', 'cAi', 'cqi')]).astype(np.complex128)
        ints = np.concatenate([np.asarray(tab, np.int32).ravel(), np.asarray(kind, np.int32)])
        blob = torch.from_numpy(np.concatenate([real, cplx.view(np.float64)])).to(dv)
        t.dr, t.d, t.dimg = blob[:n], blob[n:2 * n], blob[2 * n:3 * n]
        cx = blob[3 * n:].view(4, n, 2)
        t.cA, t.cq, t.cAi, t.cqi = cx[0], cx[1], cx[2], cx[3]
        iblob = torch.from_numpy(ints).to(dv)
        t.L = int(np.asarray(tab).shape[0])
        t.tab, t.kind = iblob[:t.L * t.L].view(t.L, t.L), iblob[t.L * t.L:]
        t.sig = np.asarray(kind, np.int32).tobytes() + np.asarray(scal['dimg'], np.float64).tobytes() + \
            np.asarray(scal['d'], np.float64).tobytes()
        t._host_scal = scal
        t._struct()
        return t

    def _struct(self):
        s = L.Tables()
        s.ntab, s.nfine, s.L = self.ntab, self.nfine, self.L
        s.AQ = self.AQ.data_ptr()
        s.dr, s.d, s.dimg = self.dr.data_ptr(), self.d.data_ptr(), self.dimg.data_ptr()
        s.cA, s.cq, s.cAi, s.cqi = (x.data_ptr() for x in (self.cA, self.cq, self.cAi, self.cqi))
        s.tab, s.kind = self.tab.data_ptr(), self.kind.data_ptr()
        self.struct = s

    @classmethod
    def from_dict(cls, tb):
        t = lambda k, dt: torch.from_numpy(np.ascontiguousarray(tb[k], dtype=dt))
        return cls(t('A', np.complex128), t('Q', np.complex128), t('dr', np.float64), t('d', np.float64),
                   t('dimg', np.float64), t('cA', np.complex128), t('cq', np.complex128), t('cAi', np.complex128),
                   t('cqi', np.complex128), t('tab', np.int32), t('kind', np.int32))

    def A(self):
        return torch.complex(self.AQ[..., 0], self.AQ[..., 1])

    def Q(self):
        return torch.complex(self.AQ[..., 2], self.AQ[..., 3])


class DevMesh:
    """The mesh, RWG and quadrature geometry on the device (the reference's g and q dicts, src/mom25d/fast.py:
    389-400). Built once per mesh."""
    G_KEYS = (('ep', i32), ('em', i32), ('lt', i32), ('te', i32), ('ap', f64), ('am', f64), ('vp', f64),
              ('vm', f64), ('cent', f64), ('area', f64), ('ta', f64), ('tv', f64))
    Q_KEYS = (('nodes', f64), ('tris', i32), ('P3', f64), ('W3', f64), ('P6', f64), ('W6', f64))

    def __init__(self, g, q=None):
        self.g = {k: on_dev(g[k], dt) for k, dt in self.G_KEYS}
        self.Ne, self.Nt = int(self.g['ep'].shape[0]), int(self.g['lt'].shape[0])
        m = L.Mesh()
        m.Ne, m.Nt = self.Ne, self.Nt
        for k, _ in self.G_KEYS:
            setattr(m, k, self.g[k].data_ptr())
        self.struct = m
        self.q, self.quad = None, None
        if q is not None:
            self.q = {k: on_dev(q[k], dt) for k, dt in self.Q_KEYS}
            s = L.Quad()
            for k, _ in self.Q_KEYS:
                setattr(s, k, self.q[k].data_ptr())
            self.quad = s


# ---- far field ----------------------------------------------------------------------------------------------
def fill_far(Z, mesh, tabs, w):
    """Far-field entries into the Ne x Ne edge block of Z (device complex128 / complex64, row-major)."""
    if Z.dim() != 2 or Z.stride(1) != 1 or Z.dtype not in (c128, c64) or Z.shape[0] < mesh.Ne or Z.shape[1] < mesh.Ne:
        raise ValueError("fill_far: Z must be a row-major complex tensor holding the edge block")
    L.call('m25x_fill_far', ptr(Z), L.C128 if Z.dtype == c128 else L.C64, Z.stride(0), C.byref(mesh.struct),
           C.byref(tabs.struct), float(w), stream())


def far_block(rows, cols, mesh, tabs, w):
    rows, cols = on_dev(rows, i32), on_dev(cols, i32)
    out = torch.empty((len(rows), len(cols)), dtype=c128, device=device())
    L.call('m25x_far_block', ptr(rows), len(rows), ptr(cols), len(cols), C.byref(mesh.struct), C.byref(tabs.struct),
           float(w), ptr(out), stream())
    return out


# ---- near field ---------------------------------------------------------------------------------------------
class NearPlan:
    """The near pairs (P, Q) and, for the deterministic scatter, every near Z entry with its contributions in a
    fixed order (pair index, then slot): geometry only, built once per mesh. stat: the pairs' analytic moments
    (12 doubles a pair), kept on the device across frequencies."""

    def __init__(self, mesh, P, Q, keep_stat=True):
        dv = device()
        self.P, self.Q = on_dev(P, i32), on_dev(Q, i32)
        self.n = int(self.P.shape[0])
        te = mesh.g['te'].long()
        n, Ne = self.n, mesh.Ne
        if n:
            rows = te[self.P.long()][:, :, None].expand(n, 3, 3)
            cols = te[self.Q.long()][:, None, :].expand(n, 3, 3)
            valid = ((rows >= 0) & (cols >= 0)).reshape(-1)
            code = torch.arange(9 * n, device=dv, dtype=torch.int64)[valid]
            key = (rows.reshape(-1) * Ne + cols.reshape(-1))[valid]
            key, order = torch.sort(key, stable=True)
            self.contrib = code[order].to(i32).contiguous()
            uniq, counts = torch.unique_consecutive(key, return_counts=True)
            self.erow, self.ecol = (uniq // Ne).to(i32).contiguous(), (uniq % Ne).to(i32).contiguous()
            self.estart = torch.zeros(len(uniq) + 1, dtype=i32, device=dv)
            self.estart[1:] = torch.cumsum(counts, 0).to(i32)
        else:
            self.contrib = self.erow = self.ecol = torch.zeros(0, dtype=i32, device=dv)
            self.estart = torch.zeros(1, dtype=i32, device=dv)
        self.nentry = int(self.erow.shape[0])
        # the entries of each row (they are sorted by row, then column): the correction as a CSR matrix
        self.rowptr = torch.searchsorted(self.erow.long(), torch.arange(Ne + 1, device=dv, dtype=torch.int64),
                                         right=False).to(i32).contiguous()
        self.stat = torch.zeros((n, 12), dtype=f64, device=dv) if keep_stat else None
        self.stat_ready = False
        self.stat_kind = None


def near_moments(mesh, tabs, plan, w, stat=True):
    """(M (n x 7 complex128), qcorr (n complex128)) of the plan's pairs; the analytic parts are kept in plan.stat
    (computed at the first call, or again when the tables' extracted parts change)."""
    M = torch.empty((plan.n, 7), dtype=c128, device=device())
    qc = torch.empty(plan.n, dtype=c128, device=device())
    use = stat and plan.stat is not None
    if use and plan.stat_kind != tabs.sig:
        plan.stat_ready, plan.stat_kind = False, tabs.sig
    L.call('m25x_near_moments', C.byref(mesh.struct), C.byref(tabs.struct), C.byref(mesh.quad), ptr(plan.P),
           ptr(plan.Q), plan.n, float(w), ptr(plan.stat) if use else None, int(use and plan.stat_ready), ptr(M),
           ptr(qc), stream())
    if use:
        plan.stat_ready = True
    return M, qc


def near_scatter(Z, mesh, plan, M, w):
    """Z[near entries] += the correction, each entry's contributions summed in a fixed order (no atomics)."""
    if Z.dim() != 2 or Z.stride(1) != 1 or Z.dtype not in (c128, c64):
        raise ValueError("near_scatter: Z must be a row-major complex tensor")
    L.call('m25x_near_scatter', ptr(Z), L.C128 if Z.dtype == c128 else L.C64, Z.stride(0), C.byref(mesh.struct),
           ptr(plan.P), ptr(plan.Q), ptr(M), plan.nentry, ptr(plan.erow), ptr(plan.ecol), ptr(plan.estart),
           ptr(plan.contrib), float(w), stream())


def near_values(mesh, plan, M, w):
    """The correction of every near entry (plan.nentry complex128, the rows / columns plan.erow / plan.ecol)."""
    vals = torch.empty(plan.nentry, dtype=c128, device=device())
    L.call('m25x_near_values', C.byref(mesh.struct), ptr(plan.P), ptr(plan.Q), ptr(M), plan.nentry, ptr(plan.estart),
           ptr(plan.contrib), float(w), ptr(vals), stream())
    return vals


def csr_matvec(plan, vals, X, Y=None):
    """Y (+)= dZ X with dZ the near correction (plan's rows and columns, vals from near_values); X (Ne, k) or
    (Ne,). Adds into Y when given."""
    one = X.dim() == 1
    X = X.reshape(X.shape[0], -1).to(c128).contiguous()
    n, k = X.shape[0], X.shape[1]
    if n < len(plan.rowptr) - 1:
        raise ValueError("csr_matvec: X has fewer rows than the edge block")
    add = Y is not None
    out = Y.reshape(Y.shape[0], -1) if add else torch.empty((len(plan.rowptr) - 1, k), dtype=c128, device=X.device)
    if not out.is_contiguous() or out.dtype != c128:
        raise ValueError("csr_matvec: Y must be a contiguous complex128 tensor")
    L.call('m25x_csr_matvec', ptr(plan.rowptr), ptr(plan.ecol), ptr(vals), len(plan.rowptr) - 1, ptr(X), k, k,
           ptr(out), k, int(add), stream())
    return out[:, 0] if one and not add else out


def matvec_far(X, mesh, tabs, w):
    """Y = Z_far X (device complex128 (Ne,) or (Ne, k)) without forming Z; columns 8 at a time."""
    one = X.dim() == 1
    X = X.reshape(X.shape[0], -1).to(c128)
    if X.shape[0] != mesh.Ne:
        raise ValueError("matvec_far: X must have Ne rows")
    k = X.shape[1]
    Y = torch.empty((mesh.Ne, k), dtype=c128, device=X.device)
    for c0 in range(0, k, 8):
        kc = min(8, k - c0)
        Xc = X[:, c0:c0 + kc].contiguous()
        Yc = torch.empty((mesh.Ne, kc), dtype=c128, device=X.device)
        wb = L.lib().m25x_matvec_far_work_bytes(mesh.Nt, kc)
        work = torch.empty(wb, dtype=torch.uint8, device=X.device)
        L.call('m25x_matvec_far', ptr(Xc), kc, kc, C.byref(mesh.struct), C.byref(tabs.struct), float(w), ptr(Yc), kc,
               ptr(work), wb, stream())
        Y[:, c0:c0 + kc] = Yc
    return Y[:, 0] if one else Y


def near_coo(mesh, plan, M, w):
    ri = torch.empty(9 * plan.n, dtype=i32, device=device())
    ci = torch.empty(9 * plan.n, dtype=i32, device=device())
    vv = torch.empty(9 * plan.n, dtype=c128, device=device())
    L.call('m25x_near_coo', C.byref(mesh.struct), ptr(plan.P), ptr(plan.Q), ptr(M), plan.n, float(w), ptr(ri),
           ptr(ci), ptr(vv), stream())
    return ri, ci, vv


def near_pairs(cent, size, z, fac):
    """(P, Q) int32 device tensors of the near triangle pairs, sorted by (p, q)."""
    cent, size, z = on_dev(cent, f64), on_dev(size, f64), on_dev(z, f64)
    Nt = int(size.shape[0])
    counts = torch.empty(Nt + 1, dtype=torch.int64, device=device())
    n = C.c_long(0)
    L.call('m25x_near_pairs', ptr(cent), ptr(size), ptr(z), Nt, float(fac), ptr(counts), C.byref(n), None, None,
           stream())
    P = torch.empty(n.value, dtype=i32, device=device())
    Q = torch.empty(n.value, dtype=i32, device=device())
    if n.value:
        L.call('m25x_near_pairs', ptr(cent), ptr(size), ptr(z), Nt, float(fac), ptr(counts), C.byref(n), ptr(P),
               ptr(Q), stream())
    return P, Q


# ---- via fans and radial sums -------------------------------------------------------------------------------
def fan_q(mesh, tabs, F):
    F = on_dev(F, i32)
    out = torch.empty((mesh.Nt, len(F)), dtype=c128, device=device())
    L.call('m25x_fan_q', C.byref(mesh.struct), C.byref(tabs.struct), ptr(F), len(F), ptr(out), stream())
    return out


def radial_table_sum(px, py, W, vx, vy, table, dr):
    px, py, W = on_dev(px, f64), on_dev(py, f64), on_dev(W, f64)
    vx, vy, table = on_dev(vx, f64), on_dev(vy, f64), on_dev(table, c128)
    if px.dim() != 2 or py.shape != px.shape or W.shape != px.shape or vy.shape != vx.shape:
        raise ValueError("radial_table_sum: px, py, W must be (nt, nq), vx and vy (nv,)")
    nt, nq = px.shape
    out = torch.empty((nt, len(vx)), dtype=c128, device=device())
    L.call('m25x_radial_table_sum', ptr(px), ptr(py), ptr(W), nt, nq, ptr(vx), ptr(vy), len(vx), ptr(table),
           len(table), float(dr), ptr(out), stream())
    return out


# ---- Sommerfeld-Hankel sums ---------------------------------------------------------------------------------
def hankel_sum(rho, kr, W):
    """out[s, i] = sum_k J0(rho_i kr_k) W[s, k] (device complex128 (ns, nr))."""
    rho, kr = on_dev(rho, f64), on_dev(kr, c128)
    W = on_dev(W, c128)
    if W.dim() == 1:
        W = W[None]
    W = W.contiguous()
    if W.shape[1] != kr.shape[0]:
        raise ValueError("hankel: W must be (nspec, len(kr))")
    out = torch.empty((W.shape[0], len(rho)), dtype=c128, device=device())
    L.call('m25x_hankel', ptr(rho), len(rho), ptr(kr), len(kr), ptr(W), W.shape[0], ptr(out), stream())
    return out


def bessel_jn_complex(n, z):
    z = on_dev(z, c128)
    out = torch.empty_like(z)
    L.call('m25x_bessel_jn_complex', int(n), ptr(z), z.numel(), ptr(out), stream())
    return out


# ---- equilibration ------------------------------------------------------------------------------------------
def diag_scales(Z):
    d = torch.empty(Z.shape[0], dtype=f64, device=device())
    L.call('m25x_diag_scales', ptr(Z), Z.shape[0], Z.stride(0), ptr(d), stream())
    return d


def scale_c64(Z, d, Z64=None):
    """Z <- D Z D in place; its complex64 copy into Z64 when given."""
    if Z.dtype != c128 or Z.dim() != 2 or Z.shape[0] != Z.shape[1] or Z.stride(1) != 1:
        raise ValueError("scale_c64: Z square complex128, row-major")
    L.call('m25x_scale_c64', ptr(Z), Z.shape[0], Z.stride(0), ptr(d), ptr(Z64) if Z64 is not None else None,
           Z64.stride(0) if Z64 is not None else 0, stream())


# ---- LU and solves ------------------------------------------------------------------------------------------
class Factor:
    """An LU factor on the device: A = Z^T (column-major view of the row-major buffer) = P L U in place, ipiv
    0-based, perm its permutation; dtype complex128 or complex64."""

    def __init__(self, LU, ipiv, perm, info):
        self.LU, self.ipiv, self.perm, self.info = LU, ipiv, perm, info
        self.n = LU.shape[0]
        self.dtype = LU.dtype

    @property
    def nbytes(self):
        return self.LU.numel() * self.LU.element_size()


def getrf(Z, pivot='tournament'):
    """LU of A = Z^T in place of the row-major square device tensor Z (complex128 or complex64). pivot:
    'tournament' (CALU: a panel's pivots by a tree of GEPPs, no per-column grid barrier) or 'partial' (LAPACK's
    pivots exactly)."""
    if pivot not in ('tournament', 'partial'):
        raise ValueError(f"getrf: pivot {pivot!r}")
    if Z.dim() != 2 or Z.shape[0] != Z.shape[1] or Z.stride(1) != 1 or Z.dtype not in (c128, c64):
        raise ValueError("getrf: a square row-major complex tensor")
    n, dv = Z.shape[0], device()
    dt = L.C128 if Z.dtype == c128 else L.C64
    ipiv = torch.empty(n, dtype=i32, device=dv)
    info = torch.zeros(1, dtype=i32, device=dv)
    wb = L.lib().m25x_getrf_work_bytes(dt, n)
    work = torch.empty(wb, dtype=torch.uint8, device=dv)
    pv = L.PIVOT_TOURNAMENT if pivot == 'tournament' else L.PIVOT_PARTIAL
    perm = torch.empty(n, dtype=i32, device=dv)
    L.call('m25x_getrf', dt, pv, ptr(Z), n, Z.stride(0), ptr(ipiv), ptr(perm), ptr(info), ptr(work), wb, stream())
    return Factor(Z, ipiv, perm, info)


def getrs(F, B, trans=1):
    """X with Z X = B (trans 1) or Z^T X = B (trans 0), B (n,) or (n, k) complex (any dtype); the solve in the
    factor's precision, complex128 out. Columns are solved 8 at a time."""
    one = B.dim() == 1
    B = B.reshape(B.shape[0], -1)
    n, k = B.shape
    out = torch.empty((n, k), dtype=c128, device=B.device)
    dt = L.C128 if F.dtype == c128 else L.C64
    for c0 in range(0, k, 8):
        kc = min(8, k - c0)
        X = B[:, c0:c0 + kc].to(F.dtype, copy=True).contiguous()      # solved in place: never the caller's B
        wb = L.lib().m25x_getrs_work_bytes(dt, n, kc)
        work = torch.empty(wb, dtype=torch.uint8, device=B.device)
        L.call('m25x_getrs', dt, ptr(F.LU), n, F.LU.stride(0), ptr(F.ipiv), ptr(F.perm), ptr(X), kc, kc,
               int(trans), ptr(work), wb, stream())
        _solve_fault(B.device).bitwise_or_(work[wb - 256:wb - 252].view(torch.int32))   # the call's abort word
        out[:, c0:c0 + kc] = X.to(c128)
    return out[:, 0] if one else out


def getrs_noalloc(F, B, trans=1, status=None):
    """getrs through m25x_getrs_noalloc: the same solve with the library's own per-device workspace (nothing is
    allocated by the library; the caller's B is solved in place, 1 <= k <= 8 columns, the factor's dtype, n within
    the workspace's limit). status: a device int32 tensor that receives the wait word (a new one when None).
    Returns (X = B, status). Calls on one device must not overlap: they share the workspace."""
    if B.dtype != F.dtype or B.dim() != 2 or not B.is_contiguous():
        raise ValueError("getrs_noalloc: B must be a contiguous (n, k) tensor of the factor's dtype")
    if status is None:
        status = torch.zeros(1, dtype=torch.int32, device=B.device)
    dt = L.C128 if F.dtype == c128 else L.C64
    L.call('m25x_getrs_noalloc', dt, ptr(F.LU), F.n, F.LU.stride(0), ptr(F.ipiv), ptr(F.perm), ptr(B), B.shape[1],
           B.shape[1], int(trans), ptr(status), stream())
    return B, status


# ---- batches ------------------------------------------------------------------------------------------------
def _item_table(cls, items):
    """A ctypes array of items on the host and the same bytes in a device tensor (kept with the call's objects)."""
    arr = (cls * len(items))(*items)
    raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
    return arr, raw.to(device(), non_blocking=False)


class BatchFactors:
    """The LU factors of a batch of matrices (m25x_getrf_batched): factors[i] is a Factor of Zs[i], in place."""

    def __init__(self, Zs, factors, keep):
        self.Zs, self.factors, self._keep = Zs, factors, keep


def getrf_batched(Zs, graph=None):
    """Tournament LUs of the square row-major device tensors Zs (one dtype, any sizes) in one sequence of launches,
    in place; returns BatchFactors. The items go to the library largest first (its grids then shrink as the smaller
    ones finish). graph: a torch.cuda.CUDAGraph to capture the launches into instead of running them (replay it to
    factor; the same tensors are factored each replay)."""
    if not Zs:
        return BatchFactors([], [], None)
    dt = Zs[0].dtype
    if dt not in (c128, c64) or any(Z.dtype != dt or Z.dim() != 2 or Z.shape[0] != Z.shape[1] or Z.stride(1) != 1
                                    for Z in Zs):
        raise ValueError("getrf_batched: square row-major complex tensors of one dtype")
    dv, code = device(), (L.C128 if dt == c128 else L.C64)
    order = sorted(range(len(Zs)), key=lambda i: -Zs[i].shape[0])
    factors, items, keep = [None] * len(Zs), [], []
    for i in order:
        Z = Zs[i]
        n = Z.shape[0]
        ipiv = torch.empty(n, dtype=i32, device=dv)
        perm = torch.empty(n, dtype=i32, device=dv)
        info = torch.zeros(1, dtype=i32, device=dv)
        wb = L.lib().m25x_getrf_work_bytes(code, n)
        work = torch.empty(max(wb, 1), dtype=torch.uint8, device=dv)
        keep.append(work)
        factors[i] = Factor(Z, ipiv, perm, info)
        items.append(L.LuItem(Z.data_ptr(), n, Z.stride(0), ipiv.data_ptr(), perm.data_ptr(), info.data_ptr(),
                              work.data_ptr(), wb))
    host, dev_tab = _item_table(L.LuItem, items)
    keep += [host, dev_tab]
    if graph is not None:
        with torch.cuda.graph(graph):
            L.call('m25x_getrf_batched', code, len(items), host, ptr(dev_tab), stream())
    else:
        L.call('m25x_getrf_batched', code, len(items), host, ptr(dev_tab), stream())
    return BatchFactors(Zs, factors, keep)


class BatchSolve:
    """Solves with the factors of a batch, each item's right-hand sides a column block of one (n_max, k_total)
    tensor: the items' tables built once (m25x_getrs_batched)."""

    def __init__(self, factors, cols):
        """factors: Factor per item; cols: (first column, columns) of each item in the stacked right-hand sides."""
        self.factors, self.cols = factors, cols
        self.dtype = factors[0].dtype
        self.code = L.C128 if self.dtype == c128 else L.C64
        dv = device()
        self.work = [torch.empty(L.lib().m25x_getrs_work_bytes(self.code, F.n, 8), dtype=torch.uint8, device=dv)
                     for F in factors]
        self.status = torch.zeros(1, dtype=i32, device=dv)

    def __call__(self, R, trans=1, subset=None):
        """X (n_max, k) complex128 with F_i X_i = R_i for each item's column block of R (n_max, k); subset: for
        each item the (first column, columns) in R instead of self.cols (columns 0 skip the item)."""
        cols = subset or self.cols
        nmax, k = R.shape
        X = R.to(self.dtype, copy=True).contiguous()            # solved in place: never the caller's R
        items = []
        for F, w, (c0, kc) in zip(self.factors, self.work, cols):
            for p in range(0, kc, 8):
                q = min(8, kc - p)
                items.append(L.RsItem(F.LU.data_ptr(), F.n, F.LU.stride(0), F.perm.data_ptr(),
                                      X.data_ptr() + (c0 + p) * X.element_size(), k, q, 0, w.data_ptr(),
                                      w.numel()))
        if items:
            host, dev_tab = _item_table(L.RsItem, items)
            L.call('m25x_getrs_batched', self.code, len(items), host, ptr(dev_tab), int(trans), ptr(self.status),
                   stream())
            _solve_fault(X.device).bitwise_or_(self.status)
        return X.to(c128)


def zgemm_nk_batched(Zs, V, cols, out=None):
    """Y with Y_i = Z_i V_i for each item's column block (first column, columns) of V (n_max, k) complex128: one
    launch for all items (rows past an item's n are zero in Y)."""
    V = V.to(c128).contiguous()
    Y = torch.zeros_like(V) if out is None else out
    items = []
    for Z, (c0, kc) in zip(Zs, cols):
        for p in range(0, kc, 8):
            q = min(8, kc - p)
            items.append(L.MvItem(Z.data_ptr(), Z.shape[0], Z.stride(0), V.data_ptr() + (c0 + p) * 16,
                                  Y.data_ptr() + (c0 + p) * 16, V.shape[1], Y.shape[1], q, 0))
    if items:
        host, dev_tab = _item_table(L.MvItem, items)
        L.call('m25x_zgemm_nk_batched', len(items), host, ptr(dev_tab), stream())
    return Y


_FAULTS = {}


def _solve_fault(dev):
    f = _FAULTS.get(dev)
    if f is None:
        f = _FAULTS[dev] = torch.zeros(1, dtype=torch.int32, device=dev)
    return f


def solve_faults(dev=None):
    """Whether a getrs since the last call ran out of a wait bound between its workgroups (its result is then not
    the solution); clears the record. Reading it synchronises with the stream."""
    f = _FAULTS.get(torch.device(dev) if dev is not None else device())
    if f is None:
        return False
    bad = bool(f.item())
    f.zero_()
    return bad


KERNELS = ('fill_far', 'near_moments', 'near_scatter', 'tslu', 'apply_moves', 'l21', 'trsm', 'gemm', 'trsv',
           'zgemm_nk', 'cgs2', 'panel', 'laswp')


def timing(kernels):
    """Open a launch-timing record for the named kernels (KERNELS; 'all' for every one), clearing the last, or close
    it (kernels empty): each recorded launch is bracketed by two HIP events on the stream it runs on."""
    if kernels == 'all':
        kernels = KERNELS
    mask = 0
    for k in kernels:
        mask |= 1 << KERNELS.index(k)
    L.call('m25x_timing', mask)


def timing_records():
    """The launches recorded since timing(...): a dict of arrays kernel (names), dtype ('c128' / 'c64'), ms, and
    the bytes and flops each launch moves and does by algorithm (0 where the kernel has no such count)."""
    lib = L.lib()
    n = lib.m25x_timing_records(None, None, None, None, None, 0)
    if n < 0:
        raise RuntimeError(lib.m25x_last_error().decode())
    k, d = np.zeros(n, np.int32), np.zeros(n, np.int32)
    ms, by, fl = np.zeros(n, np.float32), np.zeros(n, np.float64), np.zeros(n, np.float64)
    if n:
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        got = lib.m25x_timing_records(p(k), p(d), p(ms), p(by), p(fl), n)
        if got < 0:
            raise RuntimeError(lib.m25x_last_error().decode())
    return dict(kernel=np.array([KERNELS[i] for i in k], dtype=object),
                dtype=np.where(d == L.C128, 'c128', 'c64'), ms=ms.astype(np.float64), bytes=by, flops=fl)


def zgemm_nk(Z, X):
    """Y = Z X with Z (n x n complex128, row-major) and X (n,) or (n, k); columns 8 at a time."""
    one = X.dim() == 1
    X = X.reshape(X.shape[0], -1).to(c128)
    n, k = X.shape
    Y = torch.empty((n, k), dtype=c128, device=X.device)
    for c0 in range(0, k, 8):
        kc = min(8, k - c0)
        Xc = X[:, c0:c0 + kc].contiguous()
        Yc = torch.empty((n, kc), dtype=c128, device=X.device)
        L.call('m25x_zgemm_nk', ptr(Z), n, Z.stride(0), ptr(Xc), kc, ptr(Yc), stream())
        Y[:, c0:c0 + kc] = Yc
    return Y[:, 0] if one else Y


_CGS2_WORK = {}


def cgs2(V, nv, W, hcol):
    """One Arnoldi step on the device: W (n x k complex128, overwritten) orthogonalised against V[:nv] by classical
    Gram-Schmidt applied twice; hcol[:nv + 1] (nv + 1, k) receives the coefficients and |W| (real); V[nv] = W / |W|.
    V: (m, n, k) complex128 contiguous, k <= 8. Calls on one device share a workspace: one stream at a time."""
    m, n, k = V.shape
    if not (V.is_contiguous() and W.is_contiguous() and hcol.is_contiguous()) or W.shape != (n, k) or nv >= m:
        raise ValueError("cgs2: V (m, n, k), W (n, k) contiguous, nv < m")
    dev = V.device
    work = _CGS2_WORK.get(dev)
    if work is None:
        work = _CGS2_WORK[dev] = torch.empty(L.lib().m25x_cgs2_work_bytes(), dtype=torch.uint8, device=dev)
    L.call('m25x_cgs2', ptr(V), n, k, nv, ptr(W), ptr(hcol), ptr(V[nv]), ptr(work), work.numel(), stream())

