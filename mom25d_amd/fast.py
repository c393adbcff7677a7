"""The direct solver of one frequency on the GPU (src/mom25d/fast.py: FastOperator.setup / solve_direct /
_solve_kept / _refine / _reusable, 597-626, 704-850, 974-1084).

Per frequency, everything on the device:
  tables   the Green's tables of every interface pair (multilayer.green_tables)
  fill     Z (N x N complex128): far field, near-field correction (pairs and their analytic moments resident
           across frequencies), via rows / columns and block, lumped elements
  solve    'double': Z equilibrated (D Z D, d = 1/sqrt|Z_ii|), complex128 LU, solves
           'mixed' : the complex64 copy of the equilibrated Z factored (trailing updates on the FP64 matrix
                     cores, one rounding to FP32 an entry), GMRES refinement
                     against the FP64 Z (restart 30, stall (4, 0.1) for a fresh factor, converged at 100 x 1e-6);
                     a kept factor of a nearby frequency tried first (its reach learnt); where the complex64
                     factor does not converge, complex128 LU, and double precision up to MIXED_RETRY x f
The currents come back to the host (N x ports); everything else stays on the GPU.
"""
import time

import numpy as np
import torch

from . import device as dv
from .linalg import gmres_batched

HYBRID_RETRY = 1.25
MIXED_RETRY = HYBRID_RETRY
MIXED_MAX_ITER = 30
POLISH = 10.0
# a fresh factor's refinement stops (the point goes to the complex128 path) when its residual is above 0.1 |b| after
# 4 iterations, or when from 8 iterations on its rate over the last four would not reach the acceptance level
# (100 x the tolerance) within the MIXED_MAX_ITER iterations it may take: such a factor does not converge, and
# iterating on to the limit only delays the complex128 factor that replaces it
FRESH_STALL = (4, 0.1, MIXED_MAX_ITER, 100.0)
REUSE_ITER = (8, 60)
REUSE_KEEP = 4
# A complex128 fall-back factor is kept for the points after it rounded to complex64 ('c64', the default) or as
# itself ('c128'). Rounding the finished factor perturbs it by one FP32 rounding an entry, far less than a complex64
# LU's own error: on the 8k-unknown benchmark board the 1 GHz factor, rounded, preconditions 1.15-6 GHz in the same
# GMRES iterations (5 to 35 a port) as the complex128 one, with triangular solves 28 % faster and half the memory.
KEEP_DOUBLE = 'c64'


class FastOperator:
    """Builds Z on the GPU for a Simulation and solves it."""

    def __init__(self, sim, near_fac=None):
        self.sim = sim
        self.near_fac = near_fac or sim.near_fac
        self.mixed_tol = 1e-6
        self._mixed_failed = 0.0
        self._kept = []                       # [(f, Factor, dsc (device), seconds a fresh one took)]
        self._last_x = None                   # the currents (device) of the last point solved
        self._t_iter = None
        self._reach = np.log(2.0)
        self.iters = []
        self.mesh = dv.DevMesh(sim.g, sim.q)
        with sim.log.timer('near pairs', sync=True):
            z = sim.st.z[sim.tri_layer]
            P, Q = dv.near_pairs(sim.rwg.cent, sim.tri_size, z, self.near_fac)
            self.plan = dv.NearPlan(self.mesh, P, Q)
        sim.log.note(f"fast solver: {self.plan.n} near triangle pairs ({self.plan.n / max(1, len(sim.tri_size)):.0f} "
                     f"per triangle), {self.plan.nentry} near entries")
        F = sim.F
        self._fan = None
        if len(F):
            dev = dv.device()
            pos = -np.ones(len(sim.tri_size), np.int64)
            pos[F] = np.arange(len(F))
            Pn, Qn = self.plan.P.long().cpu().numpy(), self.plan.Q.long().cpu().numpy()
            sel = np.flatnonzero(pos[Qn] >= 0)
            self._fan = (torch.from_numpy(sel).to(dev), torch.from_numpy(Pn[sel]).to(dev),
                         torch.from_numpy(pos[Qn[sel]]).to(dev))

    # ---- set-up and fill ------------------------------------------------------------------------------------
    def setup(self, f):
        self.f, self.w = f, 2 * np.pi * f
        self.sim.green_tables(f)
        self.tabs = self.sim.dev_tables

    def prepare(self, f):
        """The per-frequency parts of Z that involve the host, computed before any N x N work is queued: the
        Green's tables, the near pairs' moments, the via blocks B and C (their small dense parts are assembled on
        the host, which waits for the device there) and the lumped entries, uploaded. Returns the point's state for
        assemble(); holding it keeps its tables alive, so many points can be prepared before any is assembled."""
        sim, w = self.sim, 2 * np.pi * f
        self.setup(f)
        tabs, dev = self.tabs, dv.device()
        with sim.log.timer('near-field moments'):
            M, qc = dv.near_moments(self.mesh, tabs, self.plan, w)
        B = C = lum = None
        if sim.Nv:
            with sim.log.timer('via kernels', f"{sim.Nv} chains"):
                QF = dv.fan_q(self.mesh, tabs, sim.F)
                if self._fan is not None:
                    sel, rows, cols = self._fan
                    QF[rows, cols] += qc[sel]
                B, C = sim._via_blocks(f, QF)
        le = sim.lumped_entries(f)
        if le is not None:
            r, c, v = (torch.from_numpy(x).to(dev) for x in le)
            lum = (r, c, v.to(dv.c128))
        return dict(f=f, w=w, tabs=tabs, M=M, B=B, C=C, lumped=lum)

    def assemble(self, st, dtype=dv.c128):
        """Z (device, N x N, not equilibrated) of a prepared point: complex128, or complex64 for the lean path
        (each entry computed in FP64 and rounded once; the near correction added in FP32). Only device work is
        queued (nothing waits for the device). The parts the matrix-free operator needs are kept for this
        frequency: the near entries' values, the via blocks B and C and the lumped entries."""
        sim, w = self.sim, st['w']
        Ne, Nv = sim.Ne, sim.Nv
        N = Ne + Nv
        self.f, self.w, self.tabs = st['f'], w, st['tabs']
        with sim.log.timer('direct fill', f"N={N}, {'complex128' if dtype == dv.c128 else 'complex64'}, GPU"):
            Z = torch.empty((N, N), dtype=dtype, device=dv.device())
            dv.fill_far(Z, self.mesh, self.tabs, w)
        with sim.log.timer('near-field correction'):
            dv.near_scatter(Z, self.mesh, self.plan, st['M'], w)
        self._op = dict(f=st['f'], w=w, vals=dv.near_values(self.mesh, self.plan, st['M'], w) if dtype == dv.c64 else None,
                        B=st['B'], C=st['C'], lumped=st['lumped'], M=st['M'])
        if Nv:
            B, C = st['B'], st['C']
            Z[:Ne, Ne:] = B
            Z[Ne:, :Ne] = B.T
            Z[Ne:, Ne:] = C
        if st['lumped'] is not None:
            r, c, v = st['lumped']
            Z.index_put_((r, c), v.to(dtype), accumulate=True)
        return Z

    def fill(self, f, dtype=dv.c128):
        """Z (device, N x N, not equilibrated) at f: prepare(f) then assemble (see both)."""
        return self.assemble(self.prepare(f), dtype)

    def matvec_exact(self, X):
        """Z X in FP64 without Z (the reference's matvec_exact, fast.py:571-586) at the frequency of the last fill:
        the far field triangle-aggregated (m25x_matvec_far), the near correction as a sparse matrix, the via rows
        and columns and the lumped elements. X: device (N,) or (N, k)."""
        op, Ne = self._op, self.sim.Ne
        one = X.dim() == 1
        X = X.reshape(X.shape[0], -1).to(dv.c128)
        if op.get('vals') is None:
            op['vals'] = dv.near_values(self.mesh, self.plan, op['M'], op['w'])
        Y = torch.empty_like(X)
        Xe = X[:Ne].contiguous()
        Ye = dv.matvec_far(Xe, self.mesh, self.tabs, op['w'])
        dv.csr_matvec(self.plan, op['vals'], Xe, Ye)
        Y[:Ne] = Ye
        if self.sim.Nv:
            B, C, Xv = op['B'], op['C'], X[Ne:]
            Y[:Ne] += B @ Xv
            Y[Ne:] = B.T @ Xe + C @ Xv
        if op['lumped'] is not None:
            r, c, v = op['lumped']
            Y.index_add_(0, r, v[:, None] * X[c])
        return Y[:, 0] if one else Y

    # ---- solve ---------------------------------------------------------------------------------------------
    def _refine(self, A, B, M, restart, maxiter, stall=None, x0=None):
        """GMRES to mixed_tol: (Y, converged, largest relative residual) (fast.py:658-663). Whether the answer is
        accepted (a residual within 100 x mixed_tol) decides the path as in the reference; an accepted answer whose
        residual is still above POLISH x mixed_tol gets one more restart cycle from itself (S is then within ~3e-5 of
        the double-precision one rather than ~3e-4). x0: a guess (a column it does not help starts from zero)."""
        Y, self.iters = gmres_batched(A, B, M=M, x0=x0, rtol=self.mixed_tol, restart=restart, maxiter=maxiter,
                                      stall=stall, guess=True)
        rel = self._residual(A, B, Y)
        ok = bool(np.isfinite(rel).all() and (rel <= 100 * self.mixed_tol).all())
        if ok and len(rel) and rel.max() > POLISH * self.mixed_tol:
            Y, more = gmres_batched(A, B, M=M, x0=Y, rtol=self.mixed_tol, restart=restart, maxiter=1)
            self.iters = [a + b for a, b in zip(self.iters, more)]
            rel = self._residual(A, B, Y)
        return Y, ok, float(np.max(rel)) if len(rel) else 0.0

    def _guess(self, dsc, B):
        """GMRES's first guess at a reused factor: the last solved point's currents in this point's scaling
        (D^-1 X); along a sweep they are closer to this point's currents than zero is (GMRES still starts a column
        from zero where the guess is worse). Extrapolating the last two points linearly in frequency, or taking
        the better of the two guesses column by column, measured no better on the benchmark sweep."""
        x = self._last_x
        if x is None or x.shape != B.shape:
            return None
        return x / dsc[:, None]

    def _remember(self, X):
        self._last_x = X
        return X

    @staticmethod
    def _residual(A, B, Y):
        return (torch.linalg.vector_norm(B - A(Y), dim=0) /
                torch.clamp(torch.linalg.vector_norm(B, dim=0), min=1e-300)).cpu().numpy()

    def _reusable(self, f):
        if self.sim.reuse == 'off' or not self._kept:
            return None
        k = min(self._kept, key=lambda k: abs(np.log(f / k[0])))
        return k if abs(np.log(f / k[0])) <= self._reach else None

    def _reach_learn(self, d, ok, iters, cap):
        if ok and max(iters) <= cap // 2:
            self._reach = max(self._reach, 1.25 * d)
        elif not ok:
            self._reach = min(self._reach, 0.8 * d)

    def _drop_kept(self):
        lf = np.log([k[0] for k in self._kept])
        gap = [min((abs(lf[i] - y) for j, y in enumerate(lf) if j != i), default=0.0) for i in range(len(lf))]
        del self._kept[int(np.argmin(gap))]

    def solve_direct(self, f, E, precision='double'):
        """Currents (host N x k complex128) for excitation columns E at f (fast.py:704-850)."""
        sim = self.sim
        near_failed = precision == 'mixed' and f <= MIXED_RETRY * self._mixed_failed
        t_fresh = time.time()
        if precision == 'mixed' and self.lean and not near_failed:
            X = self._solve_lean(f, E, t_fresh)
        else:
            X = self._solve_dense(f, E, precision, near_failed, t_fresh)
        if dv.solve_faults(dv.device()):
            raise RuntimeError(f"a triangular solve at {f / 1e9:.4g} GHz ran out of a wait bound between its "
                               "workgroups: its result is not the solution")
        return X.cpu().numpy()

    @property
    def lean(self):
        """The lean mixed path (the reference's streamed solve, fast.py:1120-1224, in this build's form): only the
        complex64 Z is stored (8 N^2 bytes instead of 24 N^2), and the refinement multiplies by the FP64 Z without
        forming it (matvec_exact). sim.stream 'on' asks for it; 'auto' takes it where the memory plan chose it
        (the hybrid path's FP64 Z does not fit the device)."""
        mode = getattr(self.sim, 'stream', 'off')
        return mode == 'on' or (mode == 'auto' and getattr(self.sim, 'plan', {}).get('path') == 'gpu-stream')

    def _solve_lean(self, f, E, t_fresh):
        Z64 = self.fill(f, dv.c64)
        dev = Z64.device
        dsc = 1.0 / torch.sqrt(Z64.diagonal().abs().to(dv.f64) + 1e-300)
        d32 = dsc.to(torch.float32)
        Z64.mul_(d32[:, None]).mul_(d32[None, :])               # D Z D in complex64, in place
        B = dsc[:, None] * torch.from_numpy(np.asarray(E, complex)).to(dev)
        A = lambda V: dsc[:, None] * self.matvec_exact(dsc[:, None] * V)
        return self._solve_kept(f, None, dsc, B, t_fresh, fresh='mixed', A=A, Z64=Z64)

    def _solve_dense(self, f, E, precision, near_failed, t_fresh):
        sim = self.sim
        Z = self.fill(f)
        dev = Z.device
        Ed = torch.from_numpy(np.asarray(E, complex)).to(dev)
        dsc = dv.diag_scales(Z)
        B = dsc[:, None] * Ed
        if precision != 'mixed':
            with sim.log.timer('LU factor', f"N={Z.shape[0]}, complex128, GPU"):
                dv.scale_c64(Z, dsc)
                F = dv.getrf(Z)
            self._check_info(F)
            with sim.log.timer('LU solve'):
                X = dsc[:, None] * dv.getrs(F, B, trans=1)
        else:
            X = self._solve_kept(f, Z, dsc, B, t_fresh, fresh='double' if near_failed else 'mixed')
        return X

    def _check_info(self, F):
        info = int(F.info.item())
        if info < 0:
            raise RuntimeError("the LU's panel ran out of a wait bound between its workgroups: the factor is not valid")
        if info:
            self.sim.log.note(f"WARNING the LU met an exactly zero pivot at {info - 1}: the matrix is singular")

    def _solve_kept(self, f, Z, dsc, B, t_fresh, fresh='mixed', A=None, Z64=None):
        """Mixed precision with factors kept on the GPU (fast.py:996-1084): a kept factor of a nearby frequency
        first; else a fresh complex64 one (fresh='mixed') or, within MIXED_RETRY of a frequency where complex64 did not
        converge, a fresh complex128 one (fresh='double'). A complex128 factor is kept too (rounded to complex64, see
        KEEP_DOUBLE), so the points after a failed one are preconditioned by it instead of each taking its own
        complex128 LU (the reference keeps complex64 factors only)."""
        sim = self.sim
        n = len(dsc)
        if A is None:                                           # the dense path: Z (FP64) kept for the products
            Z64 = torch.empty((n, n), dtype=dv.c64, device=Z.device) if fresh == 'mixed' else None
            dv.scale_c64(Z, dsc, Z64)                           # Z <- D Z D (kept, FP64), and its complex64 copy
            A = lambda V: dv.zgemm_nk(Z, V)
        k = self._reusable(f)
        if k is not None:
            f0, F0, d0, tf0 = k
            r = (d0 / dsc)[:, None]
            cap = int(np.clip(0.8 * tf0 / max(self._t_iter or 1e-3, 1e-6), *REUSE_ITER))
            with sim.log.timer('refinement (GMRES, reused factorisation)', f"the one of {f0 / 1e9:.4g} GHz, at most {cap}"):
                t = time.time()
                Y, ok, rel = self._refine(A, B, lambda R: r * dv.getrs(F0, r * R, trans=1), cap, 1,
                                          x0=self._guess(dsc, B))
                self._t_iter = (time.time() - t) / max(1, max(self.iters, default=1))
            self._reach_learn(abs(np.log(f / f0)), ok, self.iters, cap)
            if ok:
                sim.log.note(f"reused LU ({f0 / 1e9:.4g} GHz): GMRES iterations per port {self.iters}")
                del Z64
                return self._remember(dsc[:, None] * Y)
            sim.log.note(f"the LU of {f0 / 1e9:.4g} GHz did not precondition {f / 1e9:.4g} GHz in {cap} iterations "
                         f"(residual {rel:.1e}) - a new one")
        if fresh == 'double':
            return self._fresh_double(f, Z, dsc, B, t_fresh)
        with sim.log.timer('LU factor (GPU)', f"N={n}, complex64"):
            F = dv.getrf(Z64)
        self._check_info(F)
        with sim.log.timer('refinement (GMRES, dense double)'):
            t = time.time()
            Y, ok, rel = self._refine(A, B, lambda R: dv.getrs(F, R, trans=1), MIXED_MAX_ITER, 1, FRESH_STALL)
            self._t_iter = (time.time() - t) / max(1, max(self.iters, default=1))
        sim.log.note(f"mixed precision (GPU): GMRES iterations per port {self.iters}")
        if ok:
            if sim.reuse != 'off':
                self._kept.append((f, F, dsc, time.time() - t_fresh))
                while len(self._kept) > REUSE_KEEP:
                    self._drop_kept()
            return self._remember(dsc[:, None] * Y)
        del F, Z64
        sim.log.note(f"WARNING mixed precision did not converge at {f / 1e9:.4g} GHz (residual {rel:.1e}) - solving "
                     f"in double precision up to {MIXED_RETRY * f / 1e9:.4g} GHz")
        self._mixed_failed = max(self._mixed_failed, f)
        if Z is None:                                           # the lean path: the FP64 Z filled for this point
            Z = self.fill(f)
            dv.scale_c64(Z, dsc)
        return self._fresh_double(f, Z, dsc, B, t_fresh)

    def _fresh_double(self, f, Z, dsc, B, t_fresh):
        """The complex128 LU of the equilibrated Z and its solve; the factor is kept for the nearby points."""
        sim = self.sim
        with sim.log.timer('LU factor', f"N={Z.shape[0]}, complex128, GPU"):
            Fd = dv.getrf(Z)                                     # in place: the factor owns Z's storage from here
        self._check_info(Fd)
        X = self._remember(dsc[:, None] * dv.getrs(Fd, B, trans=1))
        if sim.reuse != 'off':
            if KEEP_DOUBLE == 'c64':
                Fd = dv.Factor(Fd.LU.to(dv.c64), Fd.ipiv, Fd.perm, Fd.info)
            self._kept.append((f, Fd, dsc, time.time() - t_fresh))
            while len(self._kept) > REUSE_KEEP:
                self._drop_kept()
        return X
