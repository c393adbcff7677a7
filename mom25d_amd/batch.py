"""Many frequency points - of one board or of many boards of uneven size - solved as one batch on the GPU.

No counterpart in the reference, which solves one point at a time (src/mom25d/run.py:55-64 calls
Simulation.sparams per frequency; src/mom25d/fast.py:704-850 fills and factors one Z). A batch:

  fill      every point's Z (FP64) on the device, equilibrated (D Z D), and its complex64 copy: first each point's
            host-side parts (tables, near moments, via blocks), then every Z queued with nothing waiting
  factor    all complex64 copies in one sequence of launches (m25x_getrf_batched: each panel step one launch per
            kernel for all points), optionally captured once in a HIP graph and replayed for the next batch of the
            same sizes
  refine    one GMRES over the ports of every point of similar size (groups within GROUP_SPREAD), side by side as
            the columns of one (n_max, k_total) block, zero-padded:
            the operator is block-diagonal - the FP64 products of all points in one launch (m25x_zgemm_nk_batched),
            the preconditioner all points' triangular solves in one launch per triangle (m25x_getrs_batched) - and
            a point whose refinement stalls stops alone
  S         per point; a point whose complex64 factor does not converge is solved again by its own board's
            single-point path (complex128 LU)

The same constants as the single-point mixed path (restart 30, stall (4, 0.1), converged at 100 x 1e-6), so each
point's answer is within the mixed path's bound of its double-precision one.
"""
import time

import numpy as np
import torch

from . import device as dv
from .fast import FRESH_STALL, MIXED_MAX_ITER, POLISH
from .linalg import gmres_batched

GROUP_SPREAD = 1.25      # points refined together: sizes within this factor of the group's smallest


class _Cols:
    """Which columns of the stacked right-hand sides belong to which point, and the per-point column blocks of the
    active subset that GMRES passes to the operators."""

    def __init__(self, ks):
        self.first = np.concatenate([[0], np.cumsum(ks)[:-1]]).astype(int)
        self.k = np.asarray(ks, int)
        self.owner = np.repeat(np.arange(len(ks)), ks)
        self.active = [(int(c), int(k)) for c, k in zip(self.first, self.k)]

    def track(self, act):
        """Active global columns (ascending) -> each point's (first position, count) among them."""
        own = self.owner[act]
        out = []
        for i in range(len(self.k)):
            pos = np.flatnonzero(own == i)
            out.append((int(pos[0]), len(pos)) if len(pos) else (0, 0))
        self.active = out


class SweepBatch:
    """Points [(sim, f), ...] solved together in mixed precision. sim: a Simulation (its mesh, near plan and tables
    stay resident on the device); one board may appear at many frequencies and many boards at once."""

    def __init__(self, points, graph=False):
        self.points = list(points)
        self.graph = graph
        self._graph = None
        self.stats = {}

    def run(self):
        """S of every point (a list of nports x nports arrays, in the order given)."""
        self.fill_factor()
        return self.refine()

    def tensors(self):
        """The device tensors the first stage leaves for the second (for a caller that runs the stages on two
        streams: record them on the second)."""
        st = self._state
        out = st['Zs'] + st['Z64s'] + st['ds'] + st['Bs']
        for f in st['F'].factors:
            out += [f.LU, f.ipiv, f.perm, f.info]
        return out + [k for k in st['F']._keep if isinstance(k, torch.Tensor)]

    def fill_factor(self):
        """Stage one, queued on the current stream: every point's Z filled on the device and equilibrated, its
        complex64 copy, and all copies factored by the batched LU."""
        t0 = time.perf_counter()
        pts = self.points
        Zs, Z64s, ds, Bs, ks = [], [], [], [], []
        # every point's host-side parts first (tables, moments, via blocks: these wait for the device), then every
        # Z, equilibration and copy queued with nothing waiting, so the fills run back to back on the device
        prep = [(sim.fast.prepare(f), torch.from_numpy(sim.excitation()).to(dv.device())) for sim, f in pts]
        for (sim, f), (st, E) in zip(pts, prep):
            Z = sim.fast.assemble(st)
            d = dv.diag_scales(Z)
            Z64 = torch.empty(Z.shape, dtype=dv.c64, device=Z.device)
            dv.scale_c64(Z, d, Z64)                          # Z <- D Z D (FP64, kept), Z64 its complex64 copy
            Zs.append(Z); Z64s.append(Z64); ds.append(d); Bs.append(d[:, None] * E); ks.append(E.shape[1])
        t1 = time.perf_counter()
        if self.graph:
            g = torch.cuda.CUDAGraph()
            F = dv.getrf_batched(Z64s, graph=g)
            g.replay()
            self._graph = (g, F)
        else:
            F = dv.getrf_batched(Z64s)
        self._state = dict(Zs=Zs, Z64s=Z64s, ds=ds, Bs=Bs, ks=ks, F=F, t0=t0, t1=t1)

    def refine(self):
        """Stage two, on the current stream (after the first stage's work): the refinement in size groups and S per
        point, a point that does not converge solved again by its board's single-point path."""
        st = self._state
        Zs, ds, Bs, ks, F, t0, t1 = st['Zs'], st['ds'], st['Bs'], st['ks'], st['F'], st['t0'], st['t1']
        pts = self.points
        torch.cuda.current_stream().synchronize()
        t2 = time.perf_counter()
        # one GMRES per group of points of similar size (the stacked block is padded to the group's largest n)
        order = np.argsort([Z.shape[0] for Z in Zs], kind='stable')
        groups, cur = [], [order[0]]
        for i in order[1:]:
            if Zs[i].shape[0] > GROUP_SPREAD * Zs[cur[0]].shape[0]:
                groups.append(cur)
                cur = []
            cur.append(i)
        groups.append(cur)
        I_of, ok, self.iters = {}, {}, [None] * len(pts)
        runs = [(grp,) + self._refine([Zs[i] for i in grp], [F.factors[i] for i in grp], [Bs[i] for i in grp],
                                      [ks[i] for i in grp]) for grp in groups]
        self._polish(runs, Zs, F.factors, Bs, ks)
        for grp, Y, okc, rel, it in runs:
            first = np.concatenate([[0], np.cumsum([ks[i] for i in grp])[:-1]])
            for j, i in enumerate(grp):
                ok[i] = bool(okc[first[j]:first[j] + ks[i]].all())
                I_of[i] = ds[i][:, None] * Y[:Zs[i].shape[0], first[j]:first[j] + ks[i]]
                self.iters[i] = it[first[j]:first[j] + ks[i]]
        torch.cuda.current_stream().synchronize()
        t3 = time.perf_counter()
        out, fallback = [], 0
        for i, (sim, f) in enumerate(pts):
            if ok[i]:
                I = I_of[i].cpu().numpy()
            else:
                fallback += 1
                I = sim.fast.solve_direct(f, sim.excitation(), precision='double')
            out.append(sim.sparams_from(f, I))
        if dv.solve_faults():
            raise RuntimeError("a batched triangular solve ran out of a wait bound: its results are not solutions")
        t4 = time.perf_counter()
        self.stats = dict(points=len(pts), fill_s=t1 - t0, factor_s=t2 - t1, refine_s=t3 - t2, s_s=t4 - t3,
                          total_s=t4 - t0, fallback=fallback, groups=len(groups),
                          iterations_max=int(max(max(x) for x in self.iters)))
        self._state = None
        return out

    def _refine(self, Zs, factors, Bs, ks):
        """GMRES over the given points' columns side by side: (Y (n_max, k_total), accepted per column, relative
        residual per column, iterations per column)."""
        nmax = max(Z.shape[0] for Z in Zs)
        cols = _Cols(ks)
        B = torch.zeros((nmax, int(cols.k.sum())), dtype=dv.c128, device=Zs[0].device)
        for i, b in enumerate(Bs):
            B[:b.shape[0], cols.first[i]:cols.first[i] + ks[i]] = b
        solve = dv.BatchSolve(factors, cols.active)
        A = lambda V: dv.zgemm_nk_batched(Zs, V, cols.active)
        M = lambda R: solve(R, trans=1, subset=cols.active)
        tol = 1e-6
        Y, it = gmres_batched(A, B, M=M, rtol=tol, restart=MIXED_MAX_ITER, maxiter=1, stall=FRESH_STALL,
                              track=cols.track, stall_each=True)
        cols.track(np.arange(B.shape[1]))
        R = B - A(Y)
        rel = (torch.linalg.vector_norm(R, dim=0) /
               torch.clamp(torch.linalg.vector_norm(B, dim=0), min=1e-300)).cpu().numpy()
        # answers accepted within 100 x tol (the single path's rule); the others are left to the single-point
        # fallback, and the accepted ones still above POLISH x tol to _polish
        okc = np.isfinite(rel) & (rel <= 100 * tol)
        return Y, okc, rel, it

    def _polish(self, runs, Zs, factors, Bs, ks):
        """The single path's polish for every group at once: one more GMRES cycle, from its answer, for each
        accepted column still above POLISH x tol, the columns of all groups side by side (one batched solve and
        product an iteration, not one a group). runs: [(group, Y, accepted, residual, iterations)], updated."""
        tol = 1e-6
        sel = []                                       # (run, point of the group, column of the point, block column)
        for r, (grp, Y, okc, rel, it) in enumerate(runs):
            first = np.concatenate([[0], np.cumsum([ks[i] for i in grp])[:-1]])
            for j, i in enumerate(grp):
                for c in range(ks[i]):
                    b = first[j] + c
                    if okc[b] and rel[b] > POLISH * tol:
                        sel.append((r, j, c, b))
        if not sel:
            return
        pts = sorted({(r, j) for r, j, _, _ in sel})
        idx = {p: n for n, p in enumerate(pts)}
        gi = [runs[r][0][j] for r, j in pts]          # the points' indices in the batch
        cnt = np.bincount([idx[(r, j)] for r, j, _, _ in sel], minlength=len(pts))
        sel.sort(key=lambda e: (idx[(e[0], e[1])], e[2]))
        nmax = max(Zs[i].shape[0] for i in gi)
        dev = Zs[gi[0]].device
        B = torch.zeros((nmax, len(sel)), dtype=dv.c128, device=dev)
        X0 = torch.zeros_like(B)
        for q, (r, j, c, b) in enumerate(sel):
            n = Zs[runs[r][0][j]].shape[0]
            B[:n, q] = Bs[runs[r][0][j]][:, c]
            X0[:n, q] = runs[r][1][:n, b]
        pc = _Cols(cnt)
        solve = dv.BatchSolve([factors[i] for i in gi], pc.active)
        A = lambda V: dv.zgemm_nk_batched([Zs[i] for i in gi], V, pc.active)
        M = lambda R: solve(R, trans=1, subset=pc.active)
        Y2, it2 = gmres_batched(A, B, M=M, x0=X0, rtol=tol, restart=MIXED_MAX_ITER, maxiter=1, track=pc.track,
                                stall_each=True)
        for q, (r, j, c, b) in enumerate(sel):
            n = Zs[runs[r][0][j]].shape[0]
            runs[r][1][:n, b] = Y2[:n, q]
            runs[r][4][b] += it2[q]
