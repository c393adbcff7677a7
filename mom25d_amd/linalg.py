"""Batched restarted GMRES on the device (src/mom25d/fast.py:1437-1500).

Right-preconditioned, k systems in lockstep: every operator / preconditioner call is one (N x k_active) product on
the GPU (the FP64 Z product of m25x_zgemm_nk, the complex64 factor's solves of m25x_getrs). The Krylov basis stays
on the device and is orthogonalised by classical Gram-Schmidt applied twice (the reference uses modified
Gram-Schmidt: the same basis to rounding), fused in one call (m25x_cgs2) for up to 8 systems; the small Hessenberg
columns come to the host once per iteration for the conjugated Givens rotations, the convergence test and the
stall cut-off, as the reference computes them. The next iteration's device work (preconditioner, product,
orthogonalisation) does not depend on those, so it is queued before the host waits for the current column
(through pinned buffers and an event), unless the residual's recent rate predicts convergence at this iteration;
an iteration queued past the end is discarded. Beyond 8 systems (a batch of points side by side) the columns are
independent systems, and once half of them are done they leave the block in the middle of a cycle, so that a
few slow ones do not carry the finished ones through every later product and solve.
"""
import numpy as np
import torch

from . import device as dv


_PINNED = {}


def _pinned(shape):
    """Two pinned host buffers of complex128 (shape), kept per shape."""
    b = _PINNED.get(shape)
    if b is None:
        b = _PINNED[shape] = [torch.empty(shape, dtype=torch.complex128, pin_memory=True) for _ in range(2)]
    return b


def _cnorm(X):
    """Column 2-norms of a device complex matrix (host float64 array)."""
    return torch.linalg.vector_norm(X, dim=0).cpu().numpy()


def gmres_batched(A, B, M=None, x0=None, rtol=1e-6, restart=30, maxiter=10, stall=None, track=None,
                  stall_each=False, guess=False):
    """(X, iterations per column) for A X = B; A, M: callables on device (N x k) complex128 tensors; B device.
    track: called with the active columns (indices into B's columns, ascending) before A and M see a new set of
    them, for operators that act on each column differently (a batch of systems side by side). stall_each: a
    column that stalls stops alone (the others go on), as for independent systems; otherwise the first stall stops
    every column, as the reference's lockstep ports do. stall (j0, ratio): a column whose residual is still above
    ratio x |b| after j0 iterations has stalled; (j0, ratio, horizon, level): also one that, from 2 j0 iterations
    on, would need more than horizon iterations in all to reach level x rtol at its rate over the last four
    (a column whose residual no longer falls needs infinitely many). guess: x0 is only a guess (the solution at
    a nearby frequency, say): a column whose guess leaves a residual above |b| starts from zero instead."""
    B = B.to(torch.complex128)
    N, k = B.shape
    dev = B.device
    X = torch.zeros((N, k), dtype=torch.complex128, device=dev) if x0 is None else x0.clone()
    M = M or (lambda R: R)
    bn = _cnorm(B)
    bn[bn == 0] = 1.0
    iters, act = np.zeros(k, int), np.arange(k)
    started = x0 is not None
    for _ in range(maxiter):
        if track is not None:
            track(act)
        Xa = X[:, torch.from_numpy(act).to(dev)]
        R = B[:, torch.from_numpy(act).to(dev)] - (A(Xa) if started else 0)
        beta = _cnorm(R)
        if guess and x0 is not None:
            guess = False
            worse = np.flatnonzero(~(beta <= bn[act]))                # (a guess that is not finite, too)
            if len(worse):                                           # those columns start from zero
                wc = torch.from_numpy(act[worse]).to(dev)
                X[:, wc] = 0
                R[:, torch.from_numpy(worse).to(dev)] = B[:, wc]
                beta[worse] = bn[act[worse]]
        done = beta <= rtol * bn[act]
        keep = ~done
        act, beta = act[keep], beta[keep]
        if not len(act):
            break
        R = R[:, torch.from_numpy(np.flatnonzero(keep)).to(dev)]
        if track is not None:
            track(act)
        ka = len(act)
        V = torch.zeros((restart + 1, N, ka), dtype=torch.complex128, device=dev)
        Zs = torch.zeros((restart, N, ka), dtype=torch.complex128, device=dev)
        H = np.zeros((restart + 1, restart, ka), complex)
        g = np.zeros((restart + 1, ka), complex)
        cs, sn = np.zeros((restart, ka), complex), np.zeros((restart, ka), complex)
        g[0] = beta
        V[0] = R / torch.from_numpy(beta).to(dev)[None, :]
        live, last, stalled = np.ones(ka, bool), np.full(ka, restart - 1), False
        rn = np.zeros((restart + 1, ka))                             # residual estimates, for the prediction
        rn[0] = beta
        hbuf = torch.empty((restart + 1, ka), dtype=torch.complex128, device=dev)
        # classical Gram-Schmidt twice (as stable as modified Gram-Schmidt): one fused device call for up to 8
        # systems and 63 basis vectors (then pipelined one iteration ahead), batched tensor products beyond
        fused = ka <= 8 and restart < 64
        if fused:
            pin, evs, queued = _pinned((restart + 1, ka)), [torch.cuda.Event(), torch.cuda.Event()], -1

            def enqueue(i):
                Zs[i] = M(V[i])
                dv.cgs2(V, i + 1, A(Zs[i]).contiguous(), hbuf)
                pin[i % 2][:i + 2].copy_(hbuf[:i + 2], non_blocking=True)
                evs[i % 2].record()
        for j in range(restart):
            if fused:
                if queued < j:
                    enqueue(j)
                    queued = j
                # queue the next iteration unless this one is predicted to converge (the last step's ratio)
                ratio = rn[j] / np.maximum(rn[j - 1], 1e-300) if j else np.ones(ka)
                soon = (rn[j] * ratio <= 2 * rtol * bn[act])[live].all()
                if j + 1 < restart and not soon:
                    enqueue(j + 1)
                    queued = j + 1
                evs[j % 2].synchronize()
                hc = pin[j % 2][:j + 2].numpy().copy()
            else:
                Zs[j] = M(V[j])
                Wv = A(Zs[j]).contiguous()
                Vj = V[:j + 1]
                h1 = (Vj.conj() * Wv[None]).sum(1)
                Wv = Wv - (Vj * h1[:, None, :]).sum(0)
                h2 = (Vj.conj() * Wv[None]).sum(1)
                Wv = Wv - (Vj * h2[:, None, :]).sum(0)
                hn = torch.linalg.vector_norm(Wv, dim=0)
                V[j + 1] = Wv / torch.where(hn > 0, hn, torch.ones_like(hn))
                hc = torch.cat([h1 + h2, hn[None].to(torch.complex128)], 0).cpu().numpy()
            H[:j + 2, j] = hc
            for i in range(j):                                       # previous Givens rotations
                H[i, j], H[i + 1, j] = (cs[i] * H[i, j] + sn[i] * H[i + 1, j],
                                        -np.conj(sn[i]) * H[i, j] + np.conj(cs[i]) * H[i + 1, j])
            den = np.sqrt(np.abs(H[j, j]) ** 2 + np.abs(H[j + 1, j]) ** 2)
            den[den == 0] = 1.0
            cs[j], sn[j] = H[j, j].conj() / den, H[j + 1, j].conj() / den
            H[j, j], H[j + 1, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j], 0
            g[j + 1], g[j] = -np.conj(sn[j]) * g[j], cs[j] * g[j]
            rn[j + 1] = np.abs(g[j + 1])
            iters[act[live]] += 1
            conv = live & (np.abs(g[j + 1]) <= rtol * bn[act])
            last[conv] = j
            live &= ~conv
            if not live.any():
                break
            if stall is not None and j + 1 >= stall[0]:
                bad = live & (np.abs(g[j + 1]) > stall[1] * bn[act])
                if len(stall) > 2 and j + 1 >= 2 * stall[0]:
                    rate = (rn[j + 1] / np.maximum(rn[j - 3], 1e-300)) ** 0.25
                    with np.errstate(divide='ignore', invalid='ignore'):
                        more = np.where(rate < 1, np.log(stall[3] * rtol * bn[act] / np.maximum(rn[j + 1], 1e-300)) /
                                        np.log(np.minimum(rate, 1 - 1e-12)), np.inf)
                    bad |= live & (j + 1 + more > stall[2])
                if stall_each:
                    last[bad] = j
                    live &= ~bad
                    stalled = stalled or bool(bad.any())
                    if not live.any():
                        break
                elif bad.any():
                    last[live], stalled = j, True
                    break
            if not fused and live.sum() <= ka // 2:
                # half the block's columns are done: they leave it (their updates now), and the rest go on alone;
                # each column's iterations are its own, so theirs do not change, but the products, solves and
                # orthogonalisation of the columns that are done stop
                for c in np.flatnonzero(~live):
                    m = last[c] + 1
                    if m:
                        y = np.linalg.solve(np.triu(H[:m, :m, c]), g[:m, c])
                        X[:, act[c]] += Zs[:m, :, c].T @ torch.from_numpy(y).to(dev)
                kn = np.flatnonzero(live)
                kt = torch.from_numpy(kn).to(dev)
                V, Zs = V[:, :, kt].contiguous(), Zs[:, :, kt].contiguous()
                H, g, cs, sn, rn = H[:, :, kn], g[:, kn], cs[:, kn], sn[:, kn], rn[:, kn]
                last, live, act, ka = last[kn], live[kn], act[kn], len(kn)
                if track is not None:
                    track(act)
        for c in range(ka):
            m = last[c] + 1
            if not m:
                continue
            y = np.linalg.solve(np.triu(H[:m, :m, c]), g[:m, c])
            X[:, act[c]] += Zs[:m, :, c].T @ torch.from_numpy(y).to(dev)
        started = True
        if stalled:
            break
    return X, iters.tolist()
