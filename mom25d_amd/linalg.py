"""Batched restarted GMRES on the device (src/mom25d/fast.py:1437-1500).

Right-preconditioned, k systems in lockstep: every operator / preconditioner call is one (N x k_active) product on
the GPU (the FP64 Z product of m25x_zgemm_nk, the complex64 factor's solves of m25x_getrs). The Krylov basis stays
on the device and is orthogonalised by classical Gram-Schmidt applied twice (the reference uses modified
Gram-Schmidt: the same basis to rounding), fused in one call (m25x_cgs2) for up to 8 systems; the small Hessenberg
columns come to the host once per iteration for the conjugated Givens rotations, the convergence test and the
stall cut-off, as the reference computes them.
"""
import numpy as np
import torch

from . import device as dv


def _cnorm(X):
    """Column 2-norms of a device complex matrix (host float64 array)."""
    return torch.linalg.vector_norm(X, dim=0).cpu().numpy()


def gmres_batched(A, B, M=None, x0=None, rtol=1e-6, restart=30, maxiter=10, stall=None):
    """(X, iterations per column) for A X = B; A, M: callables on device (N x k) complex128 tensors; B device."""
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
        Xa = X[:, torch.from_numpy(act).to(dev)]
        R = B[:, torch.from_numpy(act).to(dev)] - (A(Xa) if started else 0)
        beta = _cnorm(R)
        done = beta <= rtol * bn[act]
        keep = ~done
        act, beta = act[keep], beta[keep]
        if not len(act):
            break
        R = R[:, torch.from_numpy(np.flatnonzero(keep)).to(dev)]
        ka = len(act)
        V = torch.zeros((restart + 1, N, ka), dtype=torch.complex128, device=dev)
        Zs = torch.zeros((restart, N, ka), dtype=torch.complex128, device=dev)
        H = np.zeros((restart + 1, restart, ka), complex)
        g = np.zeros((restart + 1, ka), complex)
        cs, sn = np.zeros((restart, ka), complex), np.zeros((restart, ka), complex)
        g[0] = beta
        V[0] = R / torch.from_numpy(beta).to(dev)[None, :]
        live, last, stalled = np.ones(ka, bool), np.full(ka, restart - 1), False
        hbuf = torch.empty((restart + 1, ka), dtype=torch.complex128, device=dev)
        for j in range(restart):
            Zs[j] = M(V[j])
            Wv = A(Zs[j]).contiguous()
            # classical Gram-Schmidt twice (as stable as modified Gram-Schmidt): one fused device call for up to 8
            # systems and 63 basis vectors, batched tensor products beyond
            if ka <= 8 and j + 1 < 64:
                dv.cgs2(V, j + 1, Wv, hbuf)
                hc = hbuf[:j + 2].cpu().numpy()
            else:
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
            iters[act[live]] += 1
            conv = live & (np.abs(g[j + 1]) <= rtol * bn[act])
            last[conv] = j
            live &= ~conv
            if not live.any():
                break
            if stall is not None and j + 1 >= stall[0] and (np.abs(g[j + 1])[live] > stall[1] * bn[act][live]).any():
                last[live], stalled = j, True
                break
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
