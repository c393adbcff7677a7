"""gmres_batched restated (src/mom25d/fast.py:1437-1500). TEST INFRASTRUCTURE ONLY.

Right-preconditioned restarted GMRES for k systems in lockstep, modified Gram-Schmidt, conjugated Givens
rotations, converged columns dropping out, an optional stall cut-off (iterations, residual)."""
import numpy as np


def gmres_batched(A, B, M=None, x0=None, rtol=1e-6, restart=30, maxiter=10, stall=None):
    B = np.asarray(B, complex)
    N, k = B.shape
    X = np.zeros((N, k), complex) if x0 is None else np.array(x0, complex)
    M = M or (lambda R: R)
    bn = np.linalg.norm(B, axis=0)
    bn[bn == 0] = 1.0
    iters, act = np.zeros(k, int), np.arange(k)
    for _ in range(maxiter):
        R = B[:, act] - (A(X[:, act]) if np.any(X[:, act]) else 0)
        beta = np.linalg.norm(R, axis=0)
        done = beta <= rtol * bn[act]
        act, R, beta = act[~done], R[:, ~done], beta[~done]
        if not len(act):
            break
        ka = len(act)
        V, Z = np.zeros((restart + 1, N, ka), complex), np.zeros((restart, N, ka), complex)
        H, g = np.zeros((restart + 1, restart, ka), complex), np.zeros((restart + 1, ka), complex)
        cs, sn = np.zeros((restart, ka), complex), np.zeros((restart, ka), complex)
        g[0], V[0] = beta, R / beta
        live, last, stalled = np.ones(ka, bool), np.full(ka, restart - 1), False
        for j in range(restart):
            Z[j] = M(V[j])
            Wv = A(Z[j])
            for i in range(j + 1):
                H[i, j] = hij = np.einsum('nk,nk->k', V[i].conj(), Wv)
                Wv = Wv - V[i] * hij
            H[j + 1, j] = hn = np.linalg.norm(Wv, axis=0)
            V[j + 1] = Wv / np.where(hn > 0, hn, 1)
            for i in range(j):
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
            y = np.linalg.solve(np.triu(H[:m, :m, c]), g[:m, c]) if m else np.zeros(0)
            X[:, act[c]] += Z[:m, :, c].T @ y
        if stalled:
            break
    return X, iters.tolist()
