"""The refinement's device kernels against NumPy: the fused Arnoldi step (classical Gram-Schmidt twice) of the
batched GMRES (src/mom25d/fast.py:1437-1500 restated in oracle/gmres.py with modified Gram-Schmidt; classical twice
gives the same basis to rounding), checked against a direct NumPy evaluation of the same formulas in FP64."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k,nv", [(1, 1, 1), (100, 3, 1), (777, 4, 5), (2000, 5, 12), (7799, 8, 30), (3000, 4, 33), (7799, 4, 60)])
def test_cgs2_matches_numpy(gpu, n, k, nv):
    dv = gpu
    rng = np.random.default_rng(n + k + nv)
    V = np.zeros((nv + 1, n, k), complex)
    for c in range(k):                                     # an orthonormal basis per system
        Q, _ = np.linalg.qr(rng.standard_normal((n, nv)) + 1j * rng.standard_normal((n, nv)))
        V[:nv, :, c] = Q.T
    W = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    h, Wr = np.zeros((nv + 1, k), complex), W.copy()
    for c in range(k):
        Vc = V[:nv, :, c]
        h1 = Vc.conj() @ Wr[:, c]
        w1 = Wr[:, c] - Vc.T @ h1
        h2 = Vc.conj() @ w1
        w2 = w1 - Vc.T @ h2
        h[:nv, c], h[nv, c], Wr[:, c] = h1 + h2, np.linalg.norm(w2), w2
    Vd, Wd = torch.from_numpy(V).to(dv.device()), torch.from_numpy(W).to(dv.device())
    hb = torch.zeros((nv + 1, k), dtype=torch.complex128, device=dv.device())
    dv.cgs2(Vd, nv, Wd, hb)
    hg = hb.cpu().numpy()
    scale = np.abs(W).max() * np.sqrt(n)
    assert np.abs(hg - h).max() <= 1e-12 * scale
    assert np.abs(hg[nv].imag).max() == 0
    vn = Vd[nv].cpu().numpy()
    ref = Wr / np.where(h[nv].real > 0, h[nv].real, 1.0)[None, :]
    assert np.abs(vn - ref).max() <= 1e-11
    if n > nv:                                             # the new vector is orthonormal to the basis
        for c in range(k):
            assert abs(np.linalg.norm(vn[:, c]) - 1) < 1e-12
            assert np.abs(V[:nv, :, c].conj() @ vn[:, c]).max() < 1e-12


def _seeded_system():
    """The reference's seeded complex system (tests/test_solver.py:375-398), as the oracle's GMRES pin uses it."""
    rng = np.random.default_rng(1)
    n = 200
    ev = (1 + rng.random(n)) * np.exp(1j * rng.uniform(-1.2, 1.2, n))
    Qm = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    A = Qm @ np.diag(ev) @ Qm.conj().T + 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    good = np.linalg.inv(A + 1e-3 * rng.standard_normal((n, n)))
    bad = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return A, B, good, bad


@pytest.mark.parametrize("case", ["plain", "restarts", "good", "bad_stall", "bad_full"])
def test_device_gmres_matches_oracle(gpu, case):
    """mom25d_amd.linalg.gmres_batched (device basis, fused CGS2, the next iteration queued ahead) against the
    oracle's restatement: the same iteration counts per column and the same solution to the tolerance."""
    from oracle.gmres import gmres_batched as ref_gmres
    from mom25d_amd.linalg import gmres_batched
    A, B, good, bad = _seeded_system()
    kw = dict(plain=dict(rtol=1e-10, restart=60, maxiter=5), restarts=dict(rtol=1e-10, restart=7, maxiter=20),
              good=dict(M=good, rtol=1e-8, restart=30, maxiter=1, stall=(4, 0.1)),
              bad_stall=dict(M=bad, rtol=1e-8, restart=30, maxiter=3, stall=(4, 0.1)),
              bad_full=dict(M=bad, rtol=1e-8, restart=30, maxiter=1))[case]
    Mh = kw.pop('M', None)
    Xr, itr = ref_gmres(lambda V: A @ V, B, M=(lambda R: Mh @ R) if Mh is not None else None, **kw)
    dev = gpu.device()
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    Md = torch.from_numpy(Mh).to(dev) if Mh is not None else None
    Xd, itd = gmres_batched(lambda V: Ad @ V, Bd, M=(lambda R: Md @ R) if Md is not None else None, **kw)
    Xd = Xd.cpu().numpy()
    assert np.isfinite(Xd).all()
    assert max(abs(a - b) for a, b in zip(itd, itr)) <= 1, (itd, itr)   # (rounding may move a threshold crossing)
    if case in ("plain", "restarts", "good"):
        res = np.linalg.norm(A @ Xd - B, axis=0) / np.linalg.norm(B, axis=0)
        assert res.max() < 10 * kw['rtol']
        assert np.abs(Xd - Xr).max() <= 100 * kw['rtol'] * np.abs(Xr).max()
    else:
        assert itd == itr


def test_gmres_rate_horizon(gpu):
    """stall (j0, ratio, horizon, level): the unpreconditioned seeded system falls at ~0.71 an iteration, so from
    2 j0 = 8 iterations on it would need ~50 in all to reach 1e-8, beyond a horizon of 30: it stops at the first
    check instead of running to 30 (the ratio test alone, at 1.0 |b|, never fires). A converging run (the good
    preconditioner) is the same with and without the horizon, bit for bit."""
    from mom25d_amd.linalg import gmres_batched
    A, B, good, _ = _seeded_system()
    dev = gpu.device()
    Ad, Bd, Gd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev), torch.from_numpy(good).to(dev)
    _, it_full = gmres_batched(lambda V: Ad @ V, Bd, rtol=1e-8, restart=30, maxiter=1, stall=(4, 1.0))
    _, it_h = gmres_batched(lambda V: Ad @ V, Bd, rtol=1e-8, restart=30, maxiter=1, stall=(4, 1.0, 30, 1.0))
    assert it_full == [30, 30] and all(8 <= i <= 10 for i in it_h), (it_full, it_h)
    X0, i0 = gmres_batched(lambda V: Ad @ V, Bd, M=lambda R: Gd @ R, rtol=1e-8, restart=30, maxiter=1, stall=(4, 0.1))
    X1, i1 = gmres_batched(lambda V: Ad @ V, Bd, M=lambda R: Gd @ R, rtol=1e-8, restart=30, maxiter=1,
                           stall=(4, 0.1, 30, 100.0))
    assert i0 == i1 and torch.equal(torch.view_as_real(X0), torch.view_as_real(X1))


def test_gmres_guess(gpu):
    """guess=True: a column whose guess leaves a residual above |b|, or not finite, starts from zero (the same
    iterations and the same answer bit for bit as without a guess), a column with a good guess converges in fewer
    iterations."""
    from mom25d_amd.linalg import gmres_batched
    A, B, good, _ = _seeded_system()
    dev = gpu.device()
    Ad, Bd, Gd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev), torch.from_numpy(good).to(dev)
    kw = dict(M=lambda R: Gd @ R, rtol=1e-10, restart=30, maxiter=1)
    X0, i0 = gmres_batched(lambda V: Ad @ V, Bd, **kw)
    Xs = torch.from_numpy(np.linalg.solve(A, B)).to(dev)
    bad = 1e3 * Xs.clone()
    near = Xs * (1 + 1e-4)
    x0 = torch.stack([bad[:, 0], near[:, 1]], 1)
    X1, i1 = gmres_batched(lambda V: Ad @ V, Bd, x0=x0, guess=True, **kw)
    assert i1[0] == i0[0] and torch.equal(torch.view_as_real(X1[:, 0]), torch.view_as_real(X0[:, 0]))
    assert i1[1] < i0[1], (i0, i1)
    res = np.linalg.norm(A @ X1.cpu().numpy() - B, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() < 1e-9
    x0[:, 0] = float('nan')                              # a guess that is not finite starts from zero too
    X2, i2 = gmres_batched(lambda V: Ad @ V, Bd, x0=x0, guess=True, **kw)
    assert i2 == i1 and torch.equal(torch.view_as_real(X2), torch.view_as_real(X1))
