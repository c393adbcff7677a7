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
