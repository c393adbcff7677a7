"""The HIP backend against the reference's backend contract and its GPU test (tests/test_gpu.py:210-235 of the
reference): factor of an equilibrated complex64 Z, solves Z^-1 R in complex128, download() in lu_factor's layout
(partial pivoting: LAPACK's pivots and |lu - ref| < 1e-4 max|ref|), the admission figures."""
import numpy as np
import pytest
import scipy.linalg as sla

from mom25d_amd import gpu as G
from tests.cases import seeded_matrix


@pytest.mark.gpu
def test_hip_backend_matches_lapack(gpu):
    Z = seeded_matrix(300, dtype=np.complex64)
    b = G.HipBackend(pivot='partial')
    assert b.name == 'hip' and b.hip and b.api == 'ROCm/HIP' and b.inplace
    b.factor(Z)
    lu, piv = b.download()
    lu_ref, piv_ref = sla.lu_factor(Z.T.astype(complex))
    assert lu.flags.f_contiguous and lu.shape == (300, 300)
    assert np.array_equal(piv, piv_ref)
    assert np.abs(lu - lu_ref).max() < 1e-4 * np.abs(lu_ref).max()
    rng = np.random.default_rng(3)
    R = rng.standard_normal((300, 2)) + 1j * rng.standard_normal((300, 2))
    X = b.solve(R)
    assert X.dtype == np.complex128 and X.shape == (300, 2)
    x = np.linalg.solve(Z.astype(complex), R)
    assert np.abs(X - x).max() < 1e-4 * np.abs(x).max()
    assert b.solve(R[:, 0]).shape == (300,)
    assert b.fits(1000) and not b.fits(10 ** 6)
    assert b.vram_needed(1000) == 8.0e6 + 0.4e9
    b.release()
    assert b.lu is None and b.piv is None


@pytest.mark.gpu
def test_hip_backend_tournament_solves(gpu):
    Z = seeded_matrix(1000, dtype=np.complex64)
    b = G.HipBackend()
    b.factor(Z)
    lu, piv = b.download()
    n = Z.shape[0]
    L, U = np.tril(lu, -1) + np.eye(n), np.triu(lu)
    perm = np.arange(n)
    for i, p in enumerate(piv):
        perm[i], perm[p] = perm[p], perm[i]
    A = Z.T.astype(complex)
    assert np.abs(A[perm] - L @ U).max() < 1e-4 * np.abs(A).max()
    R = np.ones((n, 1), complex)
    assert np.abs(b.solve(R) - np.linalg.solve(Z.astype(complex), R)).max() < 1e-4


@pytest.mark.gpu
def test_hip_backend_singular_raises(gpu):
    Z = seeded_matrix(100, dtype=np.complex64)
    Z[:, 5] = 0
    Z[5, :] = 0
    with pytest.raises(FloatingPointError):
        G.HipBackend().factor(Z)


def test_registry_off_and_unknown(monkeypatch):
    monkeypatch.delenv('MOM25D_GPU', raising=False)
    assert G.get_backend('off') is None
    assert G.get_backend('nope') is None and 'unknown GPU backend' in G.get_backend.why[0]
    monkeypatch.setenv('MOM25D_GPU', 'off')
    assert G.get_backend('hip') is None


def test_registry_without_gpu_says_why(monkeypatch):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    monkeypatch.delenv('MOM25D_GPU', raising=False)
    G._cache.clear()
    assert G.get_backend('auto') is None
    assert any('hip' in w for w in G.get_backend.why)
    assert G.describe().startswith("GPU: none")
