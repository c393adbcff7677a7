"""The GPU LU and its solves against LAPACK (SciPy's lu_factor / lu_solve), the reference's factorisation
(src/mom25d/fast.py:318-358) and its GPU-backend test (tests/test_gpu.py:210-235 of the reference): the same
pivots, the factor in lu_factor's layout, and the solves of both transposes. The tournament-pivoting path against
its NumPy restatement (oracle/lu.py): the same interchanges, P A = L U, and the same solves.

Tolerances: complex128 - backward error 1e-13 and the factor within 1e-12 of LAPACK's; complex64 - the
reference's own bounds, |lu - ref| < 1e-4 max|ref| and solves within 1e-4 (test_gpu.py:231-235).
"""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import lu as olu
from tests.cases import seeded_matrix

pytestmark = pytest.mark.gpu


def _factor(dv, Z, pivot='partial'):
    Zd = torch.from_numpy(np.ascontiguousarray(Z)).to(dv.device())
    F = dv.getrf(Zd, pivot=pivot)
    torch.cuda.synchronize()
    return F


@pytest.mark.parametrize('n', [1, 7, 64, 129, 300, 517])
def test_getrf_c128_matches_lapack(gpu, n):
    dv = gpu
    Z = seeded_matrix(n)
    F = _factor(dv, Z)
    lu_ref, piv_ref = sla.lu_factor(Z.T)                   # the reference's factor of A = Z^T
    lu = F.LU.cpu().numpy().T                              # our row-major buffer is A column-major
    piv = F.ipiv.cpu().numpy()
    assert np.array_equal(piv, piv_ref)
    assert np.array_equal(F.perm.cpu().numpy(), olu.perm_of(piv))
    assert np.abs(lu - lu_ref).max() <= 1e-12 * np.abs(lu_ref).max()
    assert int(F.info.item()) == 0
    rng = np.random.default_rng(n)
    R = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    X = dv.getrs(F, torch.from_numpy(R).to(dv.device()), trans=1).cpu().numpy()
    Xr = sla.lu_solve((lu_ref, piv_ref), R, trans=1)      # Z X = R
    assert np.abs(X - Xr).max() <= 1e-11 * np.abs(Xr).max()
    assert np.abs(Z @ X - R).max() <= 1e-12 * np.abs(R).max() * np.abs(Z).max() * n ** 0.5
    X0 = dv.getrs(F, torch.from_numpy(R).to(dv.device()), trans=0).cpu().numpy()
    assert np.abs(Z.T @ X0 - R).max() <= 1e-12 * np.abs(R).max() * np.abs(Z).max() * n ** 0.5
    x1 = dv.getrs(F, torch.from_numpy(R[:, 0].copy()).to(dv.device()), trans=1).cpu().numpy()
    assert np.abs(x1 - Xr[:, 0]).max() <= 1e-11 * np.abs(Xr).max()


@pytest.mark.parametrize('n', [300, 1000])
def test_getrf_c64_matches_lapack(gpu, n):
    """complex64: the reference's bounds for its GPU LU (tests/test_gpu.py:219-235)."""
    dv = gpu
    Z = seeded_matrix(n, dtype=np.complex64)
    F = _factor(dv, Z)
    lu_ref, piv_ref = sla.lu_factor(Z.T.astype(complex))
    assert np.array_equal(F.ipiv.cpu().numpy(), piv_ref)
    lu = F.LU.cpu().numpy().T
    assert np.abs(lu - lu_ref).max() < 1e-4 * np.abs(lu_ref).max()
    rng = np.random.default_rng(3)
    R = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    x = np.linalg.solve(Z.astype(complex), R)
    X = dv.getrs(F, torch.from_numpy(R).to(dv.device()), trans=1).cpu().numpy()
    assert X.dtype == np.complex128
    assert np.abs(X - x).max() < 1e-4 * np.abs(x).max()


def test_getrf_selfcheck_matrix(gpu):
    """tools/gpu_selfcheck.py's matrix (seed 1, (randn + i randn) / sqrt(n) + 3 I): residual < 1e-4 (complex64)
    and ~1e-14 (complex128)."""
    dv = gpu
    n = 2048
    rng = np.random.default_rng(1)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64) / np.sqrt(n)
    A[np.diag_indices(n)] += 3
    b = rng.standard_normal((n, 2)).astype(complex)
    for pivot in ('partial', 'tournament'):
        for dt, tol in ((np.complex64, 1e-4), (np.complex128, 1e-12)):
            F = _factor(dv, A.astype(dt), pivot)
            x = dv.getrs(F, torch.from_numpy(b).to(dv.device()), trans=1).cpu().numpy()
            res = np.linalg.norm(A.astype(complex) @ x - b) / np.linalg.norm(b)
            assert res < tol, (pivot, dt, res)


def test_zgemm_nk(gpu):
    dv = gpu
    rng = np.random.default_rng(2)
    for n, k in ((1, 1), (300, 3), (1000, 8), (777, 11)):
        Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        X = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
        Y = dv.zgemm_nk(torch.from_numpy(Z).to(dv.device()), torch.from_numpy(X).to(dv.device())).cpu().numpy()
        assert np.abs(Y - Z @ X).max() <= 1e-13 * np.abs(Z).max() * np.abs(X).max() * n


def test_singular_pivot_reports_info(gpu):
    dv = gpu
    Z = seeded_matrix(100)
    Z[:, 5] = 0                                            # column 5 of Z = row 5 of A: zero from the start
    Z[5, :] = 0
    for pivot in ('partial', 'tournament'):
        F = _factor(dv, Z, pivot)
        assert int(F.info.item()) == 6, pivot             # LAPACK's info: 1 + the first zero pivot


def _check_tournament(dv, Z, F, RL, AR, tol_fac, tol_solve, full_ipiv):
    n = Z.shape[0]
    A = Z.T.astype(complex)
    ipiv = F.ipiv.cpu().numpy()
    assert np.all(ipiv >= np.arange(n)) and np.all(ipiv < n)
    if full_ipiv:
        LUo, ipo = olu.getrf_tournament(A, RL=RL, AR=AR)
        assert np.array_equal(ipiv, ipo)
        assert np.abs(F.LU.cpu().numpy().T - LUo).max() <= tol_fac * np.abs(LUo).max()
    else:                                                  # the first panel's election, from the same inputs
        win = olu.tslu_winners(A, 0, min(64, n), RL, AR)
        assert np.array_equal(ipiv[:len(win)], olu.interchanges(win, 0, np.arange(n)))
    assert np.array_equal(F.perm.cpu().numpy(), olu.perm_of(ipiv))
    lu = F.LU.cpu().numpy().T.astype(complex)
    Lf, Uf = np.tril(lu, -1) + np.eye(n), np.triu(lu)
    assert np.abs(A[olu.perm_of(ipiv)] - Lf @ Uf).max() <= tol_fac * np.abs(A).max()
    rng = np.random.default_rng(n)
    R = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    x = np.linalg.solve(Z.astype(complex), R)
    X = dv.getrs(F, torch.from_numpy(R).to(dv.device()), trans=1).cpu().numpy()
    assert np.abs(X - x).max() <= tol_solve * np.abs(x).max()
    X0 = dv.getrs(F, torch.from_numpy(R).to(dv.device()), trans=0).cpu().numpy()
    x0 = np.linalg.solve(A, R)
    assert np.abs(X0 - x0).max() <= tol_solve * np.abs(x0).max()


@pytest.mark.parametrize('n', [1, 7, 64, 65, 129, 300, 1000, 2000])
def test_getrf_c128_tournament_matches_oracle(gpu, n):
    """complex128 tournament pivoting (leaves of 256 rows, 4 children a node): the oracle's interchanges exactly, the
    factor within 1e-11, solves within 1e-10."""
    Z = seeded_matrix(n)
    F = _factor(gpu, Z, 'tournament')
    assert int(F.info.item()) == 0
    _check_tournament(gpu, Z, F, 256, 4, 1e-11, 1e-10, True)


@pytest.mark.parametrize('n', [300, 1000, 2000, 5000])
def test_getrf_c64_tournament(gpu, n):
    """complex64 tournament pivoting (leaves of 256 rows, 4 children a node; n = 5000 has nodes with a single
    child): the first panel's election as the oracle's, P A = L U to FP32 accuracy, solves within the reference's
    1e-4."""
    Z = seeded_matrix(n, dtype=np.complex64)
    F = _factor(gpu, Z, 'tournament')
    assert int(F.info.item()) == 0
    _check_tournament(gpu, Z, F, 256, 4, 5e-5, 1e-4, False)


@pytest.mark.parametrize('dt', [np.complex64, np.complex128])
def test_lookahead_factor_is_the_same_bit_for_bit(gpu, dt):
    """The tournament LU with look-ahead (the next panel's tournament on a second stream beside the rest of the
    update) and without it (the default) give the same factor, pivots and permutation bit for bit, from three
    panels with a one-column last one to many."""
    from mom25d_amd import _lib as L
    dv = gpu
    was = L.lib().m25x_lu_lookahead()
    for n in (129, 200, 777, 3001):
        Z = seeded_matrix(n, dtype=dt)
        outs = []
        for la in (1, 0):
            L.call('m25x_lu_set_lookahead', la)
            try:
                F = _factor(dv, Z, 'tournament')
                outs.append((F.LU.cpu(), F.ipiv.cpu(), F.perm.cpu(), int(F.info.item())))
            finally:
                L.call('m25x_lu_set_lookahead', was)
        (a, pa, qa, ia), (b, pb, qb, ib) = outs
        assert torch.equal(a, b) and torch.equal(pa, pb) and torch.equal(qa, qb) and ia == ib == 0, n


@pytest.mark.parametrize('dt', [np.complex64, np.complex128])
def test_getrs_right_hand_side_counts(gpu, dt):
    """The solves for 1 to 8 right-hand sides at sizes with and without a partial last 64-row block and with long
    chains of 64-row blocks (one workgroup each, every hand-off between them in play), both transposes, against
    NumPy's solve; the caller's B is left as it was, and no wait between workgroups ran out."""
    dv = gpu
    tol = 1e-4 if dt == np.complex64 else 1e-10
    for n in (1, 63, 64, 200, 777, 3001):
        Z = seeded_matrix(n, dtype=dt)
        F = _factor(dv, Z, 'tournament')
        Zc = Z.astype(complex)
        rng = np.random.default_rng(n)
        for k in (1, 3, 4, 5, 8):
            R = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
            Rd = torch.from_numpy(R).to(dv.device())
            R0 = Rd.clone()
            for trans in (1, 0):
                X = dv.getrs(F, Rd, trans=trans).cpu().numpy()
                assert torch.equal(Rd, R0), (n, k, trans)
                x = np.linalg.solve(Zc if trans else Zc.T, R)
                assert np.abs(X - x).max() <= tol * np.abs(x).max(), (n, k, trans)
    assert not dv.solve_faults()


def test_getrs_abort_word_covers_split_calls(gpu):
    """A complex128 solve of 8 right-hand sides runs as two parts of 4 inside one m25x_getrs call; the call's abort
    word sits at work + work_bytes - 256 for the work_bytes passed (here twice the minimum, the buffer pre-filled
    with 0xff bytes), is zeroed once by the call and still reads zero after both parts; the solution is right."""
    import ctypes as C
    from mom25d_amd import _lib as L
    dv = gpu
    n, k = 777, 8
    Z = seeded_matrix(n)
    F = _factor(dv, Z, 'tournament')
    rng = np.random.default_rng(5)
    R = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    X = torch.from_numpy(R.copy()).to(dv.device())
    wb = 2 * L.lib().m25x_getrs_work_bytes(L.C128, n, k) + 12
    work = torch.full((wb,), 255, dtype=torch.uint8, device=dv.device())
    L.call('m25x_getrs', L.C128, dv.ptr(F.LU), n, n, dv.ptr(F.ipiv), dv.ptr(F.perm), dv.ptr(X), k, k, 1,
           dv.ptr(work), wb, dv.stream())
    torch.cuda.synchronize()
    at = (wb - 256) & ~3
    assert int(work[at:at + 4].view(torch.int32).item()) == 0
    x = np.linalg.solve(Z, R)
    assert np.abs(X.cpu().numpy() - x).max() <= 1e-10 * np.abs(x).max()
    with pytest.raises(ValueError):                        # a workspace below the minimum is refused
        L.call('m25x_getrs', L.C128, dv.ptr(F.LU), n, n, dv.ptr(F.ipiv), dv.ptr(F.perm), dv.ptr(X), k, k, 1,
               dv.ptr(work), 1024, dv.stream())
