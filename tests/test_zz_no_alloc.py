"""m25x_getrs_noalloc (the solve whose workspace is the library's own device array) called through the C ABI against
m25x_getrs on the same factor: bit-identical solutions for both element types, on two calls in a row (the second
with a smaller system, so the array holds the first call's values when it starts), the wait word zero; and an n
beyond the array's size rejected with ValueError before anything runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, k, dtype, seed):
    import torch
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) + 3 * np.eye(n)
    B = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    return torch.from_numpy(A).cuda().to(dtype), torch.from_numpy(B).cuda().to(dtype)


def _solve_with_buffer(F, B, trans):
    """m25x_getrs with a caller-allocated workspace (the existing entry point), through the C ABI."""
    import torch
    from mom25d_amd import _lib as L, device as dv
    dt = L.C128 if F.dtype == torch.complex128 else L.C64
    n, k = B.shape
    X = B.clone()
    wb = L.lib().m25x_getrs_work_bytes(dt, n, k)
    work = torch.empty(wb, dtype=torch.uint8, device=B.device)
    L.call('m25x_getrs', dt, dv.ptr(F.LU), n, F.LU.stride(0), dv.ptr(F.ipiv), dv.ptr(F.perm), dv.ptr(X), k, k,
           int(trans), dv.ptr(work), wb, dv.stream())
    return X, int(work[(wb - 256) & ~3:((wb - 256) & ~3) + 4].view(torch.int32).item())


def _solve_noalloc(F, B, trans):
    import torch
    from mom25d_amd import _lib as L, device as dv
    dt = L.C128 if F.dtype == torch.complex128 else L.C64
    n, k = B.shape
    X = B.clone()
    status = torch.full((1,), 7, dtype=torch.int32, device=B.device)
    L.call('m25x_getrs_noalloc', dt, dv.ptr(F.LU), n, F.LU.stride(0), dv.ptr(F.ipiv), dv.ptr(F.perm), dv.ptr(X), k,
           k, int(trans), dv.ptr(status), dv.stream())
    return X, int(status.item())


@pytest.mark.parametrize('dtype', ['c128', 'c64'])
def test_noalloc_solve_matches_getrs(gpu, dtype):
    import torch
    from mom25d_amd import device as dv
    dt = torch.complex128 if dtype == 'c128' else torch.complex64
    for n, k, trans, seed in ((1500, 8, 1, 0), (700, 3, 0, 1)):          # the second call is smaller
        print(f"noalloc: {dtype} n={n} k={k} trans={trans}", flush=True)
        A, B = _system(n, k, dt, seed)
        F = dv.getrf(A)
        X0, w0 = _solve_with_buffer(F, B, trans)
        X1, w1 = _solve_noalloc(F, B, trans)
        torch.cuda.synchronize()
        assert w0 == 0 and w1 == 0
        assert torch.equal(torch.view_as_real(X0), torch.view_as_real(X1)), f"{dtype} n={n}: not bit-identical"
        # and it solves the system: Z X = B for trans 1, Z^T X = B for trans 0 (Z the row-major matrix factored)
        Zh = _system(n, k, torch.complex128, seed)[0].cpu().numpy()
        Xh = X1.to(torch.complex128).cpu().numpy()
        Bh = _system(n, k, torch.complex128, seed)[1].cpu().numpy()
        M = Zh if trans == 1 else Zh.T
        tol = 1e-10 if dtype == 'c128' else 1e-4
        assert np.abs(M @ Xh - Bh).max() <= tol * np.abs(Bh).max()


def test_noalloc_rejects_oversize(gpu):
    import torch
    from mom25d_amd import _lib as L, device as dv
    n = 50000                                  # complex128 at 8 columns: beyond the 16 MiB array
    need = L.lib().m25x_getrs_work_bytes(L.C128, n, 8) - 256
    assert need > L.GETRS_SCRATCH_BYTES
    tiny = torch.zeros(8, dtype=torch.complex128, device='cuda')
    status = torch.zeros(1, dtype=torch.int32, device='cuda')
    piv = torch.zeros(1, dtype=torch.int32, device='cuda')
    with pytest.raises(ValueError, match='workspace'):
        L.call('m25x_getrs_noalloc', L.C128, dv.ptr(tiny), n, n, dv.ptr(piv), None, dv.ptr(tiny), 8, 8, 1,
               dv.ptr(status), dv.stream())
