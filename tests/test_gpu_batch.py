"""Batches on the GPU: the batched LU, solves and products against the one-system calls on the same inputs (the same
kernels, so bit for bit), the batched LU captured in a HIP graph and replayed (every replay against the eager
result, with eager work between the replays), and a sweep batch of one board at several frequencies and of boards
of uneven size against each point's own double-precision solve (|dS| <= 1e-4, the mixed path's bound)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (700, 64, 1300, 200, 1)


def _mats(dtype, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    out = []
    for n in SIZES:
        A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) + 2 * np.eye(n)
        out.append(torch.from_numpy(A).cuda().to(dtype))
    return out


@pytest.mark.parametrize('dt', ['c128', 'c64'])
def test_batched_lu_is_the_single_lu(gpu, dt):
    import torch
    from mom25d_amd import device as dv
    dtype = torch.complex128 if dt == 'c128' else torch.complex64
    one = [dv.getrf(A.clone()) for A in _mats(dtype)]
    F = dv.getrf_batched([A.clone() for A in _mats(dtype)])
    torch.cuda.synchronize()
    for a, b, n in zip(one, F.factors, SIZES):
        assert torch.equal(torch.view_as_real(a.LU), torch.view_as_real(b.LU)), n
        assert torch.equal(a.ipiv, b.ipiv) and torch.equal(a.perm, b.perm) and int(b.info.item()) == 0, n


@pytest.mark.parametrize('dt', ['c128', 'c64'])
def test_batched_solves_and_products_are_the_single_ones(gpu, dt):
    import torch
    from mom25d_amd import device as dv
    dtype = torch.complex128 if dt == 'c128' else torch.complex64
    F = dv.getrf_batched(_mats(dtype))
    rng = np.random.default_rng(4)
    ks = [3, 1, 4, 2, 1]
    nmax = max(SIZES)
    R = torch.zeros((nmax, sum(ks)), dtype=torch.complex128, device='cuda')
    cols, c = [], 0
    for n, k in zip(SIZES, ks):
        R[:n, c:c + k] = torch.from_numpy(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
        cols.append((c, k))
        c += k
    for trans in (1, 0):
        X = dv.BatchSolve(F.factors, cols)(R, trans=trans)
        for f, n, (c0, k) in zip(F.factors, SIZES, cols):
            x1 = dv.getrs(f, R[:n, c0:c0 + k], trans=trans)
            assert torch.equal(torch.view_as_real(X[:n, c0:c0 + k]), torch.view_as_real(x1)), (n, trans)
    assert not dv.solve_faults()
    Zs = [A.to(torch.complex128) for A in _mats(dtype, seed=9)]
    Y = dv.zgemm_nk_batched(Zs, R, cols)
    for Z, n, (c0, k) in zip(Zs, SIZES, cols):
        y1 = dv.zgemm_nk(Z, R[:n, c0:c0 + k])
        assert torch.equal(torch.view_as_real(Y[:n, c0:c0 + k]), torch.view_as_real(y1)), n
        assert not Y[n:, c0:c0 + k].abs().any()


def test_batched_lu_graph_replays(gpu):
    """The batched LU captured once in a HIP graph and replayed on fresh inputs three times, the inputs restored by
    eager copies between the replays: every replay equals the eager batched factor."""
    import torch
    from mom25d_amd import device as dv
    src = _mats(torch.complex64, seed=2)
    eager = dv.getrf_batched([A.clone() for A in src])
    work = [A.clone() for A in src]
    g = torch.cuda.CUDAGraph()
    F = dv.getrf_batched(work, graph=g)
    for rep in range(3):
        for w, a in zip(work, src):
            w.copy_(a)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b, n in zip(eager.factors, F.factors, SIZES):
            assert torch.equal(torch.view_as_real(a.LU), torch.view_as_real(b.LU)), (rep, n)
            assert torch.equal(a.perm, b.perm), (rep, n)


def test_sweep_batch_matches_single_points(gpu):
    """A batch of the coupled-line board (N ~ 3k) at three frequencies and two one-layer trace boards at two each:
    every point's S within 1e-4 of its own double-precision solve."""
    from mom25d_amd import boards
    from mom25d_amd.batch import SweepBatch
    from mom25d_amd.sim import Simulation

    def make(board, *a):
        st, sh, po, vi, kw = board(*a)
        return Simulation(st, sh, po, vi, precision='mixed', verbose=False, **kw)
    cl = make(boards.coupled_lines)
    tb = [make(boards.trace_board, s, 0.3e-3, 4, 3) for s in (5, 6)]
    pts = [(cl, f) for f in (1.5e9, 4e9, 9e9)] + [(b, f) for b in tb for f in (2e9, 6e9)]
    batch = SweepBatch(pts)
    S = batch.run()
    assert batch.stats['points'] == len(pts)
    for (sim, f), s in zip(pts, S):
        sd = sim.fast.solve_direct(f, sim.excitation(), precision='double')
        ref = sim.sparams_from(f, sd)
        assert np.abs(s - ref).max() <= 1e-4, (f, np.abs(s - ref).max())
