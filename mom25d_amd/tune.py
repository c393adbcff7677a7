"""The LU's tile autotuner: times the trailing update's K-stage depth choices (m25x_lu_set_gemm_stage: 8, 16 or 32
columns of L21 / rows of U12 staged in LDS between two barriers) on the device and keeps the fastest for the rest of
the process. No counterpart in the reference (its LU is LAPACK's or a vendor library's).

The choice does not change any result: every stage depth accumulates the same products in the same order (the
matrix cores' accumulators persist across stages), so factors are bit-identical whichever is chosen; only the
number of barriers and the LDS footprint per workgroup differ.
"""
import numpy as np
import torch

from . import _lib as L
from . import device as dv

CHOICES = (8, 16, 32)
_CHOSEN = {}


def set_stage(ks):
    L.call('m25x_lu_set_gemm_stage', int(ks))


def stage():
    return int(L.lib().m25x_lu_gemm_stage())


def tune_lu(sizes, dtype=dv.c64, reps=2, seed=0):
    """Times the batched LU of random matrices of the given sizes (a batch like the one to be solved) with each
    stage depth, sets the fastest and returns the record {'choice', 'ms': {ks: ms}, 'sizes', 'dtype'}. A record for
    the same (sizes, dtype) is reused."""
    key = (tuple(int(n) for n in sizes), str(dtype))
    if key in _CHOSEN:
        rec = _CHOSEN[key]
        set_stage(rec['choice'])
        return rec
    rng = np.random.default_rng(seed)
    dev = dv.device()
    src = []
    for n in sizes:
        A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) + 2 * np.eye(n)
        src.append(torch.from_numpy(A).to(dev, dtype))
    s = torch.cuda.current_stream()
    times = {}
    for ks in CHOICES:
        set_stage(ks)
        best = float('inf')
        for r in range(reps + 1):
            work = [A.clone() for A in src]
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            dv.getrf_batched(work)
            b.record(s)
            torch.cuda.synchronize()
            if r:                                   # the first run of each choice loads its kernels
                best = min(best, a.elapsed_time(b))
        times[ks] = round(best, 3)
    choice = min(times, key=times.get)
    set_stage(choice)
    rec = dict(choice=choice, ms=times, sizes=list(key[0]), dtype=str(dtype))
    _CHOSEN[key] = rec
    return rec
