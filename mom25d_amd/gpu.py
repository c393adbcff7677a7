"""GPU LU backend with the reference's backend contract (src/mom25d/gpu.py:28-48, 118-146), on the HIP kernels of
libm25x.so: the complex64 (or complex128) LU of A = Z^T in place on the device, solves Z^-1 R, the factor in
scipy.linalg.lu_factor's layout, the memory admission figures, and the registry get_backend(pref) with the
MOM25D_GPU override.

Beyond the reference's contract: factor_device / solve_device take and return device tensors (no host round
trip), and `pivot` chooses the pivoting of the factor - 'tournament' (default: a panel's pivots by a tree of GEPPs,
no per-column grid barrier) or 'partial' (LAPACK's pivots exactly, the layout and pivots of lu_factor).
There is no CPU stand-in: without a GPU or without the built library, get_backend returns None and says why.
"""
import os

import numpy as np


class GPUUnavailable(RuntimeError):
    """No usable GPU backend (the reference's GPUUnavailable, src/mom25d/gpu.py:25)."""


class _Backend:
    name, device, hip = '?', 'GPU', False
    inplace = False          # LU overwrites the uploaded matrix (VRAM = 1 x N^2 of the factor's dtype)
    lu = piv = None

    @property
    def api(self):
        return 'ROCm/HIP' if self.hip else 'none'

    def vram_needed(self, N):
        """Device bytes to factor an N x N complex64 matrix (src/mom25d/gpu.py:37-38)."""
        return (1 if self.inplace else 2) * 8.0 * N * N + 0.4e9

    def fits(self, N, extra=0.0):
        return self.vram_needed(N) <= 0.9 * (self.mem_free() + extra)

    def release(self):
        self.lu = self.piv = None

    def download(self):
        raise NotImplementedError


class HipBackend(_Backend):
    """The factor lives on the device (self.factor_dev, a device.Factor); lu / piv mirror it as device tensors."""
    name, hip, inplace = 'hip', True, True

    def __init__(self, pivot='tournament'):
        try:
            import torch
            if not torch.cuda.is_available():
                raise GPUUnavailable("hip: no GPU visible")
            from . import device as dv
            dv.device()                                      # loads libm25x.so and checks for gfx950
        except GPUUnavailable:
            raise
        except Exception as e:                               # library not built, wrong architecture
            raise GPUUnavailable(f"hip: {e}")
        self.torch, self.dv = torch, dv
        self.device = torch.cuda.get_device_name(dv.device())
        self.pivot = pivot
        self.factor_dev = None

    def mem_free(self):
        return float(self.torch.cuda.mem_get_info(self.dv.device())[0])

    def mem_total(self):
        return float(self.torch.cuda.mem_get_info(self.dv.device())[1])

    def factor(self, Z):
        """Z: host complex64 (or complex128), C-contiguous, equilibrated. Factors Z^T in place on the device
        (src/mom25d/gpu.py:69-74)."""
        Z = np.asarray(Z)
        if Z.ndim != 2 or Z.shape[0] != Z.shape[1] or Z.dtype not in (np.complex64, np.complex128):
            raise ValueError("factor: a square complex64 or complex128 matrix")
        self.factor_device(self.torch.from_numpy(np.ascontiguousarray(Z)).to(self.dv.device()))

    def factor_device(self, Zd):
        """Factors the device tensor Zd (row-major N x N complex64/complex128, i.e. A = Z^T column-major) in place;
        raises FloatingPointError on an exactly zero pivot, as a singular LAPACK factor would be unusable."""
        self.release()
        F = self.dv.getrf(Zd, pivot=self.pivot)
        info = int(F.info.item())
        if info:
            raise FloatingPointError(f"factor: exactly zero pivot at {info - 1}")
        self.factor_dev, self.lu, self.piv = F, F.LU, F.ipiv

    def solve(self, R):
        """Z^-1 R for host R (N or N x k), in the factor's precision, returned as host complex128
        (src/mom25d/gpu.py:76-80)."""
        R = np.asarray(R)
        Rd = self.torch.from_numpy(np.ascontiguousarray(R.astype(complex))).to(self.dv.device())
        X = self.solve_device(Rd).cpu().numpy()
        if self.dv.solve_faults():
            raise RuntimeError("solve: a triangular solve ran out of a wait bound: its result is not the solution")
        return X

    def solve_device(self, Rd, trans=1):
        """Z^-1 R (trans 1) or Z^-T R (trans 0) for a device tensor R; complex128 device tensor out."""
        if self.factor_dev is None:
            raise RuntimeError("solve: no factor")
        return self.dv.getrs(self.factor_dev, Rd, trans=trans)

    def download(self):
        """(lu, piv) on the host as scipy.linalg.lu_factor(Z.T) returns them: lu Fortran-ordered, piv 0-based int32
        (src/mom25d/gpu.py:46-48, 82-83)."""
        if self.factor_dev is None:
            raise RuntimeError("download: no factor")
        self.torch.cuda.synchronize()
        return self.lu.cpu().numpy().T, self.piv.cpu().numpy()

    def release(self):
        super().release()
        self.factor_dev = None


_BACKENDS = {'hip': HipBackend}
_cache = {}


def get_backend(pref='auto'):
    """A backend instance or None (src/mom25d/gpu.py:122-146). pref: 'auto' | 'hip' | 'off'; the environment
    variable MOM25D_GPU overrides it. Reasons for failures are kept in get_backend.why."""
    pref = os.environ.get('MOM25D_GPU', pref) or 'auto'
    get_backend.why = []
    if pref == 'off':
        return None
    names = list(_BACKENDS) if pref == 'auto' else [pref]
    for n in names:
        if n not in _BACKENDS:
            get_backend.why.append(f"unknown GPU backend {n!r}")
            continue
        try:
            if n not in _cache:
                _cache[n] = _BACKENDS[n]()
            return _cache[n]
        except GPUUnavailable as e:
            get_backend.why.append(str(e))
    return None


get_backend.why = []


def describe():
    """One-line status for logs."""
    b = get_backend()
    if b is None:
        return "GPU: none (" + "; ".join(get_backend.why) + ")"
    return f"GPU: {b.name} on {b.device} ({b.api}), {b.mem_free() / 1e9:.1f} of {b.mem_total() / 1e9:.1f} GB free"
