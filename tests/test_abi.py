"""The C ABI (CPU: no kernel is launched): libm25x.so loads and exports every function include/m25x.h declares,
and the ctypes binding names each of them; the product refuses to run without a GPU instead of falling back."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'm25x.h')).read()
    return sorted(set(re.findall(r'\b(m25x_\w+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    from mom25d_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libm25x.so is not built (make -C mom25d_amd/csrc)")
    h = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing
    bound = set(_lib._SIGS) | {'m25x_last_error'}
    assert set(names) <= bound, sorted(set(names) - bound)
    assert h.m25x_version() == 1


def test_no_gpu_no_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from mom25d_amd import core
    import numpy as np
    with pytest.raises(RuntimeError):
        core.bessel_jn_complex(0, np.ones(3, complex))
