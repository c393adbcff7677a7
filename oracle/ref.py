"""The reference's own compiled kernels (src/mom25d/core/mom25d_core.cpp), built from its sources by
oracle/build_ref.sh into oracle/_ref/. TEST INFRASTRUCTURE ONLY: a checker of the oracle's C restatement and a
CPU baseline. None where it was not built (no reference tree in the container that built the repo)."""
import glob
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_mod = False


def core():
    """The compiled reference module, or None."""
    global _mod
    if _mod is False:
        paths = sorted(glob.glob(os.path.join(_HERE, '_ref', '_core*.so')))
        _mod = None
        if paths:
            spec = importlib.util.spec_from_file_location('_core', paths[0])
            m = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(m)
            _mod = m
    return _mod
