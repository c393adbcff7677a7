"""mom25d_amd - the per-frequency method-of-moments solve of mom25d (2.5-D MPIE, RWG basis, layered-media Green's
tables, vias) on AMD MI355X: the impedance-matrix fill and the dense complex LU as hand-written HIP kernels for
gfx950 (libm25x.so, C ABI in include/m25x.h), host code in Python on PyTorch-ROCm.

    from mom25d_amd.multilayer import Layer, LayerStack
    from mom25d_amd.sim import Simulation, Via, ViaPort, GapPort
    from mom25d_amd.run import solve_band
    import mom25d_amd.core        # the reference's compiled-kernel module, served by the GPU
"""
__version__ = '0.1.0'
