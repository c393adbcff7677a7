"""The oracle: TEST INFRASTRUCTURE, not part of the product.

A CPU restatement of the reference's per-frequency MoM solve, used only as the checker of the HIP product path
(by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg). The product (mom25d_amd) never imports,
links or runs anything from here.

  core.py      the reference's compiled-kernel module (src/mom25d/core/mom25d_core.cpp) restated in C
               (m25_oracle.c, liboracle.so) behind the same Python names and dict layouts
  greens.py    the layered-media spectral network, Sommerfeld contour, Hankel transforms, Table1D, GreenTableML,
               TMNetwork and ViaKernels (src/mom25d/multilayer.py, src/mom25d/via.py) in NumPy
  geometry.py  RWG basis, quadrature points and the kernels' geometry dicts (src/mom25d/base.py, fast.py, sim.py)
  gmres.py     gmres_batched (src/mom25d/fast.py:1437-1500)
  solve.py     the direct solve of one frequency: tables, fill, near field, via blocks, equilibration, LU, S
  ref.py       loads the reference's own compiled kernels from oracle/_ref/ where build_ref.sh made them

Each function cites the reference file:line it follows. What pins each one (golden vectors, closed forms,
the reference's compiled kernels) is in tests/test_oracle_*.py.
"""
