"""memory_plan as a pure function (CPU): the reference's plan (src/mom25d/sim.py:85-151) and its test of the
backend's own estimate (tests/test_gpu.py:260-279), with the figures of this build, whose direct solver holds
everything on the GPU: 16 n^2 for the complex128 Z factored in place, 24 n^2 + 8 n^2 a kept factor for the mixed
path, plus a bounded workspace."""
import pytest

from mom25d_amd.sim import GPU_MIN_N, WORK_BYTES, device_bytes, memory_plan


def test_precision_threshold_and_paths():
    p = memory_plan(2000, gpu=('hip', 288e9))
    assert p['precision'] == 'double' and p['path'] == 'double' and p['solver'] == 'direct' and p['fits']
    assert p['gpu'] == 16.0 * 2000 ** 2 + WORK_BYTES
    p = memory_plan(GPU_MIN_N + 1, gpu=('hip', 288e9))
    assert p['precision'] == 'mixed' and p['path'] == 'hybrid'
    # the LU's own unknowns decide (quads would make them fewer than n)
    assert memory_plan(8000, gpu=('hip', 288e9), n_lu=2000)['precision'] == 'double'
    # a requested precision is kept
    assert memory_plan(8000, gpu=('hip', 288e9), precision='double')['gpu'] == 16.0 * 8000 ** 2 + WORK_BYTES


def test_device_footprint_and_fit():
    n = 40000
    p = memory_plan(n, gpu=('hip', 288e9))
    assert p['gpu'] == 24.0 * n * n + 4 * 8.0 * n * n + WORK_BYTES and p['fits']        # 90 GB of 288
    assert not memory_plan(n, gpu=('hip', 16e9))['fits']                                # a 16 GB GPU: no
    assert memory_plan(n, gpu=('hip', 60e9), keep=0)['gpu'] == 24.0 * n * n + WORK_BYTES
    # where the FP64 Z does not fit but the complex64 one does: the lean path (the refinement without Z)
    p = memory_plan(n, gpu=('hip', 16e9), keep=0)
    assert p['path'] == 'gpu-stream' and p['fits'] and p['gpu'] == 8.0 * n * n + WORK_BYTES
    assert not memory_plan(1000, gpu=None)['fits']                                     # no GPU: nothing fits
    # 20 000 unknowns, the serving configuration's largest board: 24 n^2 = 9.6 GB and four kept factors, 22.4 GB
    assert abs(device_bytes(20000, 'mixed') - 22.4e9 - WORK_BYTES) < 1


def test_backend_estimate_is_taken():
    """The 4-tuple of gpu_info carries the backend's own estimate of its factor; the plan adds the FP64 Z."""
    n = 10000
    need = 8.0 * n * n + 0.4e9
    p = memory_plan(n, gpu=('hip', 288e9, need * 10, True), keep=0)
    assert p['gpu'] == need * 10 + 16.0 * n * n


def test_bad_arguments():
    with pytest.raises(ValueError):
        memory_plan(100, precision='half')
    with pytest.raises(ValueError):
        memory_plan(-1)
