"""The entry point that allocates nothing (bfvm_voxel_positions_noalloc), through the C ABI on an MI355X: its
indices equal those of bfvm_voxel_positions (which allocates staging buffers per call) on two calls in a row,
the second with fewer points, on 3-D and 2-D meshes; more points than its scratch holds is an argument error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 20


def _indices(lib, h, fn, pos, dims):
    count = pos.size // dims
    out = np.full(count * 3, -7, dtype=np.int32)
    rc = getattr(lib, fn)(h, count, pos.ctypes.data, out.ctypes.data)
    return rc, out.reshape(count, 3)


@pytest.mark.parametrize("dims,grid", [(3, (130, 70, 90)), (2, (300, 200, 1))])
def test_noalloc_matches_allocating_entry_point(dims, grid):
    from parabiofvm_amd import _native
    from parabiofvm_amd.solver import DeviceSolver

    lib = _native.load()
    maxs = [g * 20 if d < dims else 0 for d, g in enumerate(grid)]
    s = DeviceSolver(dims, [0, 0, 0], maxs, [20, 20, 20], 2, 0.01, np.ones(2, np.float32), np.ones(2, np.float32),
                     np.zeros(2, np.float32))
    s.initialize()
    rng = np.random.Generator(np.random.PCG64(3 + dims))
    lengths = np.array(maxs[:dims], dtype=np.float64)
    for count in (MAX_POINTS, 12345):  # the full scratch, then fewer points: nothing of the first call leaks
        pos = (rng.uniform(-0.05, 1.05, size=(count, dims)) * lengths).astype(np.float32)
        # exact voxel boundaries and points just below them
        k = min(count, 4096)
        pos[:k] = (rng.integers(0, 50, size=(k, dims)) * 20).astype(np.float32)
        pos[k:2 * k] = np.nextafter(pos[:k], np.float32(-np.inf))
        flat = np.ascontiguousarray(pos.reshape(-1))
        rc1, ref = _indices(lib, s.h, "bfvm_voxel_positions", flat, dims)
        rc2, got = _indices(lib, s.h, "bfvm_voxel_positions_noalloc", flat, dims)
        assert rc1 == 0 and rc2 == 0, (rc1, rc2, lib.bfvm_last_error())
        np.testing.assert_array_equal(got, ref)
        assert np.all(got != -7)
    big = np.zeros((MAX_POINTS + 1) * dims, dtype=np.float32)
    out = np.zeros((MAX_POINTS + 1) * 3, dtype=np.int32)
    rc = lib.bfvm_voxel_positions_noalloc(s.h, MAX_POINTS + 1, big.ctypes.data, out.ctypes.data)
    assert rc == -1
    assert b"exceeds" in lib.bfvm_last_error()
    s.close()
