"""GPU: the allocation-free update (a3d_pf_update_noalloc, working memory in a
static device array) against the filter's own update (a3d_pf_update_dev) on
the same inputs, and both against the oracle (ParticleFilter.cpp:121-196):
particles and mean bit-identical on two calls in a row, the second with fewer
particles, points and a non-zero roll, so it runs over the first call's
leftovers in the shared array."""
import ctypes

import numpy as np
import pytest

import oracle
from amcl3d_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32
MAX_PARTICLES = 131072


def _noalloc(gpu, grid, parts, bufs, m, nb, alpha, sigma, roll, pitch):
    n = parts.shape[0]
    dp = gpu.DeviceBuffer.from_array(np.ascontiguousarray(parts.T))  # 7 arrays of n floats
    dmean = gpu.DeviceBuffer(16)
    x, y, z, r = bufs
    _lib.call("a3d_pf_update_noalloc", grid.handle, dp.ptr, n, x.ptr, y.ptr, z.ptr, m, r.ptr if nb else None, nb,
              float(alpha), float(sigma), float(roll), float(pitch), dmean.ptr, None)
    _lib.call("a3d_device_synchronize")
    out = dp.download(np.float32, 7 * n).reshape(7, n).T.copy()
    return out, dmean.download(np.float32, 4)


def test_noalloc_update_matches_update_dev(gpu):
    c = W.room_case(4)
    bm = c["map"]
    grid = gpu.Grid3d()
    grid.open(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"])
    og = oracle.GridMeta(bm.min, bm.max, bm.res, grid.read_cells())
    rng = np.random.default_rng(17)
    for call, (n, m, nb, roll, pitch) in enumerate(((60000, 2000, 3, 0.0, 0.0), (1500, 700, 8, 0.03, -0.02))):
        print("no-alloc call", call, "n", n, "m", m, flush=True)
        pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), rng.uniform(-3, 3))
        cloud = W.sensor_cloud(bm, pose, m, 0.05, rng)
        ranges = W.beacons(bm, pose, nb, 0.1, rng)
        parts = W.particles_around(pose, n, rng)
        parts[: n // 20, 2] = -5.0  # some outside the map
        bufs = tuple(gpu.DeviceBuffer.from_array(np.ascontiguousarray(cloud[:, k])) for k in range(3)) + \
            (gpu.DeviceBuffer.from_array(ranges),)
        pf = gpu.ParticleFilter()
        pf.set_particles(parts)
        pf.update_dev(grid, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, m, bufs[3].ptr, nb, 0.8, 0.1, roll, pitch)
        pf.synchronize()
        ref, ref_mean = pf.particles(), pf.getMean()
        got, got_mean = _noalloc(gpu, grid, parts, bufs, m, nb, 0.8, 0.1, roll, pitch)
        assert np.array_equal(got.view(np.int32), ref.view(np.int32)), call
        assert np.array_equal(got_mean.view(np.int32), ref_mean.view(np.int32)), call
        op, om = oracle.pf_update(parts, og, cloud, ranges, 0.8, 0.1, roll, pitch)
        assert np.array_equal(got.view(np.int32), op.view(np.int32)), call
        assert np.array_equal(got_mean.view(np.int32), om.view(np.int32)), call
        pf.close()


def test_noalloc_update_rejects_more_than_its_scratch(gpu):
    c = W.room_case(4)
    bm = c["map"]
    grid = gpu.Grid3d()
    grid.open(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"])
    with pytest.raises(_lib.Amcl3dError, match="at most 131072 particles"):
        _lib.call("a3d_pf_update_noalloc", grid.handle, ctypes.c_void_p(16), MAX_PARTICLES + 1, None, None, None, 0,
                  None, 0, 0.8, 0.1, 0.0, 0.0, ctypes.c_void_p(16), None)
