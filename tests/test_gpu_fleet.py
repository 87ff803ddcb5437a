"""GPU: many uneven filters in one launch give each filter's own results."""
import ctypes

import numpy as np
import pytest

import oracle
from amcl3d_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32


def _fleet(gpu, counts):
    h = ctypes.c_void_p()
    c = _lib.u32(counts)
    _lib.call("a3d_fleet_create", len(counts), _lib.ptr(c), ctypes.byref(h))
    return h.value


def test_fleet_matches_single_filters(gpu):
    rng = np.random.default_rng(5)
    c = W.room_case(1)
    bm = c["map"]
    pts = bm.points()
    _, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
    og = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    grid = gpu.Grid3d().from_cells(prob, bm.min, bm.max, bm.res, c["sensor_dev"])
    F = 12
    counts = np.round(np.exp(rng.uniform(np.log(20), np.log(3000), F))).astype(np.uint32)
    counts[3] = 1
    ms = rng.integers(5, 3000, F).astype(np.uint32)
    nbs = rng.integers(0, 9, F).astype(np.uint32)
    parts, clouds, rngs = [], [], []
    for f in range(F):
        pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), rng.uniform(-3, 3))
        clouds.append(W.sensor_cloud(bm, pose, int(ms[f]), 0.05, rng))
        rngs.append(W.beacons(bm, pose, int(nbs[f]), 0.1, rng))
        p = W.particles_around(pose, int(counts[f]), rng)
        p[: int(counts[f]) // 10, 0] = -1.0  # some outside the map
        parts.append(p)
    alpha = rng.uniform(0, 1, F)
    sigma = rng.uniform(0.05, 0.6, F)
    roll = np.where(np.arange(F) % 4 == 1, 0.05, 0.0)
    pitch = np.zeros(F)
    u01 = rng.uniform(0, 1, F).astype(F32)
    P = np.concatenate(parts)
    cl = np.concatenate(clouds)
    R = np.concatenate(rngs) if sum(nbs) else np.zeros((0, 4), F32)
    fl = _fleet(gpu, counts)
    try:
        _lib.call("a3d_fleet_set_particles", fl, _lib.f32(P))
        x, y, z = _lib.cloud_soa(cl)
        args = [_lib.ptr(x), _lib.ptr(y), _lib.ptr(z), _lib.u32(ms), _lib.f32(R),
                _lib.u32(nbs), _lib.f64(alpha), _lib.f64(sigma),
                _lib.f64(roll), _lib.f64(pitch)]
        expect = [oracle.pf_update(parts[f], og, clouds[f], rngs[f], alpha[f], sigma[f], roll[f], pitch[f])
                  for f in range(F)]
        m0 = gpu.get_mapping()
        try:
            for mapping in (gpu.MAP_PARTICLE_LANES, gpu.MAP_POINT_LANES):
                gpu.set_mapping(mapping)
                for persistent in (1, 0, -1):
                    print("fleet mapping", mapping, "persistent", persistent, flush=True)
                    _lib.call("a3d_fleet_set_particles", fl, _lib.f32(P))
                    _lib.call("a3d_fleet_stage_update", fl, *args)
                    _lib.call("a3d_fleet_update_staged", fl, grid.handle, persistent)
                    _lib.call("a3d_fleet_synchronize", fl)
                    out = np.empty_like(P)
                    _lib.call("a3d_fleet_get_particles", fl, _lib.ptr(out))
                    means = np.empty((F, 4), F32)
                    _lib.call("a3d_fleet_get_means", fl, _lib.ptr(means))
                    off = 0
                    for f in range(F):
                        op, om = expect[f]
                        n = int(counts[f])
                        key = (mapping, persistent, f)
                        assert np.array_equal(out[off:off + n].view(np.int32), op.view(np.int32)), key
                        assert np.array_equal(means[f].view(np.int32), om.view(np.int32)), key
                        off += n
        finally:
            gpu.set_mapping(m0)
        _lib.call("a3d_fleet_resample", fl, _lib.ptr(u01), 0)
        res = np.empty_like(P)
        _lib.call("a3d_fleet_get_particles", fl, _lib.ptr(res))
        off = 0
        for f in range(F):
            n = int(counts[f])
            oq, _ = oracle.pf_resample(out[off:off + n], float(u01[f]), 0)
            assert np.array_equal(res[off:off + n].view(np.int32), oq.view(np.int32)), f
            off += n
    finally:
        _lib.call("a3d_fleet_destroy", fl)
