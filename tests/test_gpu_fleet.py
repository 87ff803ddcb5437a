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
        m0, c0, d0 = gpu.get_mapping(), gpu.get_cloud_tile(), gpu.get_deal()
        try:
            for mapping, tile, deal in ((gpu.MAP_PARTICLE_LANES, 64, 0), (gpu.MAP_PARTICLE_LANES, 256, 4),
                                        (gpu.MAP_PARTICLE_LANES, 32, 1), (gpu.MAP_POINT_LANES, 64, 3)):
                gpu.set_mapping(mapping)
                gpu.set_cloud_tile(tile)
                gpu.set_deal(deal)
                for persistent in (1, 0, -1):
                    print("fleet mapping", mapping, "tile", tile, "deal", deal, "persistent", persistent, flush=True)
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
                        key = (mapping, tile, deal, persistent, f)
                        assert np.array_equal(out[off:off + n].view(np.int32), op.view(np.int32)), key
                        assert np.array_equal(means[f].view(np.int32), om.view(np.int32)), key
                        off += n
        finally:
            gpu.set_mapping(m0)
            gpu.set_cloud_tile(c0)
            gpu.set_deal(d0)
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


def test_fleet_graph_step_matches_oracle(gpu):
    """One predict -> update -> resample of 16 uneven filters, captured in a
    HIP graph from the device-parameter forms and replayed three times with
    new draws, clouds, beacons and parameters written into the same device
    arrays between replays (with host work and synchronisation in between):
    every replay's particles and means are bit-exact against the oracle chain
    (ParticleFilter.cpp:97-222) run filter by filter."""
    from amcl3d_amd.fleet import Fleet

    rng = np.random.default_rng(11)
    c = W.room_case(2)
    bm = c["map"]
    _, prob = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
    og = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    grid = gpu.Grid3d().from_cells(prob, bm.min, bm.max, bm.res, c["sensor_dev"])
    F = 16
    counts = np.round(np.exp(rng.uniform(np.log(20), np.log(4000), F))).astype(np.uint32)
    counts[5] = 1
    ms = rng.integers(5, 3000, F).astype(np.uint32)
    nbs = rng.integers(0, 9, F).astype(np.uint32)
    poses = [np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), rng.uniform(-3, 3)) for _ in range(F)]
    parts = []
    for f in range(F):
        p = W.particles_around(poses[f], int(counts[f]), rng)
        p[: int(counts[f]) // 10, 0] = -1.0  # some outside the map
        parts.append(p)
    total = int(counts.sum())
    fl = Fleet(counts)
    graph = None
    try:
        fl.set_particles(np.concatenate(parts))
        fl.set_layout(ms, nbs)

        def inputs():
            clouds = [W.sensor_cloud(bm, poses[f], int(ms[f]), 0.05, rng) for f in range(F)]
            rngs = [W.beacons(bm, poses[f], int(nbs[f]), 0.1, rng) for f in range(F)]
            mods = rng.uniform(0.05, 0.3, (F, 4))
            deltas = rng.normal(0, 0.05, (F, 4))
            z4 = rng.standard_normal((total, 4)).astype(F32)
            params = np.stack([rng.uniform(0, 1, F), rng.uniform(0.05, 0.6, F),
                               np.where(np.arange(F) % 4 == 1, 0.05, 0.0), np.zeros(F)], 1)
            u01 = rng.uniform(0, 1, F).astype(F32)
            return clouds, rngs, mods, deltas, z4, params, u01

        first = inputs()
        cl = np.concatenate(first[0])
        R = np.concatenate([r for r in first[1] if r.shape[0]]).astype(F32)
        bufs = dict(x=gpu.DeviceBuffer(cl.shape[0] * 4), y=gpu.DeviceBuffer(cl.shape[0] * 4),
                    z=gpu.DeviceBuffer(cl.shape[0] * 4), r=gpu.DeviceBuffer(R.nbytes),
                    mods=gpu.DeviceBuffer(F * 32), deltas=gpu.DeviceBuffer(F * 32), z4=gpu.DeviceBuffer(total * 16),
                    params=gpu.DeviceBuffer(F * 32), u01=gpu.DeviceBuffer(F * 4))

        def upload(inp):
            clouds, rngs, mods, deltas, z4, params, u01 = inp
            c3 = np.concatenate(clouds).astype(F32)
            for k, name in enumerate("xyz"):
                _lib.call("a3d_buffer_upload", bufs[name].ptr, np.ascontiguousarray(c3[:, k]), c3.shape[0] * 4)
            rr = np.ascontiguousarray(np.concatenate([r for r in rngs if r.shape[0]]).astype(F32))
            _lib.call("a3d_buffer_upload", bufs["r"].ptr, rr, rr.nbytes)
            for name, a in (("mods", _lib.f64(mods)), ("deltas", _lib.f64(deltas)), ("z4", _lib.f32(z4)),
                            ("params", _lib.f64(params)), ("u01", _lib.f32(u01))):
                _lib.call("a3d_buffer_upload", bufs[name].ptr, a, a.nbytes)

        upload(first)
        fl.capture_begin()
        fl.predict_dev(bufs["mods"].ptr, bufs["deltas"].ptr, bufs["z4"].ptr)
        fl.update_dev(grid, bufs["x"].ptr, bufs["y"].ptr, bufs["z"].ptr, bufs["r"].ptr, bufs["params"].ptr)
        fl.resample_dev(bufs["u01"].ptr)
        graph = fl.capture_end()

        state = [p.copy() for p in parts]
        inp = first
        for replay in range(3):
            print("fleet graph replay", replay, flush=True)
            graph.launch()
            fl.synchronize()
            got = fl.particles()
            means = fl.means()
            clouds, rngs, mods, deltas, z4, params, u01 = inp
            off = 0
            for f in range(F):
                n = int(counts[f])
                op = oracle.pf_predict(state[f], mods[f], deltas[f], z4[off:off + n])
                op, om = oracle.pf_update(op, og, clouds[f], rngs[f], params[f, 0], params[f, 1],
                                          params[f, 2], params[f, 3])
                oq, _ = oracle.pf_resample(op, float(u01[f]), 0)
                key = (replay, f, n)
                assert np.array_equal(means[f].view(np.int32), om.view(np.int32)), key
                assert np.array_equal(got[off:off + n].view(np.int32), oq.view(np.int32)), key
                state[f] = oq
                off += n
            inp = inputs()
            upload(inp)
    finally:
        if graph is not None:
            graph.close()
        fl.close()


def test_fleet_dev_forms_refuse_without_layout(gpu):
    from amcl3d_amd.fleet import Fleet

    fl = Fleet([10, 20])
    try:
        with pytest.raises(_lib.Amcl3dError, match="set_layout"):
            fl.resample_dev(1)
        fl.set_layout([5, 6], [0, 1])
        fl.capture_begin()
        with pytest.raises(_lib.Amcl3dError, match="during capture"):
            fl.particles()
        g = fl.capture_end()
        g.close()
    finally:
        fl.close()


def test_fleet_host_predict_and_resample_match_oracle(gpu):
    """The host forms a3d_fleet_predict and a3d_fleet_resample (through
    amcl3d_amd.fleet.Fleet) on uneven filters: each filter bit for bit as the
    oracle's predict and resample of that filter alone."""
    from amcl3d_amd import fleet as fleet_mod

    rng = np.random.default_rng(8)
    F = 9
    counts = np.round(np.exp(rng.uniform(np.log(5), np.log(4000), F))).astype(np.uint32)
    parts = [W.particles_around(np.array([5.0 + f, 4.0, 1.5, rng.uniform(-3, 3)]), int(counts[f]), rng) for f in range(F)]
    for p in parts:
        p[:, 4] = rng.exponential(1.0, p.shape[0]).astype(F32)
    mods = rng.uniform(0.05, 0.3, (F, 4))
    deltas = rng.uniform(-0.2, 0.2, (F, 4))
    z4 = rng.standard_normal((int(counts.sum()), 4)).astype(F32)
    u01 = rng.uniform(0, 1, F).astype(F32)
    fl = fleet_mod.Fleet(counts)
    try:
        fl.set_particles(np.concatenate(parts))
        fl.predict(mods, deltas, z4)
        got = fl.particles()
        fl.resample(u01)
        got_rs = fl.particles()
    finally:
        fl.close()
    off = 0
    for f in range(F):
        n = int(counts[f])
        op = oracle.pf_predict(parts[f], mods[f], deltas[f], z4[off:off + n])
        assert np.array_equal(got[off:off + n].view(np.int32), op.view(np.int32)), f
        oq, _ = oracle.pf_resample(op, float(u01[f]), 0)
        assert np.array_equal(got_rs[off:off + n].view(np.int32), oq.view(np.int32)), f
        off += n


def test_fleet_predict_invalidates_the_staged_update(gpu):
    """a3d_fleet_predict replaces the descriptors a staged update uploaded:
    update_staged refuses until the update is staged again."""
    from amcl3d_amd import fleet as fleet_mod

    c = W.room_case(2)
    bm = c["map"]
    grid = gpu.Grid3d()
    grid.open(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"])
    counts = np.array([50, 70], np.uint32)
    rng = np.random.default_rng(1)
    fl = fleet_mod.Fleet(counts)
    try:
        fl.set_particles(np.concatenate([W.particles_around(c["pose"], int(n), rng) for n in counts]))
        clouds = [c["step"].cloud[:100], c["step"].cloud[100:180]]
        ranges = [c["step"].ranges, c["step"].ranges[:1]]
        fl.stage_update(clouds, ranges, 0.8, 0.1)
        fl.update_staged(grid)
        fl.predict(np.full((2, 4), 0.1), np.full((2, 4), 0.05), rng.standard_normal((120, 4)).astype(F32))
        with pytest.raises(_lib.Amcl3dError, match="stage_update"):
            fl.update_staged(grid)
        fl.stage_update(clouds, ranges, 0.8, 0.1)
        fl.update_staged(grid)
        fl.synchronize()
    finally:
        fl.close()
