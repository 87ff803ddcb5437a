"""GPU parity on the large filter's map: the seeded building of 100 x 100 x 30 m
at 0.05 m (2000 x 2000 x 600 = 2.4e9 cells, 9.6 GB of prob), built on the
device.

The map has fewer than 2^32 - 1 cells, sx < 2^24 and sy * sz < 2^24, so the
cloud weight gathers it through 32-bit cell indices; its 9.6 GB are more than
a structured buffer's 32-bit byte offsets reach, so the gathers take 64-bit
addresses.  Poses on the upper storeys land points in cells past index 2^30
and 2^31, at byte offsets past 4 GiB and 8 GiB.  The device weights, a
captured predict -> update -> resample step and a 100,000-particle update
are checked against the oracle on the grid read back from the device, and the
grid itself against the oracle's nearest distances on sampled cells, most of
them on those storeys.
"""
import sys
import time

import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32
FLOOR = 3.0
SENSOR_DEV = 0.05
# sensor heights: storeys 4 (cells from 2^30), 6, 8 and 9 (past 2^31)
STOREY_Z = (13.5, 19.5, 25.5, 28.5)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def building(gpu):
    t0 = time.time()
    bm = W.building_map(7)
    pts = bm.points(dedup=False)
    size = [int(v) for v in bm.size]
    assert size == [2000, 2000, 600]
    cells = size[0] * size[1] * size[2]
    # the library's rule for 32-bit cell indices; this map must take them
    assert cells < 2 ** 32 - 1 and size[0] < 2 ** 24 and size[1] * size[2] < 2 ** 24
    _say("building map: %d points (%.1fs); device computeGrid of %d cells" % (pts.shape[0], time.time() - t0, cells))
    g = gpu.Grid3d()
    g.open(pts, bm.min, bm.max, bm.res, SENSOR_DEV)
    t1 = time.time()
    prob = g.read_cells()
    _say("read back %.1f GB in %.1fs" % (prob.nbytes / 1e9, time.time() - t1))
    return dict(bm=bm, pts=pts, grid=g, prob=prob, og=oracle.GridMeta(bm.min, bm.max, bm.res, prob))


def _pose(bm, z, rng):
    view = W.storey_view(bm, z, FLOOR)
    p = W.free_pose(view, rng, z_range=(z, z), clearance=1.0)
    return np.append(p, rng.uniform(-np.pi, np.pi)), view


def test_grid_upper_storeys_against_oracle(building):
    """Sampled cells, two thirds of them past index 2^30, near the map's
    surfaces: prob equals the oracle's (nearest float squared distance, then
    c1 * expf(-d2 * d2 * c2) as exp of the double)."""
    bm, pts, prob = building["bm"], building["pts"], building["prob"]
    rng = np.random.default_rng(3)
    sz = bm.size.astype(np.int64)
    near = pts[rng.integers(0, pts.shape[0], 60000)].astype(np.float64)
    near += rng.uniform(-0.5, 0.5, near.shape)
    ijk = np.clip(np.floor((near - bm.min) / bm.res).astype(np.int64), 0, sz - 1)
    lin = ijk[:, 0] + sz[0] * (ijk[:, 1] + sz[1] * ijk[:, 2])
    high = lin[lin >= 2 ** 30]
    lin = np.concatenate([high[:30000], lin[lin < 2 ** 30][:10000], rng.integers(2 ** 31, int(np.prod(sz)), 2000)])
    assert (lin >= 2 ** 31).sum() > 5000
    ix, iy, iz = lin % sz[0], (lin // sz[0]) % sz[1], lin // (sz[0] * sz[1])
    q32 = np.stack([bm.min[0] + ix * bm.res, bm.min[1] + iy * bm.res, bm.min[2] + iz * bm.res], -1).astype(F32)
    t0 = time.time()
    d2 = oracle.nearest_d2(pts, bm.min, bm.max, bm.res, q32)
    _say("oracle nearest distances of %d cells (%.1fs)" % (lin.size, time.time() - t0))
    c1 = F32(1.0 / (SENSOR_DEV * np.sqrt(2 * np.pi)))
    c2 = F32(1.0 / (2.0 * SENSOR_DEV * SENSOR_DEV))
    e = (-d2 * d2 * c2).astype(F32)
    ref = (c1 * np.exp(e.astype(np.float64)).astype(F32)).astype(F32)
    got = prob[lin]
    assert (got > 0).mean() > 0.5
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("z", STOREY_Z)
def test_cloud_weights_upper_storeys(gpu, building, z):
    """Weights of 2,000 spread poses on one storey, 20,000-point clouds (noise
    0.03 m), bit-exact against the oracle; the gathered cells reach past 2^30
    (and past 2^31 on the top two storeys)."""
    bm, g, og = building["bm"], building["grid"], building["og"]
    rng = np.random.default_rng(int(z * 10))
    pose, view = _pose(bm, z, rng)
    cloud = W.sensor_cloud(view, pose, 20000, 0.03, rng)
    p = W.particles_around(pose, 2000, rng)
    cells = oracle.cloud_cells(og, cloud, p[:40, 0], p[:40, 1], p[:40, 2], p[:40, 3])
    top = int(cells.max())
    _say("storey z=%.1f: pose %s, highest gathered cell %d (byte offset %.2f GiB)" % (z, pose, top, top * 4 / 2 ** 30))
    assert top >= 2 ** 30
    if z > 25:
        assert top >= 2 ** 31
    w = g.cloud_weights(cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    ow = oracle.cloud_weight_batch(og, cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert (ow > 0).mean() > 0.9
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_cloud_weights_segment_boundary_roll_pitch(gpu, building):
    """Poses around the first 4 GiB segment boundary (cell index 2^30, at
    z = 13.4 m), a band of the scan 0.6 m high, poses sorted by height, with
    and without roll and pitch: each wavefront bounds its cells through the
    cloud's rotated z and reads them through its segment's buffer or,
    straddling, through 64-bit addresses; bit-exact against the oracle
    either way.  The oracle's cells show wavefronts of 16 poses wholly on
    one side of the boundary and across it."""
    bm, g, og = building["bm"], building["grid"], building["og"]
    rng = np.random.default_rng(41)
    pose, view = _pose(bm, 13.5, rng)
    full = W.sensor_cloud(view, pose, 20000, 0.03, rng)
    cloud = np.ascontiguousarray(full[np.abs(full[:, 2]) < 0.3])
    p = W.particles_around(pose, 640, rng, dxyz=1.0, dyaw=0.5)
    p = p[np.argsort(p[:, 2], kind="stable")]
    for roll, pitch in ((0.0, 0.0), (0.04, -0.03)):
        w = g.cloud_weights(cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        ow = oracle.cloud_weight_batch(og, cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        cells = oracle.cloud_cells(og, cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        side = []
        for w0 in range(0, p.shape[0], 16):
            c = cells[w0:w0 + 16]
            c = c[c >= 0]
            if c.size:
                side.append((int(c.min() >> 30), int(c.max() >> 30)))
        below = sum(1 for lo, hi in side if hi == 0)
        above = sum(1 for lo, hi in side if lo == 1)
        across = sum(1 for lo, hi in side if lo == 0 and hi == 1)
        _say("roll %.2f pitch %.2f: %d points; wavefronts of 16 below / above / across 2^30: %d / %d / %d"
             % (roll, pitch, cloud.shape[0], below, above, across))
        # level: wavefronts on both sides and across; tilted, the band's
        # rotated z spans more layers, and more wavefronts straddle
        assert across and (below or above) and (roll != 0.0 or (below and above))
        assert (ow > 0).mean() > 0.9
        assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_captured_step_top_storey(gpu, building):
    """One predict -> update -> resample step, captured in a HIP graph as the
    large filter runs it, with 5,000 particles on the top storey and 8
    beacons: particles, mean and resampled state bit-exact against the
    oracle chain."""
    bm, g, og = building["bm"], building["grid"], building["og"]
    rng = np.random.default_rng(29)
    pose, view = _pose(bm, 28.5, rng)
    n = 5000
    cloud = W.sensor_cloud(view, pose, 20000, 0.03, rng)
    anchors = rng.uniform(bm.min + 1, bm.max - 1, (8, 3))
    r = np.linalg.norm(anchors - pose[:3], axis=1) + rng.normal(0, 0.1, 8)
    ranges = np.concatenate([r[:, None], anchors], 1).astype(F32)
    parts = W.particles_around(pose, n, rng)
    z4 = rng.standard_normal((n, 4)).astype(F32)
    u01 = float(rng.uniform())
    mods, deltas = [0.1, 0.1, 0.1, 0.3], [0.21, -0.04, 0.0, 0.03]
    dz = gpu.DeviceBuffer.from_array(z4)
    x, y, zc = (gpu.DeviceBuffer.from_array(np.ascontiguousarray(cloud[:, k])) for k in range(3))
    dr = gpu.DeviceBuffer.from_array(ranges)
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    pf.capture_begin()
    pf.predict_dev(mods, deltas, dz)
    pf.update_dev(g, x.ptr, y.ptr, zc.ptr, cloud.shape[0], dr.ptr, 8, 0.8, 0.1)
    pf.resample_dev(u01)
    graph = pf.capture_end()
    _say("captured step: 5,000 particles on the top storey")
    graph.launch()
    pf.synchronize()
    op = oracle.pf_predict(parts, mods, deltas, z4)
    op, om = oracle.pf_update(op, og, cloud, ranges, 0.8, 0.1)
    oq, _ = oracle.pf_resample(op, u01, 0)
    assert np.array_equal(pf.getMean().view(np.int32), om.view(np.int32))
    assert np.array_equal(pf.particles().view(np.int32), oq.view(np.int32))
    graph.close()


def test_large_filter_top_storey(gpu, building):
    """The large filter's shape (100,000 particles, 20,000 points, 8 beacons)
    on the top storey: 400 sampled particles' raw weights against the oracle
    (wp = weight / wtp, so their ratios), and the resample indices of the
    device's weights against the oracle's."""
    bm, g, og = building["bm"], building["grid"], building["og"]
    rng = np.random.default_rng(31)
    pose, view = _pose(bm, 28.5, rng)
    n = 100000
    cloud = W.sensor_cloud(view, pose, 20000, 0.03, rng)
    anchors = rng.uniform(bm.min + 1, bm.max - 1, (8, 3))
    r = np.linalg.norm(anchors - pose[:3], axis=1) + rng.normal(0, 0.1, 8)
    ranges = np.concatenate([r[:, None], anchors], 1).astype(F32)
    parts = W.particles_around(pose, n, rng)
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    _say("update: 100,000 particles x 20,000 points on the top storey")
    pf.update(g, cloud, ranges, 0.8, 0.1)
    got = pf.particles()
    sel = rng.choice(n, 400, replace=False)
    ow = oracle.cloud_weight_batch(og, cloud, parts[sel, 0], parts[sel, 1], parts[sel, 2], parts[sel, 3])
    raw = got[sel, 5].astype(np.float64)
    ratio = raw / max(raw.max(), 1e-30)
    oratio = ow.astype(np.float64) / max(ow.max(), 1e-30)
    assert np.allclose(ratio, oratio, rtol=1e-6, atol=1e-9)
    assert abs(got[:, 4].astype(np.float64).sum() - 1.0) < 1e-3
    idx = pf.resample(0.37, gpu.RESAMPLE_REFERENCE, return_indices=True)
    assert np.array_equal(idx, oracle.resample_indices(got[:, 4], 0.37, 0))


def test_fleet_upper_storeys(gpu, building):
    """Eight filters on storeys 4 to 9 in one persistent fleet launch (the
    batched kernel through the same 64-bit addresses): every filter's updated
    particles and mean bit-exact against the oracle."""
    from amcl3d_amd.fleet import Fleet

    bm, g, og = building["bm"], building["grid"], building["og"]
    rng = np.random.default_rng(37)
    F = 8
    counts = rng.integers(200, 1500, F).astype(np.uint32)
    parts, clouds, ranges = [], [], []
    for f in range(F):
        pose, view = _pose(bm, STOREY_Z[f % len(STOREY_Z)], rng)
        clouds.append(W.sensor_cloud(view, pose, int(rng.integers(2000, 8000)), 0.03, rng))
        anchors = rng.uniform(bm.min + 1, bm.max - 1, (3, 3))
        r = np.linalg.norm(anchors - pose[:3], axis=1) + rng.normal(0, 0.1, 3)
        ranges.append(np.concatenate([r[:, None], anchors], 1).astype(F32))
        parts.append(W.particles_around(pose, int(counts[f]), rng))
    fl = Fleet(counts)
    try:
        fl.set_particles(np.concatenate(parts))
        fl.stage_update(clouds, ranges, 0.8, 0.1)
        _say("fleet: %d filters on the upper storeys" % F)
        fl.update_staged(g)
        fl.synchronize()
        got, means = fl.particles(), fl.means()
        off = 0
        for f in range(F):
            n = int(counts[f])
            op, om = oracle.pf_update(parts[f], og, clouds[f], ranges[f], 0.8, 0.1)
            assert np.array_equal(got[off:off + n].view(np.int32), op.view(np.int32)), f
            assert np.array_equal(means[f].view(np.int32), om.view(np.int32)), f
            off += n
    finally:
        fl.close()
