"""GPU parity: the HIP library against the oracle on the same seeded inputs.

Bar: computeGrid's distance field bit-exact and prob bit-exact against the
oracle's correctly rounded expf (exp_mode=1); per-particle weights and the
filter state bit-exact (the stated tolerance, 1e-5 relative, is checked too);
resample indices bit-exact.
"""
import sys

import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def room(gpu):
    c = W.room_case(1)
    bm = c["map"]
    pts = bm.points()
    dist, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
    g = gpu.Grid3d()
    _say("room: device computeGrid", bm.size, pts.shape)
    g.open(pts, bm.min, bm.max, bm.res, c["sensor_dev"], keep_dist=True)
    c.update(points=pts, odist=dist, oprob=prob, grid=g, ogrid=oracle.GridMeta(bm.min, bm.max, bm.res, prob))
    return c


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


# -------------------------------------------------------------- computeGrid
def test_grid_room_bit_exact(room):
    p, d = room["grid"].read_cells(dist=True)
    assert list(room["grid"].size) == [200, 200, 60]
    assert np.array_equal(d, room["odist"].reshape(-1))
    mism = (p.view(np.int32) != room["oprob"].reshape(-1).view(np.int32)).sum()
    _say("room grid prob mismatches vs oracle:", mism)
    assert mism == 0


def test_grid_kat_map(gpu, kat, kat_oracle_grid):
    ext = np.asarray(kat["map_max"]) - np.asarray(kat["map_min"])
    g = gpu.Grid3d()
    _say("kat: device computeGrid 522x384x153")
    g.open(kat["map_cloud"], np.zeros(3), ext, kat["map_res"], kat["sensor_dev"], keep_dist=True)
    prob, dist = g.read_cells(dist=True)
    odist, oprob = kat_oracle_grid
    assert np.array_equal(dist, odist.reshape(-1))
    ulps = np.abs(prob.view(np.int32).astype(np.int64) - oprob.reshape(-1).view(np.int32))
    _say("kat grid prob: cells differing from glibc expf:", int((ulps != 0).sum()), "max ulp", int(ulps.max()))
    assert ulps.max() <= 1
    # KAT-2: the nav slice
    s = oracle.slice_int8(prob.reshape(odist.shape)[20]).reshape(kat["nav"].shape).astype(np.int32)
    dd = s - kat["nav"].astype(np.int32)
    assert np.abs(dd).max() <= 1 and (dd == 0).mean() >= 0.999
    # KAT-1: computeCloudWeight at the known pose, on the device-built cells
    g2 = gpu.Grid3d().from_cells(prob, kat["map_min"], kat["map_max"], kat["map_res"], kat["sensor_dev"])
    p = kat["kat_pose"]
    w = g2.computeCloudWeight(kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], p["roll"], p["pitch"], p["yaw"])
    og = oracle.GridMeta(kat["map_min"], kat["map_max"], kat["map_res"], prob)
    ow = oracle.cloud_weight(og, kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], p["roll"], p["pitch"], p["yaw"])
    _say("KAT-1 device weight", w, "reference 3.8109049797058105")
    assert w == ow
    assert abs(w - kat["kat_weight"]) <= kat["kat_tolerance"]
    assert not g2.isIntoMap(*kat["isinto_out"]) and g2.isIntoMap(*kat["isinto_in"])


def test_grid_empty_and_tiny(gpu):
    g = gpu.Grid3d()
    g.open(np.zeros((0, 3), F32), np.zeros(3), np.ones(3), 0.25, 0.1, keep_dist=True)
    p, d = g.read_cells(dist=True)
    assert np.all(p == 0) and np.all(d == -1)
    rng = np.random.default_rng(7)
    mn = np.array([-1.3, 0.7, -0.35])
    mx = mn + np.array([1.2, 1.0, 0.8])
    pts = (mn + rng.uniform(-0.2, 1.3, (40, 3)) * (mx - mn)).astype(F32)
    pts[3] = np.nan  # invalid points are dropped, as PCL's kd-tree does
    g.open(pts, mn, mx, 0.1, 0.05, keep_dist=True)
    od, op = oracle.compute_grid(pts, mn, mx, 0.1, 0.05, exp_mode=1)
    p, d = g.read_cells(dist=True)
    assert np.array_equal(d, od.reshape(-1)) and np.array_equal(p, op.reshape(-1))


def test_grid_prob_only_sparse_map(gpu):
    """Without the dist field the search stops where prob reaches 0; on a
    sparse map (most cells farther than that from any point) prob is still
    bit-exact, and the same as a build that keeps dist."""
    bm = W.BoxMap(np.zeros(3), np.array([10.0, 8.0, 4.0]), 0.05,
                  [[1.0, 1.0, 0.0, 1.6, 1.4, 0.9], [7.2, 5.5, 1.0, 7.5, 6.9, 3.1], [4.0, 0.3, 3.7, 4.2, 0.6, 3.9]])
    pts = bm.points()
    _, oprob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, 0.05, exp_mode=1)
    g = gpu.Grid3d()
    _say("sparse map prob-only build", bm.size, pts.shape)
    g.open(pts, bm.min, bm.max, bm.res, 0.05)
    p = g.read_cells()
    assert np.array_equal(p.view(np.int32), oprob.reshape(-1).view(np.int32))
    assert (p == 0).mean() > 0.5  # most of the map is beyond the cut
    g2 = gpu.Grid3d()
    g2.open(pts, bm.min, bm.max, bm.res, 0.05, keep_dist=True)
    p2, _ = g2.read_cells(dist=True)
    assert np.array_equal(p.view(np.int32), p2.view(np.int32))


def test_grid_file_round_trip(gpu, room, tmp_path):
    """The .grid cache (Grid3d::saveGrid / loadGrid, Grid3d.cpp:210-275): the
    file holds u32 sx, sy, sz, f64 sensor_dev, then {f32 dist, f32 prob} per
    cell, byte for byte what the reference writes for the same cells; loading
    it gives the same grid; a different sensor_dev is refused."""
    bm = room["map"]
    path = tmp_path / "room.grid"
    room["grid"].save(path)
    raw = path.read_bytes()
    sx, sy, sz = np.frombuffer(raw[:12], np.uint32)
    assert (sx, sy, sz) == (200, 200, 60)
    assert np.frombuffer(raw[12:20], np.float64)[0] == room["sensor_dev"]
    cells = np.frombuffer(raw[20:], np.float32).reshape(-1, 2)
    assert np.array_equal(cells[:, 0], room["odist"].reshape(-1))
    assert np.array_equal(cells[:, 1].view(np.int32), room["oprob"].reshape(-1).view(np.int32))
    g = gpu.Grid3d().load(path, bm.min, bm.max, bm.res, room["sensor_dev"])
    p, d = g.read_cells(dist=True)
    assert np.array_equal(p, cells[:, 1]) and np.array_equal(d, cells[:, 0])
    s, parts = room["step"], room["particles"]
    w1 = room["grid"].cloud_weights(s.cloud, parts[:50, 0], parts[:50, 1], parts[:50, 2], parts[:50, 3])
    w2 = g.cloud_weights(s.cloud, parts[:50, 0], parts[:50, 1], parts[:50, 2], parts[:50, 3])
    assert np.array_equal(w1, w2)
    with pytest.raises(gpu.Amcl3dError):
        gpu.Grid3d().load(path, bm.min, bm.max, bm.res, room["sensor_dev"] * (1 + 1e-9))


# --------------------------------------------------------- computeCloudWeight
# (mapping, particles per wavefront, points per LDS cloud tile): every ppw and
# every tile size of the particle-lane kernels, and the point-lane kernels
SHAPES = [(0, 64, 64), (0, 64, 256), (0, 32, 32), (0, 32, 128), (0, 16, 32), (0, 16, 64), (0, 16, 128),
          (0, 16, 256), (1, 16, 64), (1, 32, 64), (1, 64, 64)]


@pytest.fixture(params=SHAPES, ids=lambda mt: "map%d-ppw%d-tile%d" % mt)
def shape(gpu, request):
    """Every kernel shape (mapping, particles per wavefront, cloud tile) must
    give the same bits."""
    m0, t0, c0 = gpu.get_mapping(), gpu.get_tile(), gpu.get_cloud_tile()
    gpu.set_mapping(request.param[0])
    gpu.set_tile(request.param[1])
    gpu.set_cloud_tile(request.param[2])
    _say("cloud-weight shape mapping=%d ppw=%d tile=%d" % request.param)
    yield request.param
    gpu.set_mapping(m0)
    gpu.set_tile(t0)
    gpu.set_cloud_tile(c0)


@pytest.mark.parametrize("tile", [32, 64, 128, 256])
def test_cloud_weights_ragged_clouds_every_tile(gpu, room, tile):
    """Clouds shorter than a tile, one point either side of a tile boundary,
    and several tiles plus a remainder: every tile size gives the oracle's
    bits (the points past the last whole tile are read outside LDS)."""
    s, p = room["step"], room["particles"][:300]
    saved = gpu.get_mapping(), gpu.get_tile(), gpu.get_cloud_tile()
    try:
        gpu.set_mapping(gpu.MAP_PARTICLE_LANES)
        gpu.set_cloud_tile(tile)
        for ppw in (64, 16):
            gpu.set_tile(ppw)
            for m in (1, 11, 12, tile - 1, tile, tile + 1, 2 * tile + 7, 1999):
                cloud = s.cloud[:m]
                w = room["grid"].cloud_weights(cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
                ow = oracle.cloud_weight_batch(room["ogrid"], cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
                assert np.array_equal(w.view(np.int32), ow.view(np.int32)), (tile, ppw, m)
    finally:
        gpu.set_mapping(saved[0])
        gpu.set_tile(saved[1])
        gpu.set_cloud_tile(saved[2])


def test_cloud_weights_bit_exact(gpu, room, shape):
    s, p = room["step"], room["particles"]
    w = room["grid"].cloud_weights(s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))
    # ragged cloud sizes: the unrolled main loop plus every tail length
    for m in (11, 12, 37, 64 * 8 + 5):
        w = room["grid"].cloud_weights(s.cloud[:m], p[:200, 0], p[:200, 1], p[:200, 2], p[:200, 3])
        ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud[:m], p[:200, 0], p[:200, 1], p[:200, 2], p[:200, 3])
        assert np.array_equal(w.view(np.int32), ow.view(np.int32)), m


def test_cloud_weights_roll_pitch_and_edges(gpu, room, shape):
    rng = np.random.default_rng(11)
    s = room["step"]
    n = 300
    p = W.particles_around(room["pose"], n, rng, dxyz=2.0, dyaw=3.0)
    p[:20, 0] = -3.0  # pose outside the map: still weighted (computeCloudWeight has no map check)
    for roll, pitch in ((0.05, -0.02), (-0.3, 0.2)):
        w = room["grid"].cloud_weights(s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        assert np.array_equal(w.view(np.int32), ow.view(np.int32))
    # points exactly on cell faces (yaw 0, lattice translation): the division fallback
    bm = room["map"]
    lat = (np.stack(np.meshgrid(np.arange(-20, 20), np.arange(-20, 20), np.arange(-5, 5)), -1).reshape(-1, 3) *
           bm.res).astype(F32)
    tx = np.array([10.0, 10.05, 9.95, 0.0, 19.999], F32)
    ty = np.array([10.0, 10.0, 10.1, 0.0, 0.001], F32)
    tz = np.array([3.0, 3.0, 2.9, 0.0, 5.999], F32)
    ya = np.zeros(5, F32)
    w = room["grid"].cloud_weights(lat, tx, ty, tz, ya)
    ow = oracle.cloud_weight_batch(room["ogrid"], lat, tx, ty, tz, ya)
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))
    # <= 10 in-map points -> 0; nothing in the map -> 0
    w = room["grid"].cloud_weights(s.cloud[:10], p[:5, 0], p[:5, 1], p[:5, 2], p[:5, 3])
    assert np.all(w == 0)
    w = room["grid"].cloud_weights(s.cloud + 100, p[:5, 0], p[:5, 1], p[:5, 2], p[:5, 3])
    assert np.all(w == 0)
    # non-finite points are rejected like the reference's NaN compares
    c = s.cloud.copy()
    c[::7, 2] = np.nan
    c[::11, 0] = np.inf
    w = room["grid"].cloud_weights(c, p[:64, 0], p[:64, 1], p[:64, 2], p[:64, 3])
    ow = oracle.cloud_weight_batch(room["ogrid"], c, p[:64, 0], p[:64, 1], p[:64, 2], p[:64, 3])
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_cloud_weights_64bit_index_path(gpu, room, shape, monkeypatch):
    monkeypatch.setenv("A3D_FORCE_IDX64", "1")
    bm = room["map"]
    g = gpu.Grid3d().from_cells(room["oprob"], bm.min, bm.max, bm.res, room["sensor_dev"])
    s, p = room["step"], room["particles"]
    w = g.cloud_weights(s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_cloud_weights_32bit_index_64bit_address_path(gpu, room, shape, monkeypatch):
    """The gather of grids of 2^30 to 2^32 - 2 cells (32-bit cell indices,
    64-bit addresses), forced on the room map."""
    monkeypatch.setenv("A3D_FORCE_IDX32", "1")
    bm = room["map"]
    g = gpu.Grid3d().from_cells(room["oprob"], bm.min, bm.max, bm.res, room["sensor_dev"])
    s, p = room["step"], room["particles"]
    w = g.cloud_weights(s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def _segment_waves(bm, cloud, tz, per_wave, bits):
    """Which wavefronts of the poses kernel (per_wave consecutive poses each)
    read through segment loads under A3D_SEG_TEST_BITS=bits, with roll =
    pitch = 0: the device's decision restated.  The cells of a wavefront lie
    in layers [iz(lo + oz_min) - 1, iz(hi + oz_max) + 1] (lo, hi: the cloud's
    finite z range; oz = tz - min z), and it takes the segment loads when
    those layers' cells lie in one 2^bits-cell block."""
    z = cloud[:, 2][np.isfinite(cloud[:, 2])].astype(np.float64)
    lo, hi = z.min(), z.max()
    size = [int(v) for v in bm.size]
    layer = size[0] * size[1]
    inv = 1.0 / bm.res
    out = []
    for w0 in range(0, tz.size, per_wave):
        oz = tz[w0:w0 + per_wave].astype(np.float64) - np.float64(bm.min[2])
        qlo = np.float64(np.float32(lo + oz.min()))
        qhi = np.float64(np.float32(hi + oz.max()))
        zl = max(np.floor(qlo * inv) - 1.0, 0.0)
        zh = min(np.floor(qhi * inv) + 1.0, size[2] - 1.0)
        if zl > zh:
            out.append(True)
            continue
        a, b = int(zl) * layer, (int(zh) + 1) * layer - 1
        out.append((a >> bits) == (b >> bits))
    return np.array(out)


@pytest.mark.parametrize("bits", [19, 20])
def test_cloud_weights_segment_loads_mixed_wavefronts(gpu, room, shape, monkeypatch, bits):
    """Segment loads (a structured buffer based at the 2^30-cell segment that
    holds all of a wavefront's cells) on the room map through the 32-bit
    index kernels, with a test block of 2^bits cells so that some wavefronts
    take them and the others the 64-bit addresses: bit-exact against the
    oracle.  A band of the scan 0.5 m high and poses sorted by height keep
    each wavefront's cells within a few layers."""
    monkeypatch.setenv("A3D_FORCE_IDX32", "2")
    monkeypatch.setenv("A3D_SEG_TEST_BITS", str(bits))
    bm = room["map"]
    g = gpu.Grid3d().from_cells(room["oprob"], bm.min, bm.max, bm.res, room["sensor_dev"])
    s = room["step"]
    c = np.ascontiguousarray(s.cloud[(s.cloud[:, 2] > -0.3) & (s.cloud[:, 2] < 0.2)])
    rng = np.random.default_rng(40 + bits)
    p = W.particles_around(room["pose"], 1280, rng, dxyz=1.5, dyaw=3.0)
    p = p[np.argsort(p[:, 2], kind="stable")]
    w = g.cloud_weights(c, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    ow = oracle.cloud_weight_batch(room["ogrid"], c, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert (ow > 0).mean() > 0.5
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))
    if shape[0] == gpu.MAP_PARTICLE_LANES:
        seg = _segment_waves(bm, c, p[:, 2], shape[1], bits)
        _say("segment loads in %d of %d wavefronts" % (seg.sum(), seg.size))
        assert seg.any() and not seg.all()
    # with roll and pitch the rotated z mixes x and y in (the range the
    # wavefronts are bounded by is that of px*r20 + py*r21 + pz*r22)
    w = g.cloud_weights(c, p[:, 0], p[:, 1], p[:, 2], p[:, 3], 0.05, -0.03)
    ow = oracle.cloud_weight_batch(room["ogrid"], c, p[:, 0], p[:, 1], p[:, 2], p[:, 3], 0.05, -0.03)
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_cloud_weights_64bit_index_segment_loads(gpu, room, shape, monkeypatch):
    """The 64-bit index kernels with their segment loads (every wavefront of
    the room map is in segment 0), roll and pitch included."""
    monkeypatch.setenv("A3D_FORCE_IDX64", "2")
    bm = room["map"]
    g = gpu.Grid3d().from_cells(room["oprob"], bm.min, bm.max, bm.res, room["sensor_dev"])
    s, p = room["step"], room["particles"]
    for roll, pitch in ((0.0, 0.0), (0.05, -0.02)):
        w = g.cloud_weights(s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3], roll, pitch)
        assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_cloud_weights_window_test_on_clean_map(gpu, room, shape, monkeypatch):
    """The room's 0.1 m grid is clean (the reciprocal product floors like the
    division for every float), so it normally runs without the window test;
    the knob forces the window kernels on it."""
    monkeypatch.setenv("A3D_FORCE_NEAR_TEST", "1")
    bm = room["map"]
    g = gpu.Grid3d().from_cells(room["oprob"], bm.min, bm.max, bm.res, room["sensor_dev"])
    s, p = room["step"], room["particles"]
    w = g.cloud_weights(s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    ow = oracle.cloud_weight_batch(room["ogrid"], s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


@pytest.fixture(scope="module")
def long_axis(gpu):
    """A map 15.9 million cells long in x (one cell in y and z) whose resolution
    r = RN(1120.125 / 1354172) makes q / r land on an integer for q = k * 1120.125
    while q * RN(1/r) rounds just below it: the cells where a product with the
    reciprocal floors one lower than the reference's division.  Cell values are
    signed (a map loaded from cells may hold any float)."""
    res = float(np.float64(F32(1120.125)) / 1354172)
    mx = (15_900_000 * res, res, res)
    size = oracle.grid_sizes((0.0, 0.0, 0.0), mx, res)
    assert size[0] < (1 << 24) and size[1] == 1 and size[2] == 1
    prob = np.random.default_rng(21).normal(size=int(size[0])).astype(F32)
    q = (np.arange(1, 12, dtype=np.float64) * 1120.125).astype(F32)
    d_fast, d_exact = np.floor(q.astype(np.float64) * (1.0 / res)), np.floor(q.astype(np.float64) / res)
    # the product floors lower beyond 5e6, where a fixed 1e-9 window around
    # integers would no longer cover the rounding of q * (1/r)
    assert np.any((d_fast != d_exact) & (d_exact > 5e6))
    rng = np.random.default_rng(22)
    rand = rng.uniform(2.0, mx[0] - 2.0, 53).astype(F32)
    qx = np.concatenate([q, rand, q[::-1]])
    # pose (1, r/2, r/2), yaw 0: sx = px, so q = (float)(px + 1.0)
    cloud = np.zeros((qx.shape[0], 3), F32)
    cloud[:, 0] = qx - F32(1.0)
    assert np.array_equal((cloud[:, 0].astype(np.float64) + 1.0).astype(F32), qx)
    g = gpu.Grid3d().from_cells(prob, (0.0, 0.0, 0.0), mx, res, 0.05)
    return dict(res=res, mx=mx, prob=prob, cloud=cloud, grid=g,
                ogrid=oracle.GridMeta((0.0, 0.0, 0.0), mx, res, prob))


def _long_axis_poses(res, n=96):
    tx = np.full(n, 1.0, F32)
    tx[n // 2:] = np.random.default_rng(23).uniform(0.5, 2.0, n - n // 2).astype(F32)
    t = np.full(n, F32(res / 2), F32)
    return tx, t, t.copy(), np.zeros(n, F32)


def test_cloud_weights_long_axis_near_integer_cells(gpu, long_axis, shape):
    la = long_axis
    tx, ty, tz, ya = _long_axis_poses(la["res"])
    w = la["grid"].cloud_weights(la["cloud"], tx, ty, tz, ya)
    ow = oracle.cloud_weight_batch(la["ogrid"], la["cloud"], tx, ty, tz, ya)
    assert np.all(ow[:48] != 0)
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


def test_cloud_weights_long_axis_64bit_index_path(gpu, long_axis, monkeypatch):
    monkeypatch.setenv("A3D_FORCE_IDX64", "1")
    la = long_axis
    g = gpu.Grid3d().from_cells(la["prob"], (0.0, 0.0, 0.0), la["mx"], la["res"], 0.05)
    tx, ty, tz, ya = _long_axis_poses(la["res"])
    w = g.cloud_weights(la["cloud"], tx, ty, tz, ya)
    ow = oracle.cloud_weight_batch(la["ogrid"], la["cloud"], tx, ty, tz, ya)
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


@pytest.mark.parametrize("mn", [0.0, 1e-50])
def test_cloud_weights_zero_offsets(gpu, shape, mn):
    """Poses at (or within 1e-45 of) the map's minimum corner: the offsets
    t - min are 0, -0 or tiny, and a point at the origin gives q = -0.0 or a
    tiny negative double that rounds to -0.0f, which the reference accepts
    (q >= 0) into cell 0."""
    res = 0.25
    mnv = (mn, mn, mn)
    mx = (8.0, 8.0, 8.0)
    size = oracle.grid_sizes(mnv, mx, res)
    prob = np.random.default_rng(24).uniform(0.0, 1.0, int(np.prod(size))).astype(F32)
    g = gpu.Grid3d().from_cells(prob, mnv, mx, res, 0.05)
    og = oracle.GridMeta(mnv, mx, res, prob)
    rng = np.random.default_rng(25)
    base = rng.uniform(0.0, 7.9, (40, 3)).astype(F32)
    special = np.array([[0.0, 0.0, 0.0], [-0.0, -0.0, -0.0], [-0.0, 0.0, -0.0], [1e-45, -1e-45, 0.0],
                        [-1e-45, 0.0, 0.0], [0.0, 0.125, 0.0], [-1e-30, 0.0, 0.0]], F32)
    cloud = np.concatenate([special, base, special])
    t = np.array([0.0, -0.0, 1e-45, -1e-45, 0.5, 2.0], F32)
    tx, ty, tz = (a.reshape(-1) for a in np.meshgrid(t, t, t, indexing="ij"))
    ya = np.zeros(tx.shape[0], F32)
    w = g.cloud_weights(cloud, tx, ty, tz, ya)
    ow = oracle.cloud_weight_batch(og, cloud, tx, ty, tz, ya)
    assert np.array_equal(w.view(np.int32), ow.view(np.int32))


# ---------------------------------------------------------------- the filter
def test_init_bit_exact(gpu):
    rng = np.random.default_rng(2)
    for n in (1, 7, 1000, 5000):
        z = rng.standard_normal((n, 4)).astype(F32)
        pf = gpu.ParticleFilter()
        assert not pf.isInitialized()
        pf.init(n, 2.87080164308, 1.63820032984, 0.21460881682, 0.144, 0.05, 0.05, 0.05, 0.1, z)
        assert pf.isInitialized()
        op, om = oracle.pf_init(n, (2.87080164308, 1.63820032984, 0.21460881682, 0.144), (0.05, 0.05, 0.05, 0.1), z)
        assert np.array_equal(pf.particles().view(np.int32), op.view(np.int32)), n
        assert np.array_equal(pf.getMean().view(np.int32), om.view(np.int32)), n
        # p_.resize(abs(num_particles)) (ParticleFilter.cpp:51): -n gives the same filter
        pf2 = gpu.ParticleFilter()
        pf2.init(-n, 2.87080164308, 1.63820032984, 0.21460881682, 0.144, 0.05, 0.05, 0.05, 0.1, z)
        assert np.array_equal(pf2.particles().view(np.int32), op.view(np.int32)), -n


def test_predict_keeps_mean_and_is_bit_exact(gpu):
    # ParticleFilterTest.cpp:39-72: predict does not change getMean()
    rng = np.random.default_rng(3)
    n = 600
    z = rng.standard_normal((n, 4)).astype(F32)
    pf = gpu.ParticleFilter()
    pf.init(n, 2.87080164308, 1.63820032984, 0.21460881682, 0.144, 0.05, 0.05, 0.05, 0.1, z)
    m0 = pf.getMean()
    before = pf.particles()
    z2 = rng.standard_normal((n, 4)).astype(F32)
    pf.predict(0.1, 0.1, 0.1, 0.3, -0.067421, -0.006161, 0.130909, 0.052421, z2)
    assert np.array_equal(pf.getMean(), m0)
    op = oracle.pf_predict(before, [0.1, 0.1, 0.1, 0.3], [-0.067421, -0.006161, 0.130909, 0.052421], z2)
    assert np.array_equal(pf.particles().view(np.int32), op.view(np.int32))


def _full_step_case(gpu, room, seed, n=1000, out_of_map=0, nb=3):
    rng = np.random.default_rng(seed)
    bm = room["map"]
    pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), rng.uniform(-np.pi, np.pi))
    cloud = W.sensor_cloud(bm, pose, 2000, 0.05, rng)
    ranges = W.beacons(bm, pose, nb, 0.1, rng)
    parts = W.particles_around(pose, n, rng)
    parts[:out_of_map, 2] = -1.0
    parts[:, 5] = rng.uniform(0, 1, n)  # stale wp/wr values are part of the state
    parts[:, 6] = rng.uniform(0, 1, n)
    z4 = rng.standard_normal((n, 4)).astype(F32)
    return pose, cloud, ranges, parts, z4, float(rng.uniform())


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_full_step_bit_exact(gpu, room, seed):
    _say("full step seed", seed)
    pose, cloud, ranges, parts, z4, u01 = _full_step_case(gpu, room, seed, out_of_map=37 if seed == 2 else 0,
                                                          nb=0 if seed == 3 else 3)
    mods, deltas = [0.1, 0.1, 0.1, 0.3], [0.05, -0.03, 0.01, 0.02]
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    pf.predict(*mods, *deltas, z4)
    pf.update(room["grid"], cloud, ranges, 0.8, 0.1, 0.0, 0.0)
    got = pf.particles()
    mean = pf.getMean()
    op = oracle.pf_predict(parts, mods, deltas, z4)
    op, om = oracle.pf_update(op, room["ogrid"], cloud, ranges, 0.8, 0.1, 0.0, 0.0)
    rel = _rel(got[:, 4], op[:, 4])[op[:, 4] > 0]
    _say("seed", seed, "max rel err of w:", rel.max() if rel.size else 0.0)
    assert rel.size == 0 or rel.max() <= 1e-5  # the stated tolerance
    assert np.array_equal(got.view(np.int32), op.view(np.int32))  # and in fact bit-exact
    assert np.array_equal(mean.view(np.int32), om.view(np.int32))
    for mode in (gpu.RESAMPLE_REFERENCE, gpu.RESAMPLE_FP64):
        pf2 = gpu.ParticleFilter()
        pf2.set_particles(got)
        idx = pf2.resample(u01, mode, return_indices=True)
        oq, oidx = oracle.pf_resample(op, u01, mode)
        assert np.array_equal(idx, oidx), mode
        assert np.array_equal(pf2.particles().view(np.int32), oq.view(np.int32))
    ref_idx = oracle.resample_indices(op[:, 4], u01, 0)
    f64_idx = oracle.resample_indices(op[:, 4], u01, 1)
    _say("seed", seed, "resample indices differing between float and FP64 prefix:", int((ref_idx != f64_idx).sum()))


@pytest.mark.parametrize("n", [3000, 30000])
def test_update_mean_of_mixed_sign_terms(gpu, room, n):
    """Yaw spread over (-pi, pi]: the mean's yaw sum has terms of both signs,
    whose running sum crosses float binades again and again (the device's
    exact parallel sum restarts at each crossing and goes serial after many);
    the mean must still be the reference's sequential float sum."""
    rng = np.random.default_rng(41 + n)
    bm = room["map"]
    pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), 0.0)
    cloud = W.sensor_cloud(bm, pose, 1500, 0.05, rng)
    ranges = W.beacons(bm, pose, 2, 0.1, rng)
    parts = W.particles_around(pose, n, rng, dxyz=3.0)
    parts[:, 3] = rng.uniform(-np.pi, np.pi, n).astype(F32)
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    pf.update(room["grid"], cloud, ranges, 0.8, 0.1, 0.0, 0.0)
    op, om = oracle.pf_update(parts, room["ogrid"], cloud, ranges, 0.8, 0.1, 0.0, 0.0)
    assert np.array_equal(pf.particles().view(np.int32), op.view(np.int32))
    assert np.array_equal(pf.getMean().view(np.int32), om.view(np.int32))
    idx = pf.resample(0.61, gpu.RESAMPLE_REFERENCE, return_indices=True)
    assert np.array_equal(idx, oracle.resample_indices(op[:, 4], 0.61, 0))


@pytest.mark.parametrize("mapping,ppw", [(0, 64), (0, 16), (1, 32)])
def test_update_visiting_order_invariance(gpu, room, mapping, ppw):
    """The particle visiting order (index or spatial) and the kernel shape
    change nothing: the update stays bit-exact."""
    pose, cloud, ranges, parts, z4, u01 = _full_step_case(gpu, room, 41, n=3001, out_of_map=5)
    op, om = oracle.pf_update(parts, room["ogrid"], cloud, ranges, 0.8, 0.1)
    saved = gpu.get_mapping(), gpu.get_tile(), gpu.get_order()
    try:
        gpu.set_mapping(mapping)
        gpu.set_tile(ppw)
        for order in (gpu.ORDER_INDEX, gpu.ORDER_SPATIAL):
            _say("update with mapping", mapping, "ppw", ppw, "order", order)
            gpu.set_order(order)
            pf = gpu.ParticleFilter()
            pf.set_particles(parts)
            pf.update(room["grid"], cloud, ranges, 0.8, 0.1)
            assert np.array_equal(pf.particles().view(np.int32), op.view(np.int32)), order
            assert np.array_equal(pf.getMean().view(np.int32), om.view(np.int32)), order
    finally:
        gpu.set_mapping(saved[0])
        gpu.set_tile(saved[1])
        gpu.set_order(saved[2])


def test_update_roll_pitch_and_no_map(gpu, room):
    pose, cloud, ranges, parts, z4, u01 = _full_step_case(gpu, room, 21)
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    pf.update(room["grid"], cloud, ranges, 0.5, 0.53, 0.07, -0.04)
    op, om = oracle.pf_update(parts, room["ogrid"], cloud, ranges, 0.5, 0.53, 0.07, -0.04)
    assert np.array_equal(pf.particles().view(np.int32), op.view(np.int32))
    assert np.array_equal(pf.getMean().view(np.int32), om.view(np.int32))
    pf.set_particles(parts)
    pf.update(None, cloud, ranges, 0.5, 0.53)
    op, om = oracle.pf_update(parts, oracle.GridMeta(room["map"].min, room["map"].max, 0.1, None), cloud, ranges,
                              0.5, 0.53)
    assert np.array_equal(pf.particles().view(np.int32), op.view(np.int32))
    assert np.all(pf.getMean() == 0)


def test_large_filter_sampled(gpu, room):
    # 100,000 particles x 20,000 points: sampled particles against the oracle,
    # and size-independent properties over all of them
    rng = np.random.default_rng(77)
    bm = room["map"]
    pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), 0.3)
    n, m = 100000, 20000
    cloud = W.sensor_cloud(bm, pose, m, 0.03, rng)
    ranges = W.beacons(bm, pose, 8, 0.1, rng)
    parts = W.particles_around(pose, n, rng)
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    _say("large update n=1e5 m=2e4")
    pf.update(room["grid"], cloud, ranges, 0.8, 0.1)
    got = pf.particles()
    sel = rng.choice(n, 300, replace=False)
    ow = oracle.cloud_weight_batch(room["ogrid"], cloud, parts[sel, 0], parts[sel, 1], parts[sel, 2],
                                   parts[sel, 3])
    # wp after normalisation is wp_raw / wtp: compare the ratios of sampled pairs
    raw = got[sel, 5].astype(np.float64)
    ratio = raw / max(raw.max(), 1e-30)
    oratio = ow.astype(np.float64) / max(ow.max(), 1e-30)
    assert np.allclose(ratio, oratio, rtol=1e-6, atol=1e-9)
    assert abs(got[:, 4].astype(np.float64).sum() - 1.0) < 1e-3
    idx = pf.resample(0.37, gpu.RESAMPLE_REFERENCE, return_indices=True)
    assert np.all(np.diff(idx.astype(np.int64)) >= 0) and idx.max() < n
    oidx = oracle.resample_indices(got[:, 4], 0.37, 0)
    assert np.array_equal(idx, oidx)


def _adversarial_weights(case, n, rng):
    """Weight vectors that stress a parallel restatement of the float running
    sum: binade crossings, exact ties, zero runs, subnormals, huge jumps, terms
    of both signs."""
    if case == "lognormal":
        return np.exp(rng.uniform(np.log(1e-30), np.log(1e10), n)).astype(F32)
    if case == "ties":  # 1 + k 2^-24 ... : adding 2^-24 to a sum in [1, 2) is an exact tie
        w = np.full(n, 2.0 ** -24, F32)
        w[0] = 1.0
        w[1::7] = 3 * 2.0 ** -24
        w[2::11] = 2.0 ** -23
        return w
    if case == "zero_runs":
        w = rng.uniform(0, 1, n).astype(F32)
        w[: n // 3] = 0.0
        w[n // 2: n // 2 + 5000] = 0.0
        w[n // 3 + 1::5] = -0.0
        return w
    if case == "subnormal":
        w = (rng.uniform(0, 1, n) * 1e-41).astype(F32)
        w[n // 2:] = rng.uniform(0, 1e-3, n - n // 2).astype(F32)
        return w
    if case == "jumps":
        w = rng.uniform(0, 1e-6, n).astype(F32)
        w[::997] = 1e6
        w[min(5, n - 1)] = 3e30
        return w
    if case == "signed":
        return rng.normal(0, 1, n).astype(F32)
    if case == "powers":  # sums landing exactly on powers of two
        w = np.full(n, 1.0, F32)
        w[1::3] = 0.5
        w[2::3] = 0.5
        return w
    raise ValueError(case)


@pytest.mark.parametrize("case", ["lognormal", "ties", "zero_runs", "subnormal", "jumps", "signed", "powers"])
def test_resample_running_sum_adversarial(gpu, case):
    """The running sum c of resample (and through it every index) on weights
    chosen to exercise each branch of the device's exact parallel sum."""
    rng = np.random.default_rng(hash(case) % 2 ** 32)
    for n in (1, 63, 64, 65, 4095, 4097, 100_000):
        w = _adversarial_weights(case, n, rng)
        p = np.zeros((n, 7), F32)
        p[:, 0] = np.arange(n)
        p[:, 4] = w
        pf = gpu.ParticleFilter()
        pf.set_particles(p)
        for u01 in (0.0, 0.37, 0.999):
            pf.set_particles(p)
            idx = pf.resample(u01, gpu.RESAMPLE_REFERENCE, return_indices=True)
            oidx = oracle.resample_indices(w, u01, 0)
            assert np.array_equal(idx, oidx), (case, n, u01, int((idx != oidx).sum()))
        pf.close()


def test_graph_capture_replays_match_eager(gpu, room):
    pose, cloud, ranges, parts, z4, u01 = _full_step_case(gpu, room, 31)
    mods, deltas = [0.1, 0.1, 0.1, 0.3], [0.05, -0.03, 0.01, 0.02]
    dz = gpu.DeviceBuffer.from_array(z4)
    x, y, z = (gpu.DeviceBuffer.from_array(np.ascontiguousarray(cloud[:, k])) for k in range(3))
    dr = gpu.DeviceBuffer.from_array(ranges)
    eager = gpu.ParticleFilter()
    eager.set_particles(parts)
    graphed = gpu.ParticleFilter()
    graphed.set_particles(parts)
    graphed.capture_begin()
    graphed.predict_dev(mods, deltas, dz)
    graphed.update_dev(room["grid"], x.ptr, y.ptr, z.ptr, cloud.shape[0], dr.ptr, ranges.shape[0], 0.8, 0.1)
    graphed.resample_dev(u01)
    graph = graphed.capture_end()
    for rep in range(3):
        eager.predict(*mods, *deltas, z4)
        eager.update(room["grid"], cloud, ranges, 0.8, 0.1)
        eager.resample(u01)
        graph.launch()
        graphed.synchronize()
        assert np.array_equal(graphed.particles().view(np.int32), eager.particles().view(np.int32)), rep
        assert np.array_equal(graphed.getMean(), eager.getMean()), rep
    graph.close()


def test_grid_room_every_cell_against_fp64(room):
    """computeGrid against an FP64 nearest neighbour (scipy cKDTree) on all
    2.4 million cells of the room map: the error of the float field, stated
    and bounded (the field itself equals the oracle's bit for bit)."""
    from scipy.spatial import cKDTree

    bm = room["map"]
    prob, dist = room["grid"].read_cells(dist=True)
    sx, sy, sz = (int(v) for v in room["grid"].size)
    i = np.arange(sx * sy * sz)
    ix, iy, iz = i % sx, (i // sx) % sy, i // (sx * sy)
    q = np.stack([bm.min[0] + ix * bm.res, bm.min[1] + iy * bm.res, bm.min[2] + iz * bm.res], -1)
    q = q.astype(F32).astype(np.float64)  # the sample point as the reference forms it (a float)
    d64, _ = cKDTree(room["points"].astype(np.float64)).query(q)
    e64 = d64 ** 2
    d2 = dist.astype(np.float64)
    abs_err = np.abs(d2 - e64)
    rel_err = abs_err / np.maximum(e64, 1e-12)
    sd = room["sensor_dev"]
    p64 = 1.0 / (sd * np.sqrt(2 * np.pi)) * np.exp(-e64 * e64 / (2 * sd * sd))
    dprob = np.abs(prob.astype(np.float64) - p64)
    _say("computeGrid vs FP64 over %d cells: dist2 max abs %.3g mean abs %.3g max rel %.3g mean rel %.3g; "
         "max |dprob| %.3g (peak prob %.3g)" % (i.size, abs_err.max(), abs_err.mean(), rel_err.max(), rel_err.mean(),
                                                dprob.max(), p64.max()))
    assert i.size >= 1_000_000
    assert rel_err.max() < 1e-5
    assert dprob.max() < 1e-5 * p64.max()
