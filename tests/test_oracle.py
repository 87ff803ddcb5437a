"""Pins of the oracle (the CPU restatement in oracle/), on the CPU.

Every oracle function is checked against something that does not restate it:
the reference's own known-answer values and fixtures (amcl3d/tests/*), closed
forms, brute force on tiny inputs, and an FP64 evaluation.
"""
import ctypes

import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

F32 = np.float32


# ----------------------------------------------------------- reference pins
def test_fixture_counts_and_bounds(kat):
    # PointCloudToolsTest.cpp:166-173: 179,551 occupied points and the bounds
    assert kat["map_cloud"].shape == (179551, 3)
    assert kat["map_points"] == 179551
    size = oracle.grid_sizes(kat["map_min"], kat["map_max"], kat["map_res"])
    assert list(size) == [522, 384, 153]
    # the nav slice is one z-plane of that grid (Grid3d.cpp:100-118)
    assert kat["nav"].shape == (384, 522)
    assert kat["particle_info_poses"] == 0  # computeCloudWeightParticlesTest asserts nothing


def test_kat_cloud_weight(kat, kat_oracle_grid):
    # Grid3dTest.cpp:128-132,167-168: computeCloudWeight at a known pose
    _, prob = kat_oracle_grid
    g = oracle.GridMeta(kat["map_min"], kat["map_max"], kat["map_res"], prob)
    p = kat["kat_pose"]
    w = oracle.cloud_weight(g, kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], p["roll"], p["pitch"], p["yaw"])
    assert abs(w - kat["kat_weight"]) <= kat["kat_tolerance"]
    assert abs(w - kat["kat_weight"]) / kat["kat_weight"] < 1e-5  # observed 1.6e-6
    # Grid3dTest.cpp:160-161: 0 before open
    g0 = oracle.GridMeta(kat["map_min"], kat["map_max"], kat["map_res"], None)
    assert oracle.cloud_weight(g0, kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], 0, 0, p["yaw"]) == 0.0


def test_kat_nav_slice(kat, kat_oracle_grid):
    # nav_msg.bin is slice iz=20 of the d^4 likelihood field scaled to int8
    _, prob = kat_oracle_grid
    s = oracle.slice_int8(prob[20]).reshape(kat["nav"].shape).astype(np.int32)
    d = s - kat["nav"].astype(np.int32)
    assert np.abs(d).max() <= 1
    assert (d == 0).mean() >= 0.999
    # every mismatch is an int8 truncation tie: prob*100/max within 1e-4 of an integer
    plane = prob[20].reshape(kat["nav"].shape)
    scale = np.float32(100.0) / np.float32(plane.max())
    v = (plane * scale)[d != 0]
    assert np.all(np.abs(v - np.round(v)) < 1e-4)
    # the textbook N(d; 0, sigma) field does not reproduce the slice
    dist, _ = kat_oracle_grid
    sd = kat["sensor_dev"]
    textbook = (1 / (sd * np.sqrt(2 * np.pi)) * np.exp(-dist[20].astype(np.float64) / (2 * sd * sd))).astype(F32)
    t = oracle.slice_int8(textbook).reshape(kat["nav"].shape).astype(np.int32)
    assert (t == kat["nav"]).mean() < 0.9


def test_is_into_map(kat):
    # Grid3dTest.cpp:247-265
    g = oracle.GridMeta(kat["map_min"], kat["map_max"], kat["map_res"], np.zeros(1, F32))
    assert oracle.is_into_map(g, *kat["isinto_in"])
    assert not oracle.is_into_map(g, *kat["isinto_out"])
    assert not oracle.is_into_map(g, *kat["isinto_in"], open_=False)
    # float vs double compares, half-open: (float)-17.35 lies below -17.35, (float)-17.3 inside
    assert float(F32(-17.35)) < -17.35
    assert not oracle.is_into_map(g, F32(-17.35), F32(0.0), F32(0.0))
    assert oracle.is_into_map(g, F32(-17.3), F32(-9.5), F32(-1.4))
    assert not oracle.is_into_map(g, 8.75, 0.0, 0.0)


# ------------------------------------------------------------- computeGrid
def _brute_d2(points, queries):
    """float32 brute force in FLANN L2_Simple order: ((dx*dx + dy*dy) + dz*dz)."""
    p = points.astype(F32)
    out = np.empty(queries.shape[0], F32)
    for i, q in enumerate(queries.astype(F32)):
        dx = q[0] - p[:, 0]
        dy = q[1] - p[:, 1]
        dz = q[2] - p[:, 2]
        d = dx * dx
        d = d + dy * dy
        d = d + dz * dz
        out[i] = d.min()
    return out


def test_grid_brute_force_tiny():
    rng = np.random.default_rng(7)
    mn = np.array([-1.3, 0.7, -0.35])
    mx = mn + np.array([1.2, 1.0, 0.8])
    res = 0.1
    pts = (mn + rng.uniform(-0.2, 1.3, (40, 3)) * (mx - mn)).astype(F32)
    dist, prob = oracle.compute_grid(pts, mn, mx, res, 0.05, exp_mode=1)
    sz = oracle.grid_sizes(mn, mx, res)
    iz, iy, ix = np.meshgrid(np.arange(sz[2]), np.arange(sz[1]), np.arange(sz[0]), indexing="ij")
    q = np.stack([(mn[0] + ix * res), (mn[1] + iy * res), (mn[2] + iz * res)], -1).reshape(-1, 3).astype(F32)
    ref = _brute_d2(pts, q)
    assert np.array_equal(dist.reshape(-1), ref)
    c1 = F32(1.0 / (0.05 * np.sqrt(2 * np.pi)))
    c2 = F32(1.0 / (2.0 * 0.05 * 0.05))
    e = (-ref * ref * c2).astype(F32)
    assert np.array_equal(prob.reshape(-1), (c1 * np.exp(e.astype(np.float64)).astype(F32)).astype(F32))


def test_grid_fp64_and_modes():
    # FP64 nearest distances (scipy) bound the float field; exp modes agree to 1 ulp
    from scipy.spatial import cKDTree

    c = W.room_case(2)
    bm = c["map"]
    pts = bm.points()
    dist, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=0)
    _, prob1 = oracle.compute_grid(pts, bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
    sz = bm.size
    rng = np.random.default_rng(0)
    lin = rng.integers(0, dist.size, 20000)
    ix, iy, iz = lin % sz[0], (lin // sz[0]) % sz[1], lin // (sz[0] * sz[1])
    q = np.stack([bm.min[0] + ix * bm.res, bm.min[1] + iy * bm.res, bm.min[2] + iz * bm.res], -1)
    q32 = q.astype(F32).astype(np.float64)
    d64, _ = cKDTree(pts.astype(np.float64)).query(q32)
    d2 = dist.reshape(-1)[lin].astype(np.float64)
    rel = np.abs(d2 - d64 ** 2) / np.maximum(d64 ** 2, 1e-12)
    assert rel.max() < 1e-5, rel.max()
    ulps = np.abs(prob.reshape(-1).view(np.int32).astype(np.int64) - prob1.reshape(-1).view(np.int32))
    assert ulps.max() <= 1
    assert (ulps != 0).mean() < 1e-3


def test_grid_empty_cloud():
    dist, prob = oracle.compute_grid(np.zeros((0, 3), F32), np.zeros(3), np.ones(3), 0.25, 0.1)
    assert np.all(dist == -1.0) and np.all(prob == 0.0)


# ------------------------------------------------------ computeCloudWeight
def _index_grid(mn, mx, res):
    g = oracle.GridMeta(mn, mx, res)
    n = int(np.prod(g.size))
    g.prob = (np.arange(n, dtype=np.float64) * 2.0 ** -20).astype(F32)  # cell id is readable from prob
    return g


def _expected_weight(g, cloud, t, R):
    q = (R @ cloud.astype(np.float64).T).T + np.asarray(t, np.float64) - g.min
    ok = np.all((q >= 0) & (q < (g.max - g.min)), 1)
    cell = np.floor(q[ok] / g.res).astype(np.int64)
    lin = cell[:, 0] + g.size[0] * (cell[:, 1] + g.size[1] * cell[:, 2])
    return g.prob[lin].astype(np.float64).sum() / max(1, len(lin)), len(lin)


def _rot(roll, pitch, yaw):
    cr, sr, cp, sp, cy, sy = np.cos(roll), np.sin(roll), np.cos(pitch), np.sin(pitch), np.cos(yaw), np.sin(yaw)
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    return Rz @ Ry @ Rx


@pytest.mark.parametrize("roll,pitch,yaw", [(0, 0, 0), (0, 0, np.pi / 2), (0, 0, -2.3), (0.2, -0.3, 1.1)])
def test_cloud_weight_cells(roll, pitch, yaw):
    # The cell of each point is fixed by R = Rz(yaw) Ry(pitch) Rx(roll) and the
    # translation; points are kept away from cell faces so rounding cannot move them.
    mn, mx, res = np.array([-2.0, -1.5, 0.5]), np.array([3.0, 2.5, 3.5]), 0.1
    g = _index_grid(mn, mx, res)
    rng = np.random.default_rng(3)
    R = _rot(roll, pitch, yaw)
    t = np.array([0.43, 0.61, 1.97])
    cells = rng.integers([0, 0, 0], g.size, (500, 3))
    world = mn + (cells + rng.uniform(0.2, 0.8, (500, 3))) * res
    cloud = ((world - t) @ R).astype(F32)  # R^T (world - t)
    w = oracle.cloud_weight(g, cloud, F32(t[0]), F32(t[1]), F32(t[2]), F32(roll), F32(pitch), F32(yaw))
    exp, n = _expected_weight(g, cloud, t, R)
    assert n == 500
    assert abs(w - exp) <= 1e-6 * exp
    # the gathered cells themselves are the analytic ones
    cl = oracle.cloud_cells(g, cloud, [t[0]], [t[1]], [t[2]], [yaw], roll, pitch)[0]
    assert np.array_equal(cl, cells[:, 0] + g.size[0] * (cells[:, 1] + g.size[1] * cells[:, 2]))


def test_cloud_cells_sum_to_the_weight():
    """cloud_cells lists the cells cloud_weight adds: their prob summed in
    cloud order as floats, over more than 10 points, divided by the count, is
    the weight; points outside the map (here a third of them) are -1."""
    c = W.room_case(2, n=64, m=900)
    bm, s, p = c["map"], c["step"], c["particles"]
    rng = np.random.default_rng(4)
    g = oracle.GridMeta(bm.min, bm.max, bm.res, rng.uniform(0, 4, int(np.prod(bm.size))).astype(F32))
    cloud = s.cloud.copy()
    cloud[::3] += F32(40.0)
    w = oracle.cloud_weight_batch(g, cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    cells = oracle.cloud_cells(g, cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    assert np.all(cells[:, ::3] == -1) and (cells >= 0).sum() > 0.3 * cells.size
    for i in range(p.shape[0]):
        acc, n = F32(0.0), 0
        for c_ in cells[i][cells[i] >= 0]:
            acc = F32(acc + g.prob[c_])
            n += 1
        assert w[i] == (F32(acc / F32(n)) if n > 10 else 0.0), i


def test_cloud_weight_constant_grid_and_min_points():
    mn, mx, res = np.zeros(3), np.array([2.0, 2.0, 2.0]), 0.1
    g = oracle.GridMeta(mn, mx, res, np.full(20 * 20 * 20, 0.25, F32))
    rng = np.random.default_rng(1)
    cloud = rng.uniform(-0.5, 0.5, (300, 3)).astype(F32)
    assert oracle.cloud_weight(g, cloud, 1.0, 1.0, 1.0, 0, 0, 0.3) == 0.25
    # n <= 10 points in the map -> 0 (Grid3d.cpp:198)
    assert oracle.cloud_weight(g, cloud[:10], 1.0, 1.0, 1.0, 0, 0, 0.3) == 0.0
    assert oracle.cloud_weight(g, cloud[:11], 1.0, 1.0, 1.0, 0, 0, 0.3) == 0.25
    # all points outside -> 0
    assert oracle.cloud_weight(g, cloud + 10, 1.0, 1.0, 1.0, 0, 0, 0.3) == 0.0


def test_cloud_weight_fp64_error():
    # FP64 evaluation of the same weight: the float transform moves a point to a
    # neighbouring cell only when it lies within rounding of a cell face.
    c = W.room_case(3)
    bm, s, p = c["map"], c["step"], c["particles"]
    _, prob = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"])
    g = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    w = oracle.cloud_weight_batch(g, s.cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
    rel = []
    for i in range(0, p.shape[0], 50):
        e, n = _expected_weight(g, s.cloud, p[i, :3].astype(np.float64), _rot(0, 0, float(p[i, 3])))
        if n > 10 and e > 0:
            rel.append(abs(w[i] - e) / e)
    rel = np.array(rel)
    assert np.median(rel) < 1e-5 and rel.max() < 5e-3


# ------------------------------------------------------------- range weight
def test_range_weight():
    assert oracle.range_weight(1, 2, 3, np.zeros((0, 4)), 0.5) == 0.0
    sigma = 0.53
    k1 = F32(1.0 / (sigma * np.sqrt(2 * np.pi)))
    # exact range -> exp(0) = 1 -> k1
    assert oracle.range_weight(0, 0, 0, np.array([[5.0, 3.0, 4.0, 0.0]]), sigma) == k1
    rng = np.random.default_rng(5)
    for _ in range(50):
        x = rng.uniform(-5, 5, 3)
        r = np.concatenate([rng.uniform(0, 10, (4, 1)), rng.uniform(-5, 5, (4, 3))], 1).astype(F32)
        d = np.linalg.norm(x.astype(F32).astype(np.float64) - r[:, 1:].astype(np.float64), axis=1)
        exp = np.prod(1 / (sigma * np.sqrt(2 * np.pi)) * np.exp(-(d - r[:, 0]) ** 2 / (2 * sigma ** 2)))
        got = oracle.range_weight(F32(x[0]), F32(x[1]), F32(x[2]), r, sigma)
        assert abs(got - exp) <= 1e-5 * exp + 1e-30


# ------------------------------------------------------------------ filter
def test_init_closed_form():
    n = 64
    p, mean = oracle.pf_init(n, (1.5, -2.0, 0.5, 0.3), (0.05, 0.05, 0.05, 0.1), np.zeros((n, 4), F32))
    assert np.all(p[:, 0] == F32(1.5)) and np.all(p[:, 3] == F32(0.3))
    assert abs(p[:, 4].sum() - 1) < 1e-6
    assert np.allclose(mean, [1.5, -2.0, 0.5, 0.3], atol=1e-5)
    rng = np.random.default_rng(0)
    z = rng.standard_normal((n, 4)).astype(F32)
    devs = (0.05, 0.07, 0.03, 0.1)
    p, mean = oracle.pf_init(n, (1.5, -2.0, 0.5, 0.3), devs, z)
    # draw scaling: x_i = x0 + (float)(z * dev), particle 0 is the pose
    assert p[0, 0] == F32(1.5)
    assert np.array_equal(p[1:, 1], F32(-2.0) + z[1:, 1] * F32(0.07))
    # weights: Gaussian of the distance with dev = max(xdev, ydev, zdev), normalised
    d2 = ((p[:, :3].astype(np.float64) - p[0, :3]) ** 2).sum(1)
    w = np.exp(-d2 / (2 * 0.07 ** 2))
    assert np.allclose(p[:, 4], w / w.sum(), rtol=1e-5)
    assert np.allclose(mean, (p[:, :4].astype(np.float64) * (w / w.sum())[:, None]).sum(0), rtol=1e-5)


def test_predict_closed_form():
    n = 16
    p = np.zeros((n, 7), F32)
    p[:, 0] = 1.0
    p[:, 3] = 0.0
    mods = np.array([0.1, 0.1, 0.1, 0.3])
    deltas = np.array([0.2, -0.1, 0.05, 0.02])
    q = oracle.pf_predict(p, mods, deltas, np.zeros((n, 4), F32))
    # yaw 0, no noise: x += dx, y += dy, z += dz, a += da
    assert np.all(q[:, 0] == F32(1.0) + F32(0.2)) and np.all(q[:, 1] == F32(-0.1))
    assert np.all(q[:, 2] == F32(0.05)) and np.all(q[:, 3] == F32(0.02))
    # yaw pi/2: the step turns with the particle (x -= dy, y += dx)
    p[:, 3] = np.pi / 2
    q = oracle.pf_predict(p, mods, deltas, np.ones((n, 4), F32))
    sx = F32(abs(0.2 * 0.1))
    sy = F32(abs(-0.1 * 0.1))
    assert np.allclose(q[:, 0], 1.0 - (-0.1 + sy), atol=1e-6)
    assert np.allclose(q[:, 1], 0.2 + sx, atol=1e-6)
    assert np.allclose(q[:, 3], np.pi / 2 + 0.02 + 0.02 * 0.3, atol=1e-6)


def test_update_properties():
    c = W.room_case(4)
    bm, s, p = c["map"], c["step"], c["particles"].copy()
    _, prob = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"])
    g = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    p[:50, 0] = -5.0  # outside the map: w = 0, wp/wr keep their old values until normalised
    p[:50, 5] = 0.7
    q, mean = oracle.pf_update(p, g, s.cloud, s.ranges, s.alpha, s.sigma)
    assert np.all(q[:50, 4] == 0)
    assert abs(q[:, 4].astype(np.float64).sum() - 1) < 1e-5
    # the blend in FP64 from the oracle's own terms
    wp = q[:, 5].astype(np.float64)
    wr = q[:, 6].astype(np.float64)
    w = np.where(np.arange(len(q)) < 50, 0.0, wp * s.alpha + wr * (1 - s.alpha))
    assert np.allclose(q[:, 4], w / w.sum(), rtol=1e-5, atol=1e-12)
    assert np.allclose(mean, (q[:, :4].astype(np.float64) * q[:, 4:5]).sum(0), rtol=1e-5)
    # stale wp of out-of-map particles is divided by wtp like the others (ParticleFilter.cpp:162-165)
    assert np.all(q[:50, 5] == q[0, 5]) and q[0, 5] > 0 and q[0, 5] != F32(0.7)
    # no map -> everything 0
    q0, m0 = oracle.pf_update(p, oracle.GridMeta(bm.min, bm.max, bm.res, None), s.cloud, s.ranges, 0.8, 0.1)
    assert np.all(q0[:, 4] == 0) and np.all(m0 == 0) and np.all(q0[:, 5] == 0)


def test_resample_closed_forms():
    n = 8
    idx = oracle.resample_indices(np.full(n, 1 / n, F32), 0.5)
    assert list(idx) == list(range(n))  # equal weights: each particle once
    w = np.zeros(n, F32)
    w[5] = 1.0
    assert np.all(oracle.resample_indices(w, 0.3) == 5)  # one-hot: all copies of it
    assert np.all(oracle.resample_indices(np.zeros(n, F32), 0.3) == n - 1)  # clamped
    rng = np.random.default_rng(9)
    for mode in (0, 1):
        w = rng.uniform(0, 1, 1000).astype(F32)
        w /= w.sum()
        idx = oracle.resample_indices(w, float(rng.uniform()), mode)
        assert np.all(np.diff(idx.astype(np.int64)) >= 0)
        counts = np.bincount(idx, minlength=1000)
        assert np.all(np.abs(counts - 1000 * w) <= 1.0 + 1e-3)  # low-variance property
        assert list(np.unique(oracle.resample_indices(np.full(1024, 1 / 1024, F32), 0.5, mode))) == list(range(1024))
    p = np.zeros((n, 7), F32)
    p[:, 0] = np.arange(n)
    p[:, 4] = w[:n] / w[:n].sum()
    q, idx = oracle.pf_resample(p, 0.4)
    assert np.all(q[:, 0] == idx) and np.all(q[:, 4] == F32(1) / F32(n))


# ---------------------------------------------------------------- VoxelGrid
def test_voxel_grid_closed_forms():
    """pcl::VoxelGrid restated (Node.cpp:130-141): pins from the definition."""
    leaf = 0.1
    # one point per voxel: the output is the input, in voxel-index order (x fastest)
    cells = np.array([[3, 1, 0], [0, 0, 0], [1, 0, 0], [0, 2, 1], [2, 2, 2]])
    pts = ((cells + 0.5) * leaf).astype(F32)
    out = oracle.voxel_grid(pts, leaf)
    key = cells[:, 0] + 4 * (cells[:, 1] + 3 * cells[:, 2])  # div = (4, 3, 3) from the bounds
    assert np.array_equal(out, pts[np.argsort(key)])
    # two points in a voxel: their float mean (a sum of two is order-free)
    p2 = np.array([[0.01, 0.02, 0.03], [0.05, 0.07, 0.09], [0.31, 0.0, 0.0]], F32)
    out = oracle.voxel_grid(p2, leaf)
    assert out.shape == (2, 3)
    assert np.array_equal(out[0], ((p2[0] + p2[1]) / F32(2)).astype(F32))
    # non-finite points are skipped; an empty or all-NaN cloud gives nothing
    p3 = np.vstack([p2, [[np.nan, 0, 0], [0, np.inf, 0]]]).astype(F32)
    assert np.array_equal(oracle.voxel_grid(p3, leaf), out)
    assert oracle.voxel_grid(np.zeros((0, 3), F32), leaf).shape == (0, 3)
    assert oracle.voxel_grid(np.full((4, 3), np.nan, F32), leaf).shape == (0, 3)
    # more voxels than int32 indices: the input comes back unchanged
    far = np.array([[0, 0, 0], [1000, 1000, 1000], [1, 2, 3]], F32)
    assert np.array_equal(oracle.voxel_grid(far, 1e-3), far)


def test_voxel_grid_against_brute_force():
    # every voxel's centroid equals the mean of its points up to float
    # summation order; the voxel set is exactly the set of occupied voxels
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2.0, 3.0, (20000, 3)).astype(F32)
    leaf = np.float32(0.25)
    out = oracle.voxel_grid(pts, leaf)
    inv = np.float32(1) / leaf
    ijk = np.floor(pts * inv).astype(np.int64)
    uniq, inv_idx, counts = np.unique(ijk, axis=0, return_inverse=True, return_counts=True)
    assert out.shape[0] == uniq.shape[0]
    mean = np.zeros((uniq.shape[0], 3))
    np.add.at(mean, inv_idx.reshape(-1), pts.astype(np.float64))
    mean /= counts[:, None]
    # output order is increasing voxel index, x fastest
    lo = ijk.min(0)
    div = ijk.max(0) - lo + 1
    key = (uniq[:, 0] - lo[0]) + div[0] * ((uniq[:, 1] - lo[1]) + div[1] * (uniq[:, 2] - lo[2]))
    assert np.allclose(out, mean[np.argsort(key)], rtol=1e-6, atol=1e-6)
    # PCL's sort order and point order differ only in the last bits
    out1 = oracle.voxel_grid(pts, leaf, order_mode=1)
    ulps = np.abs(out.view(np.int32).astype(np.int64) - out1.view(np.int32))
    assert ulps.max() <= 8


def test_libm_choice_effect_on_room_steps():
    """Which libm the reference's float calls resolve to (expf for prob,
    PointCloudTools.cpp:137; sin/cos of the yaw, Grid3d.cpp:139-144 and
    ParticleFilter.cpp:110-111) is fixed by its GCC 5 / glibc platform, not
    by its source.  Measured here: how often glibc's expf and sinf/cosf differ
    from the double routine rounded once to float (what the device computes),
    and what the expf difference does to a full step on the room map (five
    seeds): weights that change and resample indices that change."""
    libm = ctypes.CDLL("libm.so.6")
    for name in ("expf", "sinf", "cosf"):
        getattr(libm, name).restype = ctypes.c_float
        getattr(libm, name).argtypes = [ctypes.c_float]
    for name in ("sin", "cos"):
        getattr(libm, name).restype = ctypes.c_double
        getattr(libm, name).argtypes = [ctypes.c_double]
    rng = np.random.default_rng(3)
    yaw = rng.uniform(-np.pi, np.pi, 50000).astype(F32)
    # glibc's double sin/cos rounded once (numpy's own float64 sin may differ)
    sin_d = np.array([libm.sin(float(a)) for a in yaw], np.float64).astype(F32)
    cos_d = np.array([libm.cos(float(a)) for a in yaw], np.float64).astype(F32)
    sin_f = np.array([libm.sinf(float(a)) for a in yaw], F32)
    cos_f = np.array([libm.cosf(float(a)) for a in yaw], F32)
    rate_sin = float((sin_d.view(np.int32) != sin_f.view(np.int32)).mean())
    rate_cos = float((cos_d.view(np.int32) != cos_f.view(np.int32)).mean())
    cells = weights = indices = total_w = 0
    for seed in range(1, 6):
        c = W.room_case(seed)
        bm = c["map"]
        _, p0 = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=0)
        _, p1 = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
        cells += int((p0.view(np.int32) != p1.view(np.int32)).sum())
        s, parts = c["step"], c["particles"]
        out = []
        for prob in (p0, p1):
            g = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
            op = oracle.pf_predict(parts, s.mods, s.deltas, s.z4)
            op, _ = oracle.pf_update(op, g, s.cloud, s.ranges, s.alpha, s.sigma)
            out.append((op[:, 4].copy(), oracle.resample_indices(op[:, 4], s.u01, 0)))
        weights += int((out[0][0].view(np.int32) != out[1][0].view(np.int32)).sum())
        indices += int((out[0][1] != out[1][1]).sum())
        total_w += parts.shape[0]
    print("libm: sinf vs (float)sin(double) differ on %.2e of yaws, cosf on %.2e; "
          "expf vs (float)exp(double): %d of %d room cells, %d of %d weights, %d resample indices"
          % (rate_sin, rate_cos, cells, 5 * 2400000, weights, total_w, indices))
    assert cells < 1e-3 * 5 * 2400000


def test_voxel_sum_order_effect_on_room_steps():
    """The device sums a voxel's points in point order; PCL sums them in the
    order its unstable std::sort leaves them (oracle.voxel_grid order_mode 0
    and 1).  Measured here on full steps of the room map (five seeds, a
    100,000-point raw scan voxelised at 0.1 m as Node.cpp:130-141 does): the
    centroid coordinates that differ, and what that does downstream to the
    particle weights (relative difference) and the resample indices.  A
    last-bit change of a point can move it to the next cell, so weights can
    differ by more than the 1e-5 relative that holds for identical inputs:
    the measured worst is about 2e-4, with no resample index changed; the
    bounds below are a loose fence around those figures."""
    coords = total_c = weights = total_w = indices = 0
    worst = 0.0
    for seed in range(1, 6):
        c = W.room_case(seed)
        bm, s, parts = c["map"], c["step"], c["particles"]
        rng = np.random.default_rng(500 + seed)
        raw = W.sensor_cloud(bm, c["pose"], 100000, 0.05, rng)
        clouds = [oracle.voxel_grid(raw, 0.1, mode) for mode in (0, 1)]
        assert clouds[0].shape == clouds[1].shape
        coords += int((clouds[0].view(np.int32) != clouds[1].view(np.int32)).sum())
        total_c += clouds[0].size
        _, prob = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
        g = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
        out = []
        for cloud in clouds:
            op = oracle.pf_predict(parts, s.mods, s.deltas, s.z4)
            op, _ = oracle.pf_update(op, g, cloud, s.ranges, s.alpha, s.sigma)
            out.append((op[:, 4].astype(np.float64), oracle.resample_indices(op[:, 4], s.u01, 0)))
        w0, w1 = out[0][0], out[1][0]
        weights += int((w0 != w1).sum())
        total_w += w0.size
        nz = np.maximum(np.abs(w0), np.abs(w1)) > 0
        if nz.any():
            worst = max(worst, float((np.abs(w0 - w1)[nz] / np.maximum(np.abs(w0), np.abs(w1))[nz]).max()))
        indices += int((out[0][1] != out[1][1]).sum())
    print("voxel sum order: %d of %d centroid coordinates differ; %d of %d weights differ, worst relative %.2e; "
          "%d resample indices differ" % (coords, total_c, weights, total_w, worst, indices))
    assert coords > 0  # the two orders do differ
    assert worst < 1e-3
    assert indices <= 0.01 * total_w
