"""GPU, full scale: the 200 x 200 x 40 m city block at 0.05 m (4000 x 4000 x
800 = 1.28e10 cells, 51.2 GB of prob, 64-bit cell indices) from ~49M seeded
occupied points, built on the device by computeGrid (PointCloudTools.cpp:
84-149).

* Sampled cells, half near the surfaces: prob bit-exact against the oracle's
  float nearest distance and c1 * exp(-d2 * d2 * c2); the FP64 nearest
  distance gives the stated error.  The oracle and the FP64 tree search only
  the points within 1 m of the sampled cells (every point of a 27-bucket
  neighbourhood of 1 m buckets): a cell whose nearest point is 1 m or more
  away has d2 >= 1 and prob c1 * exp(-200) = 0 in float, whichever point is
  nearest, so the subset gives every cell's exact prob.
* Cloud weights (Grid3d.cpp:133-199) of sampled particles of a global
  localisation session (uniform over the free space) and of two tracking
  sessions, 10,000-point clouds, on the device's grid, against the oracle's
  weight composed from its own cells (the cells computeCloudWeight gathers)
  and its own computeGrid of those cells (nearest neighbour over all the
  map's points): bit-exact.

About a minute of host work (the oracle's tree over 49M points) beside a
0.7 s device build; the grid stays on the device (no read-back).
"""
import sys
import time

import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32
SENSOR_DEV = 0.05
C1 = F32(1.0 / (SENSOR_DEV * np.sqrt(2 * np.pi)))
C2 = F32(1.0 / (2.0 * SENSOR_DEV * SENSOR_DEV))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _prob(d2):
    """c1 * exp(-d2 * d2 * c2) with the exponent's float product, exp in double."""
    e = (-d2 * d2 * C2).astype(F32)
    return (C1 * np.exp(e.astype(np.float64)).astype(F32)).astype(F32)


def _cell_centres(bm, lin):
    sz = bm.size.astype(np.int64)
    ix, iy, iz = lin % sz[0], (lin // sz[0]) % sz[1], lin // (sz[0] * sz[1])
    return np.stack([bm.min[0] + ix * bm.res, bm.min[1] + iy * bm.res, bm.min[2] + iz * bm.res], -1).astype(F32)


def _near_subset(pts, keys, q):
    """The points in the 1 m buckets around each query (every point within 1 m)."""
    qk = np.floor(q.astype(np.float64)).astype(np.int64)
    offs = np.stack(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"), -1).reshape(-1, 3)
    nk = (qk[:, None, :] + offs[None]).reshape(-1, 3)
    need = np.unique((nk[:, 0] * 1024 + nk[:, 1]) * 1024 + nk[:, 2])
    return pts[np.isin(keys, need)]


@pytest.fixture(scope="module")
def city(gpu):
    t0 = time.time()
    bm = W.city_map(5)
    pts = bm.points(dedup=False)
    _say("city map", bm.size, pts.shape, "(%.1fs)" % (time.time() - t0))
    g = gpu.Grid3d()
    t1 = time.time()
    g.open(pts, bm.min, bm.max, bm.res, SENSOR_DEV)
    gpu.synchronize()
    _say("device computeGrid: %d cells in %.1fs" % (g.cells, time.time() - t1))
    assert g.cells == 4000 * 4000 * 800
    k = np.floor(pts.astype(np.float64)).astype(np.int64)
    keys = (k[:, 0] * 1024 + k[:, 1]) * 1024 + k[:, 2]
    return dict(bm=bm, pts=pts, grid=g, keys=keys)


def test_full_scale_city_grid(gpu, city):
    from scipy.spatial import cKDTree

    bm, pts, g = city["bm"], city["pts"], city["grid"]
    rng = np.random.default_rng(5)
    sz = bm.size.astype(np.int64)
    nsample = 4000
    lin_r = rng.integers(0, int(np.prod(sz)), nsample // 2)
    near = pts[rng.integers(0, pts.shape[0], nsample // 2)].astype(np.float64)
    near += rng.uniform(-0.6, 0.6, near.shape)
    ijk = np.clip(np.floor((near - bm.min) / bm.res).astype(np.int64), 0, sz - 1)
    lin = np.concatenate([lin_r, ijk[:, 0] + sz[0] * (ijk[:, 1] + sz[1] * ijk[:, 2])])
    q32 = _cell_centres(bm, lin)
    got = np.array([g.read_cells(int(c), 1)[0] for c in lin], np.float32)
    assert (got > 0).mean() > 0.2  # the near-surface half has prob > 0
    assert (lin >= 2 ** 32).sum() > nsample // 4  # cells past 32-bit indices
    sub = _near_subset(pts, city["keys"], q32)
    t0 = time.time()
    d2 = oracle.nearest_d2(sub, bm.min, bm.max, bm.res, q32)
    d2 = np.where(d2 < 0, np.float32(np.inf), d2)  # no point within 1 m: prob 0
    _say("oracle nearest distances of %d cells over %d nearby points (%.1fs)" % (lin.size, sub.shape[0],
                                                                                time.time() - t0))
    ref = _prob(d2)
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))
    d64, _ = cKDTree(sub.astype(np.float64)).query(q32.astype(np.float64), distance_upper_bound=1.0)
    p64 = float(C1) * np.exp(-(d64 ** 2) ** 2 * float(C2))
    err = np.abs(got.astype(np.float64) - p64)
    _say("prob vs FP64 nearest distance: max abs err %.3g (peak %.3g)" % (err.max(), C1))
    assert err.max() < 1e-5 * float(C1)


def _weights_from_cells(cells, prob_of):
    """computeCloudWeight from the gathered cells: float sum in cloud order of
    prob over the in-map points, divided by their count when it exceeds 10."""
    out = np.empty(cells.shape[0], F32)
    for i in range(cells.shape[0]):
        c = cells[i]
        c = c[c >= 0]
        vals = np.concatenate([np.zeros(1, F32), prob_of(c)])
        s = np.add.accumulate(vals, dtype=F32)[-1]
        out[i] = F32(0) if c.size <= 10 else F32(s / F32(c.size))
    return out


def test_full_scale_session_weights(gpu, city):
    """A global session's and two tracking sessions' sampled particles:
    device weights on the 1.28e10-cell grid equal the oracle's composition."""
    from amcl3d_amd.full_scale import free_particles

    bm, pts, g = city["bm"], city["pts"], city["grid"]
    rng = np.random.default_rng(17)
    og = oracle.GridMeta(bm.min, bm.max, bm.res)  # geometry only
    cases = []
    true = np.append(W.free_pose(bm, rng, z_range=(1.5, 1.5), clearance=1.0), rng.uniform(-np.pi, np.pi))
    cases.append(("global", free_particles(bm, 48, rng), W.sensor_cloud(bm, true, 10000, 0.05, rng)))
    for z in (1.5, 12.0):
        pose = np.append(W.free_pose(bm, rng, z_range=(z, z), clearance=1.0), rng.uniform(-np.pi, np.pi))
        cases.append(("tracking z=%.1f" % z, W.particles_around(pose, 64, rng), W.sensor_cloud(bm, pose, 10000, 0.05,
                                                                                                rng)))
    gathered = []
    for name, p, cloud in cases:
        cells = oracle.cloud_cells(og, cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
        gathered.append(cells)
        _say("%s: %d particles, %d in-map points, highest cell %d" % (name, p.shape[0], (cells >= 0).sum(),
                                                                      cells.max()))
    assert max(int(c.max()) for c in gathered) >= 2 ** 32
    uniq = np.unique(np.concatenate([c[c >= 0] for c in gathered]))
    t0 = time.time()
    d2 = oracle.nearest_d2(pts, bm.min, bm.max, bm.res, _cell_centres(bm, uniq))
    _say("oracle computeGrid of %d gathered cells over all %d points (%.1fs)" % (uniq.size, pts.shape[0],
                                                                                time.time() - t0))
    probs = _prob(d2)

    def prob_of(c):
        return probs[np.searchsorted(uniq, c)]

    for (name, p, cloud), cells in zip(cases, gathered):
        w = g.cloud_weights(cloud, p[:, 0], p[:, 1], p[:, 2], p[:, 3])
        ow = _weights_from_cells(cells, prob_of)
        _say("%s: weights > 0: %d of %d" % (name, (ow > 0).sum(), ow.size))
        assert np.array_equal(w.view(np.int32), ow.view(np.int32)), name
    assert (ow > 0).mean() > 0.5  # the last tracking session sees its surfaces
