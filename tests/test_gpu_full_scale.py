"""GPU, full scale (opt-in: A3D_FULL_SCALE=1; minutes and 51 GB of HBM):
computeGrid of the 4000 x 4000 x 800 city map (1.28e10 cells, 64-bit cell
indices) on sampled cells against the oracle's float nearest distance (prob
bit-exact) and an FP64 nearest distance (error stated)."""
import os
import sys
import time

import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("A3D_FULL_SCALE") != "1",
                                 reason="full-scale map (51 GB, minutes): set A3D_FULL_SCALE=1")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_full_scale_city_grid(gpu):
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(5)
    bm = W.city_map(5)
    pts = bm.points(dedup=False)
    _say("city map", bm.size, pts.shape)
    g = gpu.Grid3d()
    t0 = time.time()
    g.open(pts, bm.min, bm.max, bm.res, 0.05)
    gpu.synchronize()
    _say("device computeGrid: %d cells in %.1fs" % (g.cells, time.time() - t0))
    assert g.cells == 4000 * 4000 * 800
    sz = np.asarray(g.size, np.int64)
    nsample = 4000
    lin_r = rng.integers(0, int(np.prod(sz)), nsample // 2)
    near = pts[rng.integers(0, pts.shape[0], nsample // 2)].astype(np.float64)
    near += rng.uniform(-0.6, 0.6, near.shape)
    ijk = np.clip(np.floor((near - bm.min) / bm.res).astype(np.int64), 0, sz - 1)
    lin = np.concatenate([lin_r, ijk[:, 0] + sz[0] * (ijk[:, 1] + sz[1] * ijk[:, 2])])
    ix, iy, iz = lin % sz[0], (lin // sz[0]) % sz[1], lin // (sz[0] * sz[1])
    q32 = np.stack([bm.min[0] + ix * bm.res, bm.min[1] + iy * bm.res, bm.min[2] + iz * bm.res], -1).astype(np.float32)
    got = np.array([g.read_cells(int(c), 1)[0] for c in lin], np.float32)
    assert (got > 0).mean() > 0.2  # the near-surface half has prob > 0
    d2 = oracle.nearest_d2(pts, bm.min, bm.max, bm.res, q32)
    c1 = np.float32(1.0 / (0.05 * np.sqrt(2 * np.pi)))
    c2 = np.float32(1.0 / (2.0 * 0.05 * 0.05))
    e = (-d2 * d2 * c2).astype(np.float32)
    ref = (c1 * np.exp(e.astype(np.float64)).astype(np.float32)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32))
    _say("prob vs oracle: exact %.5f, max ulp %d" % ((ulps == 0).mean(), ulps.max()))
    assert ulps.max() <= 1 and (ulps == 0).mean() > 0.999
    d64, _ = cKDTree(pts.astype(np.float64)).query(q32.astype(np.float64))
    p64 = float(c1) * np.exp(-(d64 ** 2) ** 2 * float(c2))
    _say("prob vs FP64 nearest distance: max abs err %.3g (peak %.3g)" % (np.abs(got - p64).max(), c1))
    assert np.abs(got - p64).max() < 1e-5 * float(c1)
