"""GPU: the sums of a single large filter (16,384 particles or more), which
run tiled across the chip, against the oracle's sequential float sums
(ParticleFilter.cpp:150-222): full predict -> update -> resample steps bit
for bit, with one-signed and mixed-sign mean terms (yaws on both sides of 0,
a map at negative coordinates) and resample's running sum over weights
built to cross many binades."""
import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def room(gpu):
    c = W.room_case(6)
    bm = c["map"]
    _, prob = oracle.compute_grid(bm.points(), bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
    return dict(case=c, bm=bm, prob=prob)


@pytest.mark.parametrize("n,yaw,shift", [(16384, 0.9, 0.0), (40000, 0.0, 0.0), (100000, -2.1, 0.0),
                                         (65000, 0.4, -30.0)])
def test_tiled_full_step(gpu, room, n, yaw, shift):
    """yaw 0.0: the particles' yaws straddle 0 (the yaw mean's terms have both
    signs); shift -30: the map and everything on it moved 30 m along -x and
    -y (negative x, y mean terms)."""
    c, bm = room["case"], room["bm"]
    rng = np.random.default_rng(n)
    s = c["step"]
    sh = np.array([shift, shift, 0.0])
    mn, mx = bm.min + sh, bm.max + sh
    grid = gpu.Grid3d().from_cells(room["prob"], mn, mx, bm.res, c["sensor_dev"])
    og = oracle.GridMeta(mn, mx, bm.res, room["prob"])
    pose = np.array([10.0 + shift, 10.0 + shift, 1.5, yaw])
    parts = W.particles_around(pose, n, rng)
    parts[: n // 50, 2] = -50.0  # some outside the map
    z4 = rng.standard_normal((n, 4)).astype(F32)
    ranges = s.ranges.copy()
    ranges[:, 1:3] += shift
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    pf.predict(*s.mods, *s.deltas, z4)
    pf.update(grid, s.cloud, ranges, s.alpha, s.sigma)
    upd, mean = pf.particles(), pf.getMean()
    idx = pf.resample(s.u01, gpu.RESAMPLE_REFERENCE, return_indices=True)
    op = oracle.pf_predict(parts, s.mods, s.deltas, z4)
    op, om = oracle.pf_update(op, og, s.cloud, ranges, s.alpha, s.sigma)
    oq, oidx = oracle.pf_resample(op, s.u01, 0)
    print("n %d: mean %s" % (n, om), flush=True)
    assert np.array_equal(upd.view(np.int32), op.view(np.int32))
    assert np.array_equal(mean.view(np.int32), om.view(np.int32))
    assert np.array_equal(idx, oidx)
    assert np.array_equal(pf.particles().view(np.int32), oq.view(np.int32))


@pytest.mark.parametrize("kind", ["exponential", "growing", "ties", "zeros_first", "subnormal", "one_hot", "nan_after_zeros",
                                  "inf_mid", "negative_run"])
def test_tiled_resample_running_sum(gpu, kind):
    """resample alone (its running sum over the set weights, tiled) against
    the reference's loop, for weights that cross many binades, sit on exact
    ties, start with long zero runs, stay subnormal, hold a NaN or an inf, or
    change sign for a run."""
    n = 70000
    rng = np.random.default_rng(3)
    if kind == "exponential":
        w = rng.exponential(1.0, n)
    elif kind == "growing":
        w = np.exp(np.linspace(-40, 10, n))
    elif kind == "ties":
        w = rng.integers(0, 8, n) * 2.0 ** -20
    elif kind == "zeros_first":
        w = np.concatenate([np.zeros(30000), rng.uniform(0, 1, n - 30000)])
    elif kind == "subnormal":
        w = rng.uniform(0, 1, n) * 1e-41
    elif kind == "one_hot":
        w = np.zeros(n)
        w[54321] = 1.0
    elif kind == "nan_after_zeros":
        w = np.concatenate([np.zeros(30000), rng.uniform(0, 1, n - 30000)])
        w[30000] = np.nan
    elif kind == "inf_mid":
        w = rng.uniform(0, 1, n)
        w[40000] = np.inf
    else:  # a run of negative weights between positive ones
        w = rng.uniform(0, 1, n)
        w[20000:26000] *= -1.0
    p = np.zeros((n, 7), F32)
    p[:, 0] = np.arange(n)
    p[:, 4] = w.astype(F32)
    pf = gpu.ParticleFilter()
    pf.set_particles(p)
    idx = pf.resample(0.41, gpu.RESAMPLE_REFERENCE, return_indices=True)
    assert np.array_equal(idx, oracle.resample_indices(p[:, 4], 0.41, 0)), kind


@pytest.mark.parametrize("n", [1_048_576 + 1, 2_500_000])
def test_tiled_resample_past_one_chunk_of_tile_records(gpu, n):
    """More than 1,024 tiles (the walk holds the tile records of one chunk
    in LDS at a time): resample indices over several chunks, weights growing
    across binades, against the reference's loop."""
    rng = np.random.default_rng(n)
    w = (np.exp(np.linspace(-20, 5, n)) * rng.uniform(0.5, 1.5, n)).astype(F32)
    p = np.zeros((n, 7), F32)
    p[:, 0] = np.arange(n) % 1000
    p[:, 4] = w
    pf = gpu.ParticleFilter()
    pf.set_particles(p)
    idx = pf.resample(0.37, gpu.RESAMPLE_REFERENCE, return_indices=True)
    assert np.array_equal(idx, oracle.resample_indices(w, 0.37, 0))


def test_tiled_full_step_past_one_chunk(gpu, room):
    """A full step of 1.2M particles (1,172 tiles, two chunks of tile records)
    with a 200-point cloud: every sum of the update tail and resample bit for
    bit against the oracle."""
    c, bm = room["case"], room["bm"]
    n = 1_200_000
    rng = np.random.default_rng(12)
    s = c["step"]
    grid = gpu.Grid3d().from_cells(room["prob"], bm.min, bm.max, bm.res, c["sensor_dev"])
    og = oracle.GridMeta(bm.min, bm.max, bm.res, room["prob"])
    parts = W.particles_around(np.array([10.0, 10.0, 1.5, 0.2]), n, rng)
    z4 = rng.standard_normal((n, 4)).astype(F32)
    cloud = s.cloud[:200]
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    pf.predict(*s.mods, *s.deltas, z4)
    pf.update(grid, cloud, s.ranges, s.alpha, s.sigma)
    upd, mean = pf.particles(), pf.getMean()
    idx = pf.resample(s.u01, gpu.RESAMPLE_REFERENCE, return_indices=True)
    op = oracle.pf_predict(parts, s.mods, s.deltas, z4)
    op, om = oracle.pf_update(op, og, cloud, s.ranges, s.alpha, s.sigma)
    _, oidx = oracle.pf_resample(op, s.u01, 0)
    assert np.array_equal(upd.view(np.int32), op.view(np.int32))
    assert np.array_equal(mean.view(np.int32), om.view(np.int32))
    assert np.array_equal(idx, oidx)
