"""The library's node glue (a3d_node_*) against the oracle node
(oracle/node.py) on the same scripted flight: odometry with roll and pitch,
beacon ranges, raw clouds voxelised by the node, update gating, take-off, a
resample interval and an odometry jump.  Everything the node publishes, its
deltas, means and particles are compared bit for bit."""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle
from amcl3d_amd import _lib
from amcl3d_amd import workloads as W
from oracle import node as on

pytestmark = pytest.mark.gpu

PARAMS = dict(num_particles=600, resample_interval=1, set_initial_pose=1, init_x=6.0, init_y=7.0, init_z=0.3,
              init_a=0.4, init_x_dev=0.3, init_y_dev=0.3, init_z_dev=0.1, init_a_dev=0.15, sensor_range=0.1,
              voxel_size=0.1, odom_x_mod=0.1, odom_y_mod=0.1, odom_z_mod=0.1, odom_a_mod=0.3, update_rate=10.0,
              d_th=0.1, a_th=0.05, take_off_height=0.5, alpha=0.8)
SEED = 77
SCAN_SIZES = [1500, 1500, 3000, 700, 700, 5200, 1200, 4000, 300]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _draws(n):
    cache = {}

    def d(step):
        if step not in cache:
            z4 = np.empty(4 * n, np.float32)
            u = ctypes.c_float()
            _lib.call("a3d_session_draws", SEED, int(step), n, _lib.ptr(z4), ctypes.byref(u))
            cache[step] = (z4, float(u.value))
        return cache[step]
    return d


def _world(odom, init=(6.0, 7.0, 0.3, 0.4)):
    c, s = math.cos(init[3]), math.sin(init[3])
    return (init[0] + c * odom[0] - s * odom[1], init[1] + s * odom[0] + c * odom[1], init[2] + odom[2],
            init[3] + odom[3])


def test_node_matches_the_oracle_on_a_scripted_flight(gpu):
    bm = W.room_map(1)
    pts = bm.points()
    g = gpu.Grid3d()
    g.open(pts, bm.min, bm.max, bm.res, 0.1)
    _, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, 0.1, exp_mode=1)
    og = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    from amcl3d_amd.node import Node

    nd = Node(PARAMS, g, seed=SEED)
    ref = on.Node(PARAMS, og, _draws(PARAMS["num_particles"]))
    rng = np.random.default_rng(3)
    anchors = [(1.0, 1.0, 2.0), (18.0, 2.0, 1.0), (9.0, 18.0, 2.5)]
    # odometry: climb to 1.5 m, fly a curve, then a 3 m jump of the odometry
    path = []
    for k in range(40):
        z = min(1.5, 0.05 * k)
        path.append((0.08 * k, 0.03 * k + 0.2 * math.sin(0.2 * k), z, 0.02 * k, 0.01 * math.sin(k), 0.015))
    for k in range(40, 46):
        o = path[-1]
        path.append((o[0] + (3.0 if k == 42 else 0.1), o[1], o[2], o[3], o[4], o[5]))
    now = 0
    updates = 0
    for k, (x, y, z, yaw, roll, pitch) in enumerate(path):
        q = on.quat_rpy(roll, pitch, yaw)
        _say("step", k, "odometry", (x, y, z, yaw))
        a = nd.odomCallback([x, y, z], q)
        b = ref.odometry([x, y, z], q)
        assert (a is None) == (b is None)
        if a is not None:
            for key in ("base_t", "base_q", "odom_t", "odom_q"):
                assert a[key] == list(b[key]), (k, key, a[key], b[key])
            assert a["amcl_out"] == b["amcl_out"] and a["has_takenoff"] == b["has_takenoff"]
        w = _world((x, y, z, yaw))
        for an in anchors:
            r = math.dist(w[:3], an) + rng.normal(0, 0.05)
            nd.rangeCallback(r, *an)
            ref.range(r, *an)
        # scan sizes grow, shrink and grow again (the node's device buffers
        # and VoxelGrid workspace regrow, then run with capacity above the
        # scan size); every other scan arrives as pcl::PointXYZ (16 bytes)
        m = SCAN_SIZES[k % len(SCAN_SIZES)]
        cloud = W.sensor_cloud(bm, (w[0], w[1], max(w[2], 0.3), w[3]), m, 0.02, rng)
        if k % 2:
            cloud = np.concatenate([cloud, np.ones((m, 1), np.float32)], 1)
        now += 60_000_000 if k % 3 else 30_000_000
        ua = nd.pointcloudCallback(now, cloud)
        ub = ref.point_cloud(now, cloud[:, :3])
        assert ua == ub, k
        updates += ua
        st = nd.state()
        assert np.array_equal(st["mean"], ref.mean), (k, st["mean"], ref.mean)
        assert list(st["deltas"]) == list(ref.deltas)
        assert st["roll"] == ref.roll and st["pitch"] == ref.pitch
        assert st["amcl_out"] == ref.amcl_out and st["draw_step"] == ref.step
    _say("updates", updates, "amcl_out", ref.amcl_out, "mean", ref.mean)
    assert updates >= 10
    assert ref.amcl_out  # the jump was seen
    assert np.array_equal(nd.particles(), ref.particles)


def test_node_callbacks_keep_timing_bounded(gpu):
    """Hundreds of point-cloud callbacks that never read the filter's kernel
    timing: the timing events are recycled (the pool holds at most 64 pairs),
    and every update is still timed and counted."""
    bm = W.room_map(2)
    g = gpu.Grid3d()
    g.open(bm.points(), bm.min, bm.max, bm.res, 0.1)
    from amcl3d_amd.node import Node

    params = dict(PARAMS, num_particles=200, resample_interval=0)
    nd = Node(params, g, seed=5)
    rng = np.random.default_rng(8)
    q = on.quat_rpy(0.0, 0.0, 0.4)
    nd.odomCallback([0.0, 0.0, 1.0], q)
    cloud = W.sensor_cloud(bm, (6.0, 7.0, 1.3, 0.4), 800, 0.02, rng)  # the same scan each time
    now, updates = 0, 0
    _say("300 node callbacks")
    for k in range(300):
        nd.odomCallback([0.15 * (k + 1), 0.0, 1.0], q)
        now += 120_000_000
        updates += nd.pointcloudCallback(now, cloud)
    assert updates >= 250
    pf = ctypes.c_void_p()
    _lib.call("a3d_node_get_filter", nd._h, ctypes.byref(pf))
    ms, launches, evals = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
    _lib.call("a3d_pf_timing", pf.value, 1, ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(evals))
    _say("updates", updates, "timed launches", launches.value, "avg ms", ms.value)
    assert launches.value == updates and ms.value > 0
