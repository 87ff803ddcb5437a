"""The node glue's oracle (oracle/node.py) on the CPU: the tf / LinearMath
routines against closed forms and round trips, ros::Duration's rounding,
and the node's gating, take-off and jump rules on scripted inputs.  The
library's node is checked against this oracle on the GPU
(tests/test_gpu_node.py)."""
import math

import numpy as np
import pytest

import oracle
from oracle import node as on


def _close(a, b, tol=1e-12):
    return all(abs(x - y) <= tol for x, y in zip(np.ravel(a), np.ravel(b)))


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("ang", [0.0, 0.3, -1.2, 2.9, math.pi / 2 - 1e-3])
def test_single_axis_rotation_matches_the_closed_form(axis, ang):
    q = [0.0, 0.0, 0.0, math.cos(ang / 2)]
    q[axis] = math.sin(ang / 2)
    b = on.quat_basis(q)
    c, s = math.cos(ang), math.sin(ang)
    R = np.eye(3)
    j, k = (axis + 1) % 3, (axis + 2) % 3
    R[j][j], R[j][k], R[k][j], R[k][k] = c, -s, s, c
    assert _close(b, R, 4e-16 * 4)
    r, p, y = on.Tf(b).rpy()
    if axis == 1 and abs(ang) > math.pi / 2:
        # pitch beyond +-pi/2 comes back as the equivalent (pi, pi - ang, pi)
        assert _close(on.quat_basis(on.quat_rpy(r, p, y)), b, 1e-15)
        assert abs(abs(r) - math.pi) < 1e-12 and abs(abs(y) - math.pi) < 1e-12
        return
    want = [0.0, 0.0, 0.0]
    want[axis] = ang
    assert _close((r, p, y), want, 1e-12)


def test_euler_quaternion_basis_round_trips():
    rng = np.random.default_rng(0)
    for _ in range(200):
        r, p, y = rng.uniform(-math.pi, math.pi), rng.uniform(-1.5, 1.5), rng.uniform(-math.pi, math.pi)
        q = on.quat_rpy(r, p, y)
        assert abs(sum(v * v for v in q) - 1.0) < 1e-15
        t = on.Tf(on.quat_basis(q))
        assert _close(t.rpy(), (r, p, y), 1e-12)
        q2 = t.quat()
        sgn = 1.0 if q2[3] * q[3] + q2[0] * q[0] + q2[1] * q[1] + q2[2] * q[2] > 0 else -1.0
        assert _close([sgn * v for v in q2], q, 1e-14)


def test_transform_inverse_and_product():
    rng = np.random.default_rng(1)
    for _ in range(100):
        q = on.quat_rpy(*rng.uniform(-1, 1, 3))
        a = on.Tf.from_tq(rng.uniform(-10, 10, 3), q)
        i = a * a.inverse()
        assert _close(i.b, np.eye(3), 1e-15) and _close(i.o, [0, 0, 0], 1e-13)
        b = on.Tf.from_tq(rng.uniform(-10, 10, 3), on.quat_rpy(*rng.uniform(-1, 1, 3)))
        M = lambda t: np.block([[np.array(t.b), np.array(t.o)[:, None]], [np.zeros((1, 3)), np.ones((1, 1))]])  # noqa
        assert np.allclose(M(a * b), M(a) @ M(b), atol=1e-12)


def test_gimbal_lock_branches():
    for sgn in (1, -1):
        b = on.quat_basis(on.quat_rpy(0.4, sgn * math.pi / 2, 0.0))
        b[2][0] = -float(sgn)  # exactly at the singularity
        r, p, y = on.Tf(b).rpy()
        assert y == 0.0 and p == sgn * on.PI / 2.0


def test_duration_rounding():
    assert on.duration_ns(0.1) == 100000000
    assert on.duration_ns(1.0 / 3.0) == 333333333
    assert on.duration_ns(2.0 / 3.0) == 666666667
    assert on.duration_ns(1.5) == 1500000000
    assert on.duration_ns(1e-10) == 0


def _room():
    from amcl3d_amd import workloads as W

    bm = W.room_map(1)
    pts = bm.points()
    _, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, 0.1, exp_mode=1)
    return bm, oracle.GridMeta(bm.min, bm.max, bm.res, prob)


PARAMS = dict(num_particles=200, resample_interval=1, set_initial_pose=1, init_x=5.0, init_y=5.0, init_z=0.2,
              init_a=0.3, init_x_dev=0.3, init_y_dev=0.3, init_z_dev=0.1, init_a_dev=0.2, sensor_range=0.1,
              voxel_size=0.1, odom_x_mod=0.1, odom_y_mod=0.1, odom_z_mod=0.1, odom_a_mod=0.3, update_rate=10.0,
              d_th=0.1, a_th=0.05, take_off_height=0.5, alpha=0.8)


def _draws(n):
    rng = np.random.default_rng(42)
    cache = {}

    def d(step):
        if step not in cache:
            cache[step] = (rng.standard_normal(4 * n).astype(np.float32), float(rng.uniform()))
        return cache[step]
    return d


def test_node_rules_on_a_scripted_flight():
    """Hand-fixed outcomes: nothing is published before the filter exists;
    the pose follows odometry until take-off; a cloud updates only when the
    time and a distance or yaw threshold allow it; resample every second
    update; a jump of more than 1 m in the mean sets amcl_out for good."""
    bm, g = _room()
    nd = on.Node(PARAMS, g, _draws(PARAMS["num_particles"]))
    q0 = [0.0, 0.0, 0.0, 1.0]
    anchors = [(1.0, 1.0, 2.0), (9.0, 2.0, 1.0), (5.0, 9.0, 2.5)]

    def ranges(odom):  # beacon ranges from the world pose init * odom
        c, s_ = math.cos(0.3), math.sin(0.3)
        w = (5.0 + c * odom[0] - s_ * odom[1], 5.0 + s_ * odom[0] + c * odom[1], 0.2 + odom[2])
        for a in anchors:
            nd.range(math.dist(w, a), *a)
    assert nd.odometry([0.0, 0.0, 0.0], q0) is None  # initialises, publishes nothing
    assert nd.particles is not None and nd.step == 1
    out = nd.odometry([0.0, 0.0, 0.1], q0)
    assert not out["has_takenoff"] and _close(out["base_t"], [5.0, 5.0, 0.3], 1e-12)
    cloud = np.array([[1.0, 0.0, 0.0]], np.float32)  # <= 10 points: every cloud weight is 0

    def cloud_at(now, odom):
        ranges(odom)
        return nd.point_cloud(now, cloud)

    # moved 0.1 m, not more than d_th: no update
    assert not cloud_at(0, [0.0, 0.0, 0.1])
    nd.ranges = []
    nd.odometry([0.2, 0.0, 0.1], q0)
    assert cloud_at(0, [0.2, 0.0, 0.1]) and nd.step == 2 and nd.n_updates == 1
    # the increment since the pose the filter was initialised at (Node.cpp:394)
    assert _close(nd.deltas, [0.2, 0.0, 0.1, 0.0], 1e-15)
    # moved again, but within 1/update_rate of the last update
    nd.odometry([0.4, 0.0, 0.1], q0)
    assert not cloud_at(99_999_999, [0.4, 0.0, 0.1])
    nd.ranges = []
    assert cloud_at(100_000_000, [0.4, 0.0, 0.1]) and nd.n_updates == 0  # second update resampled
    # a yaw-only move above a_th updates
    a = 0.06
    nd.odometry([0.4, 0.0, 0.1], [0.0, 0.0, math.sin(a / 2), math.cos(a / 2)])
    assert cloud_at(200_000_000, [0.4, 0.0, 0.1])
    assert abs(nd.deltas[3] - a) < 1e-12
    # take off; the mean then follows the beacons without jumping
    out = nd.odometry([0.4, 0.0, 1.0], q0)
    assert out["has_takenoff"]
    assert cloud_at(300_000_000, [0.4, 0.0, 1.0])
    out = nd.odometry([0.4, 0.0, 1.0], q0)
    assert not out["amcl_out"]
    # a 3 m odometry jump the beacons do not see moves the mean by more than 1 m
    nd.odometry([3.4, 0.0, 1.0], q0)
    assert cloud_at(400_000_000, [0.4, 0.0, 1.0])
    assert abs(float(nd.mean[0] - nd.lastmean[0])) > 1.0
    out = nd.odometry([3.4, 0.0, 1.0], q0)
    assert out["amcl_out"]
    out = nd.odometry([3.5, 0.0, 1.0], q0)
    assert out["amcl_out"]  # stays set
