"""Oracle for the node's glue (test infrastructure only): amcl3d/src/Node.cpp's
odometry, point-cloud and range callbacks (Node.cpp:112-191,193-324,326-398)
restated in Python over the oracle's filter (``oracle.pf_*``) and VoxelGrid,
with the tf / Bullet LinearMath double arithmetic the node uses (tf is not
vendored and is absent here; its published Matrix3x3 / Quaternion /
Transform routines are followed).

Pins (tests/test_node.py): closed forms of the transform routines (rotations
about one axis, round trips Euler -> quaternion -> basis -> Euler and
quaternion -> basis -> quaternion, T * T^-1 = I, the gimbal-lock branches),
ros::Duration's rounding, and the gating / take-off / jump rules on scripted
inputs whose outcome is fixed by hand.
"""
import math

import numpy as np

import oracle

PI = 3.1415926535897932384626433832795029


def quat_basis(q):
    """Matrix3x3::setRotation(q), q = (x, y, z, w)."""
    x, y, z, w = q
    d = x * x + y * y + z * z + w * w
    s = 2.0 / d
    xs, ys, zs = x * s, y * s, z * s
    wx, wy, wz = w * xs, w * ys, w * zs
    xx, xy, xz = x * xs, x * ys, x * zs
    yy, yz, zz = y * ys, y * zs, z * zs
    return [[1.0 - (yy + zz), xy - wz, xz + wy],
            [xy + wz, 1.0 - (xx + zz), yz - wx],
            [xz - wy, yz + wx, 1.0 - (xx + yy)]]


class Tf:
    """tf::Transform: basis (rows) and origin."""

    def __init__(self, basis=None, origin=(0.0, 0.0, 0.0)):
        self.b = basis if basis is not None else [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
        self.o = list(origin)

    @staticmethod
    def from_tq(t, q):
        return Tf(quat_basis(q), [float(v) for v in t])

    def inverse(self):
        inv = [[self.b[j][i] for j in range(3)] for i in range(3)]
        n = [-v for v in self.o]
        return Tf(inv, [inv[i][0] * n[0] + inv[i][1] * n[1] + inv[i][2] * n[2] for i in range(3)])

    def __mul__(self, t):
        a, c = self.b, t.b
        # Matrix3x3 operator*: m2.tdotx(m1[i]) = m2[0][j]*m1[i][0] + m2[1][j]*m1[i][1] + m2[2][j]*m1[i][2]
        m = [[c[0][j] * a[i][0] + c[1][j] * a[i][1] + c[2][j] * a[i][2] for j in range(3)] for i in range(3)]
        o = [a[i][0] * t.o[0] + a[i][1] * t.o[1] + a[i][2] * t.o[2] + self.o[i] for i in range(3)]
        return Tf(m, o)

    def rpy(self):
        """Matrix3x3::getRPY -> getEulerYPR, solution 1: (roll, pitch, yaw)."""
        m = self.b
        if abs(m[2][0]) >= 1:
            delta = math.atan2(m[2][1], m[2][2])
            pitch = PI / 2.0 if m[2][0] < 0 else -PI / 2.0
            return delta, pitch, 0.0
        pitch = -math.asin(m[2][0])
        roll = math.atan2(m[2][1] / math.cos(pitch), m[2][2] / math.cos(pitch))
        yaw = math.atan2(m[1][0] / math.cos(pitch), m[0][0] / math.cos(pitch))
        return roll, pitch, yaw

    def quat(self):
        """Matrix3x3::getRotation."""
        m = self.b
        trace = m[0][0] + m[1][1] + m[2][2]
        t = [0.0] * 4
        if trace > 0.0:
            s = math.sqrt(trace + 1.0)
            t[3] = s * 0.5
            s = 0.5 / s
            t[0] = (m[2][1] - m[1][2]) * s
            t[1] = (m[0][2] - m[2][0]) * s
            t[2] = (m[1][0] - m[0][1]) * s
        else:
            i = (2 if m[1][1] < m[2][2] else 1) if m[0][0] < m[1][1] else (2 if m[0][0] < m[2][2] else 0)
            j, k = (i + 1) % 3, (i + 2) % 3
            s = math.sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0)
            t[i] = s * 0.5
            s = 0.5 / s
            t[3] = (m[k][j] - m[j][k]) * s
            t[j] = (m[j][i] + m[i][j]) * s
            t[k] = (m[k][i] + m[i][k]) * s
        return t


def quat_rpy(roll, pitch, yaw):
    """Quaternion::setRPY."""
    hy, hp, hr = yaw * 0.5, pitch * 0.5, roll * 0.5
    cy, sy = math.cos(hy), math.sin(hy)
    cp, sp = math.cos(hp), math.sin(hp)
    cr, sr = math.cos(hr), math.sin(hr)
    return [sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy, cr * cp * sy - sr * sp * cy,
            cr * cp * cy + sr * sp * sy]


def duration_ns(d):
    """ros::Duration(d) (DurationBase::fromSec) in nanoseconds."""
    sec = math.floor(d)
    frac = (d - sec) * 1e9
    nsec = math.floor(frac + 0.5) if frac >= 0 else -math.floor(-frac + 0.5)
    return int(sec) * 1000000000 + int(nsec)


class Node:
    """Node's callbacks over the oracle filter.  draws(step) -> (z4, u01) are
    the filter's draws (inputs); grid is an ``oracle.GridMeta``."""

    def __init__(self, params, grid, draws):
        self.p = dict(params)
        self.g = grid
        self.draws = draws
        self.step = 0
        self.particles = None
        self.is_odom = False
        self.amcl_out = False
        self.has_takenoff = False
        self.n_updates = 0
        self.roll = 0.0
        self.pitch = 0.0
        self.mean = np.zeros(4, np.float32)
        self.lastmean = np.zeros(4, np.float32)
        self.deltas = [0.0] * 4
        I = Tf()  # noqa: E741
        self.lastbase_2_world = self.initodom_2_world = self.lastodom_2_world = I
        self.amcl_out_lastbase_2_odom = self.lastupdatebase_2_odom = self.base_2_odom = I
        self.nextupdate_ns = 0
        self.ranges = []

    def set_initial_pose(self, init_pose, xdev, ydev, zdev, adev):
        self.initodom_2_world = init_pose
        x, y, z = (float(np.float32(v)) for v in init_pose.o)
        a = float(np.float32(init_pose.rpy()[2]))
        z4, _ = self.draws(0)
        self.step = 1
        f = lambda v: float(np.float32(v))  # noqa: E731
        self.particles, self.mean = oracle.pf_init(self.p["num_particles"], (x, y, z, a),
                                                   (f(xdev), f(ydev), f(zdev), f(adev)), z4)
        self.lastmean = self.mean.copy()
        self.lastupdatebase_2_odom = self.base_2_odom

    def odometry(self, t, q):
        self.base_2_odom = Tf.from_tq(t, q)
        if self.particles is None:
            if self.p["set_initial_pose"]:
                a = self.p["init_a"]
                pose = Tf.from_tq((self.p["init_x"], self.p["init_y"], self.p["init_z"]),
                                  (0.0, 0.0, math.sin(a * 0.5), math.cos(a * 0.5)))
                self.set_initial_pose(pose, self.p["init_x_dev"], self.p["init_y_dev"], self.p["init_z_dev"],
                                      self.p["init_a_dev"])
            return None
        self.roll, self.pitch, _ = self.base_2_odom.rpy()
        if not self.is_odom:
            self.is_odom = True
            self.lastbase_2_world = self.initodom_2_world * self.base_2_odom
            self.lastodom_2_world = self.initodom_2_world
        if not self.has_takenoff:
            self.has_takenoff = self.base_2_odom.o[2] > self.p["take_off_height"]
            self.lastbase_2_world = self.initodom_2_world * self.base_2_odom
            self.lastodom_2_world = self.initodom_2_world
            self.lastmean = self.mean.copy()
        else:
            m = self.mean
            if np.any(np.isnan(m)) or np.any(np.isinf(m)):
                self.amcl_out = True
            if np.any(np.abs(m - self.lastmean) > 1.0):  # float differences, as the node's
                self.amcl_out = True
            if not self.amcl_out:
                b2w = Tf.from_tq([float(v) for v in m[:3]], quat_rpy(self.roll, self.pitch, float(m[3])))
                b2w = b2w * self.lastupdatebase_2_odom.inverse() * self.base_2_odom
                self.lastmean = m.copy()
                self.lastbase_2_world = b2w
                self.lastodom_2_world = b2w * self.base_2_odom.inverse()
                self.amcl_out_lastbase_2_odom = self.lastupdatebase_2_odom
            else:
                self.lastbase_2_world = self.lastbase_2_world * self.amcl_out_lastbase_2_odom.inverse() * \
                    self.base_2_odom
                self.amcl_out_lastbase_2_odom = self.base_2_odom
        return {"base_t": list(self.lastbase_2_world.o), "base_q": self.lastbase_2_world.quat(),
                "odom_t": list(self.lastodom_2_world.o), "odom_q": self.lastodom_2_world.quat(),
                "amcl_out": self.amcl_out, "has_takenoff": self.has_takenoff}

    def range(self, r, ax, ay, az):
        self.ranges.append([np.float32(r), np.float32(ax), np.float32(ay), np.float32(az)])

    def _due(self, now_ns):
        if now_ns < self.nextupdate_ns:
            return False
        inc = self.lastupdatebase_2_odom.inverse() * self.base_2_odom
        o = inc.o
        if math.sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) > self.p["d_th"]:
            return True
        return abs(inc.rpy()[2]) > self.p["a_th"]

    def point_cloud(self, now_ns, xyz):
        if not self.is_odom or not self._due(now_ns):
            return False
        self.nextupdate_ns = now_ns + duration_ns(1.0 / self.p["update_rate"])
        cloud = oracle.voxel_grid(xyz, self.p["voxel_size"], order_mode=1)
        inc = self.lastupdatebase_2_odom.inverse() * self.base_2_odom
        self.deltas = [inc.o[0], inc.o[1], inc.o[2], inc.rpy()[2]]
        z4, u01 = self.draws(self.step)
        self.step += 1
        mods = [self.p[k] for k in ("odom_x_mod", "odom_y_mod", "odom_z_mod", "odom_a_mod")]
        self.particles = oracle.pf_predict(self.particles, mods, self.deltas, z4)
        ranges = np.array(self.ranges, np.float32).reshape(-1, 4)
        self.particles, self.mean = oracle.pf_update(self.particles, self.g, cloud, ranges, self.p["alpha"],
                                                     self.p["sensor_range"], self.roll, self.pitch)
        self.ranges = []
        self.lastupdatebase_2_odom = self.base_2_odom
        self.n_updates += 1
        if self.n_updates > self.p["resample_interval"]:
            self.n_updates = 0
            self.particles, _ = oracle.pf_resample(self.particles, u01)
        return True
