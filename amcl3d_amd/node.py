"""Node: the ROS node's glue around the filter, without ROS.

Mirrors amcl3d::Node's callbacks (amcl3d/src/Node.cpp:112-398) over the C
ABI: ``odomCallback`` (odometry transform in, the published base and odom
poses in the world out), ``rangeCallback``, ``pointcloudCallback`` (update
gating by time, distance and yaw thresholds, VoxelGrid, predict from the
odometry increment, update, resample every ``resample_interval + 1``
updates) and ``setInitialPose``.  ROS time is the caller's ``now_ns``.
"""
import ctypes

import numpy as np

from . import _lib

PARAM_NAMES = [f[0] for f in _lib.NodeParams._fields_]


class Node:
    def __init__(self, params, grid, seed=0):
        """params: dict of amcl3d::Parameters values (Parameters.h:33-73) by the
        field names of a3d_node_params."""
        p = _lib.NodeParams(**{k: params[k] for k in PARAM_NAMES})
        h = ctypes.c_void_p()
        _lib.call("a3d_node_create", ctypes.byref(p), grid.handle, int(seed), ctypes.byref(h))
        self._h = h.value
        self.grid = grid  # kept alive while the node uses it

    def close(self):
        if self._h:
            _lib.call("a3d_node_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_node_destroy(self._h)
        except Exception:
            pass

    def odomCallback(self, t, q):
        """Returns what the node publishes, or None before the filter is initialised."""
        out = _lib.NodeOutput()
        _lib.call("a3d_node_odometry", self._h, _lib.f64(t), _lib.f64(q), ctypes.byref(out))
        if not out.published:
            return None
        return {"base_t": list(out.base_t), "base_q": list(out.base_q), "odom_t": list(out.odom_t),
                "odom_q": list(out.odom_q), "amcl_out": bool(out.amcl_out), "has_takenoff": bool(out.has_takenoff)}

    def rangeCallback(self, r, ax, ay, az):
        _lib.call("a3d_node_range", self._h, float(r), float(ax), float(ay), float(az))

    def pointcloudCallback(self, now_ns, points):
        """points: (n, 3) or (n, 4) float32 (pcl::PointXYZ).  True when an update ran."""
        pts = _lib.f32(points)
        pts = pts.reshape(pts.shape[0], -1) if pts.size else np.zeros((0, 3), np.float32)
        upd = ctypes.c_int()
        _lib.call("a3d_node_point_cloud", self._h, int(now_ns), _lib.ptr(pts) if pts.shape[0] else None,
                  pts.shape[0], pts.shape[1] * 4, ctypes.byref(upd))
        return bool(upd.value)

    def setInitialPose(self, t, q, xdev, ydev, zdev, adev):
        _lib.call("a3d_node_set_initial_pose", self._h, _lib.f64(t), _lib.f64(q), float(xdev), float(ydev),
                  float(zdev), float(adev))

    def state(self):
        mean = np.zeros(4, np.float32)
        deltas = np.zeros(4)
        roll, pitch = ctypes.c_double(), ctypes.c_double()
        is_odom, out = ctypes.c_int(), ctypes.c_int()
        step = ctypes.c_uint64()
        _lib.call("a3d_node_get_state", self._h, _lib.ptr(mean), _lib.ptr(deltas), ctypes.byref(roll),
                  ctypes.byref(pitch), ctypes.byref(is_odom), ctypes.byref(out), ctypes.byref(step))
        return {"mean": mean, "deltas": deltas, "roll": roll.value, "pitch": pitch.value,
                "is_odom": bool(is_odom.value), "amcl_out": bool(out.value), "draw_step": step.value}

    def particles(self):
        pf = ctypes.c_void_p()
        _lib.call("a3d_node_get_filter", self._h, ctypes.byref(pf))
        n = ctypes.c_size_t()
        _lib.call("a3d_pf_size", pf.value, ctypes.byref(n))
        out = np.empty((n.value, 7), np.float32)
        if n.value:
            _lib.call("a3d_pf_get_particles", pf.value, _lib.ptr(out), n.value)
        return out
