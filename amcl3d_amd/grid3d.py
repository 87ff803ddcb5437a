"""Grid3d: the device-resident likelihood grid.

Mirrors ``amcl3d::Grid3d`` (amcl3d/src/Grid3d.h:34-159) on top of the C ABI.
``open(map_path, sensor_dev)`` is the reference's: the OcTree file (.bt/.ot),
then its .grid cache or computeGrid (amcl3d/src/PointCloudTools.cpp:84-149) on
the GPU.  ``open(points, min, max, res, sensor_dev)`` builds the grid from a
map already read (the occupied points and bounds computePointCloud gives).
"""
import ctypes
import os

import numpy as np

from . import _lib


class Grid3d:
    def __init__(self):
        self._h = None
        self.size = None
        self.min = None
        self.max = None
        self.res = None
        self.sensor_dev = None
        self.last_error = None

    # -- construction ------------------------------------------------------
    def open(self, points, min_xyz=None, max_xyz=None, res=None, sensor_dev=None, keep_dist=False):
        """Grid3d::open (Grid3d.cpp:25-78).

        ``open(map_path, sensor_dev)``: the OcTree file, the .grid cache beside
        it, else computeGrid on the device and the cache written.  Returns
        False when the file cannot be used (the reason is in ``last_error``),
        like the reference; device failures raise.

        ``open(points, min, max, res, sensor_dev)``: computeGrid on the device
        from the map's occupied points.  Returns True."""
        self.close()
        if isinstance(points, (str, bytes, os.PathLike)):
            sd = min_xyz if sensor_dev is None else sensor_dev
            h = ctypes.c_void_p()
            path = os.fsencode(points)
            try:
                _lib.call("a3d_grid_open", path, float(sd), 1 if keep_dist else 0, ctypes.byref(h))
            except _lib.Amcl3dError as e:
                if e.code != _lib.A3D_ERR_IO:
                    raise
                self.last_error = str(e)
                return False
            self._h = h.value
            self._read_info()
            return True
        pts = _lib.f32(points).reshape(-1, 3)
        h = ctypes.c_void_p()
        _lib.call("a3d_grid_create_from_points", _lib.ptr(pts) if pts.shape[0] else None, pts.shape[0],
                  _lib.f64(min_xyz), _lib.f64(max_xyz), float(res), float(sensor_dev),
                  1 if keep_dist else 0, ctypes.byref(h))
        self._h = h.value
        self._read_info()
        return True

    def from_cells(self, prob, min_xyz, max_xyz, res, sensor_dev, dist=None):
        """Grid from precomputed cells, x fastest (the loadGrid path, Grid3d.cpp:239-275)."""
        self.close()
        p = _lib.f32(prob).reshape(-1)
        d = None if dist is None else _lib.f32(dist).reshape(-1)
        size = np.ceil((np.asarray(max_xyz, np.float64) - np.asarray(min_xyz, np.float64)) / res)
        sz = _lib.u32(size)
        h = ctypes.c_void_p()
        _lib.call("a3d_grid_create_from_cells", _lib.ptr(p), _lib.ptr(d), _lib.ptr(sz), _lib.f64(min_xyz),
                  _lib.f64(max_xyz), float(res), float(sensor_dev), ctypes.byref(h))
        self._h = h.value
        self._read_info()
        return self

    def load(self, path, min_xyz, max_xyz, res, sensor_dev):
        """Grid3d::loadGrid (Grid3d.cpp:239-275): the reference's .grid file."""
        self.close()
        h = ctypes.c_void_p()
        _lib.call("a3d_grid_load", str(path).encode(), _lib.f64(min_xyz), _lib.f64(max_xyz),
                  float(res), float(sensor_dev), ctypes.byref(h))
        self._h = h.value
        self._read_info()
        return self

    def save(self, path):
        """Grid3d::saveGrid (Grid3d.cpp:210-237)."""
        _lib.call("a3d_grid_save", self.handle, str(path).encode())

    def close(self):
        if self._h:
            _lib.call("a3d_grid_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_grid_destroy(self._h)
        except Exception:
            pass

    def _read_info(self):
        sz = np.zeros(3, np.uint32)
        mn = np.zeros(3)
        mx = np.zeros(3)
        res = ctypes.c_double()
        sd = ctypes.c_double()
        _lib.call("a3d_grid_get_info", self._h, _lib.ptr(sz), _lib.ptr(mn), _lib.ptr(mx), ctypes.byref(res),
                  ctypes.byref(sd))
        self.size, self.min, self.max, self.res, self.sensor_dev = sz, mn, mx, res.value, sd.value

    @property
    def handle(self):
        if not self._h:
            raise _lib.Amcl3dError(_lib.A3D_ERR_STATE, "grid is not open")
        return self._h

    @property
    def is_open(self):
        return bool(self._h)

    @property
    def cells(self):
        return int(self.size[0]) * int(self.size[1]) * int(self.size[2])

    # -- queries -----------------------------------------------------------
    def read_cells(self, first=0, count=None, dist=False):
        """prob (and dist) of cells [first, first+count), x fastest."""
        count = self.cells - first if count is None else count
        p = np.empty(count, np.float32)
        d = np.empty(count, np.float32) if dist else None
        _lib.call("a3d_grid_read_cells", self.handle, int(first), int(count), _lib.ptr(p), _lib.ptr(d))
        return (p, d) if dist else p

    def prob_volume(self):
        return self.read_cells().reshape(int(self.size[2]), int(self.size[1]), int(self.size[0]))

    def isIntoMap(self, x, y, z):
        """Grid3d::isIntoMap (Grid3d.cpp:201-208); False when not open."""
        out = ctypes.c_int()
        _lib.call("a3d_grid_is_into_map", self._h, float(x), float(y), float(z), ctypes.byref(out))
        return bool(out.value)

    def computeCloudWeight(self, cloud, tx, ty, tz, roll, pitch, yaw):
        """Grid3d::computeCloudWeight (Grid3d.cpp:133-199) for one pose; 0 when not open."""
        return float(self.cloud_weights(cloud, [tx], [ty], [tz], [yaw], roll, pitch)[0])

    def cloud_weights(self, cloud, tx, ty, tz, yaw, roll=0.0, pitch=0.0):
        """computeCloudWeight for many poses sharing one cloud, in one launch."""
        x, y, z = _lib.cloud_soa(cloud)
        tx, ty, tz, yaw = (_lib.f32(v).reshape(-1) for v in (tx, ty, tz, yaw))
        out = np.empty(tx.shape[0], np.float32)
        _lib.call("a3d_grid_cloud_weight", self._h, _lib.ptr(x), _lib.ptr(y), _lib.ptr(z), x.shape[0],
                  _lib.ptr(tx), _lib.ptr(ty), _lib.ptr(tz), _lib.ptr(yaw), tx.shape[0], float(roll), float(pitch),
                  _lib.ptr(out))
        return out
