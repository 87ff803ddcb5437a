"""The oracle: a CPU restatement of amcl3d's measurement-update hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline).  The product path (``amcl3d_amd``) never
imports it and fails loudly when its HIP library is missing.

``liboracle.so`` is built from ``amcl3d_oracle.cpp`` (``make -C oracle``) and
restates, with the reference's float/double mixing:

* ``computeGrid``          amcl3d/src/PointCloudTools.cpp:84-149
* ``computeCloudWeight``   amcl3d/src/Grid3d.cpp:133-199
* ``isIntoMap``            amcl3d/src/Grid3d.cpp:201-208
* ``init``                 amcl3d/src/ParticleFilter.cpp:46-95
* ``predict``              amcl3d/src/ParticleFilter.cpp:97-119
* ``update``               amcl3d/src/ParticleFilter.cpp:121-196
* ``resample``             amcl3d/src/ParticleFilter.cpp:198-222
* ``computeRangeWeight``   amcl3d/src/ParticleFilter.cpp:224-244
* ``voxel_grid``           pcl::VoxelGrid as amcl3d/src/Node.cpp:130-141 calls it
                           (PCL 1.7.2, not vendored; restated)

``fp64`` holds an independent double-precision oracle used only to state
errors.  How each function is pinned (known-answer values from the reference's
own tests, closed forms, brute force) is listed in the tests that use them.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_grid_sizes.argtypes = [_f64p, _f64p, ctypes.c_double, _u32p]
        L.orc_compute_grid.argtypes = [_f32p, _sz, _f64p, _f64p, ctypes.c_double, ctypes.c_double, _f32p,
                                       _f32p, ctypes.c_int, ctypes.c_int]
        L.orc_compute_grid.restype = ctypes.c_int64
        L.orc_nearest_d2.argtypes = [_f32p, _sz, _f64p, _f64p, ctypes.c_double, _f32p, _sz, _f32p]
        L.orc_cloud_weight.argtypes = [_vp, _u32p, _f64p, _f64p, ctypes.c_double, _f32p, _f32p, _f32p, _sz,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float]
        L.orc_cloud_weight.restype = ctypes.c_float
        L.orc_cloud_weight_batch.argtypes = [_f32p, _u32p, _f64p, _f64p, ctypes.c_double, _f32p, _f32p, _f32p,
                                             _sz, _f32p, _f32p, _f32p, _f32p, _sz, ctypes.c_float,
                                             ctypes.c_float, _f32p, ctypes.c_int]
        L.orc_cloud_cells.argtypes = [_u32p, _f64p, _f64p, ctypes.c_double, _f32p, _f32p, _f32p, _sz, _f32p, _f32p,
                                      _f32p, _f32p, _sz, ctypes.c_float, ctypes.c_float,
                                      np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]
        L.orc_is_into_map.argtypes = [_u32p, _f64p, _f64p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float]
        L.orc_is_into_map.restype = ctypes.c_int
        L.orc_range_weight.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p, _sz,
                                       ctypes.c_double]
        L.orc_range_weight.restype = ctypes.c_float
        L.orc_pf_init.argtypes = [_f32p, ctypes.c_int] + [ctypes.c_float] * 8 + [_f32p, _f32p]
        L.orc_pf_predict.argtypes = [_f32p, ctypes.c_int, _f64p, _f64p, _f32p]
        L.orc_pf_update.argtypes = [_f32p, ctypes.c_int, _vp, _u32p, _f64p, _f64p, ctypes.c_double, _f32p,
                                    _f32p, _f32p, _sz, _vp, _sz, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _f32p, ctypes.c_int]
        L.orc_resample_indices.argtypes = [_f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int, _u32p]
        L.orc_pf_resample.argtypes = [_f32p, ctypes.c_int, ctypes.c_float, ctypes.c_int, _vp]
        L.orc_slice_int8.argtypes = [_f32p, _sz, _i8p]
        L.orc_voxel_grid.argtypes = [_f32p, _sz, ctypes.c_float, ctypes.c_int, _f32p]
        L.orc_voxel_grid.restype = _sz
        _LIB = L
    return _LIB


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _nthreads(n):
    return int(n) if n else max(1, min(16, os.cpu_count() or 1))


def grid_sizes(mn, mx, res):
    out = np.zeros(3, np.uint32)
    lib().orc_grid_sizes(_f64(mn), _f64(mx), float(res), out)
    return out


def compute_grid(points, mn, mx, res, sensor_dev, nthreads=None, exp_mode=0):
    """computeGrid: returns (dist, prob), each shaped (sz, sy, sx) float32."""
    pts = _f32(points).reshape(-1, 3)
    s = grid_sizes(mn, mx, res)
    ncell = int(s[0]) * int(s[1]) * int(s[2])
    dist = np.empty(ncell, np.float32)
    prob = np.empty(ncell, np.float32)
    lib().orc_compute_grid(pts.reshape(-1), pts.shape[0], _f64(mn), _f64(mx), float(res), float(sensor_dev),
                           dist, prob, _nthreads(nthreads), int(exp_mode))
    shape = (int(s[2]), int(s[1]), int(s[0]))
    return dist.reshape(shape), prob.reshape(shape)


def nearest_d2(points, mn, mx, res, queries):
    pts = _f32(points).reshape(-1, 3)
    q = _f32(queries).reshape(-1, 3)
    out = np.empty(q.shape[0], np.float32)
    lib().orc_nearest_d2(pts.reshape(-1), pts.shape[0], _f64(mn), _f64(mx), float(res), q.reshape(-1),
                         q.shape[0], out)
    return out


class GridMeta:
    """Bounds, resolution and sizes of a likelihood grid (Grid3dInfo + PointCloudInfo)."""

    def __init__(self, mn, mx, res, prob=None):
        self.min = _f64(mn)
        self.max = _f64(mx)
        self.res = float(res)
        self.size = grid_sizes(mn, mx, res)
        self.prob = None if prob is None else _f32(prob).reshape(-1)

    def _probp(self):
        return None if self.prob is None else self.prob.ctypes.data


def cloud_weight(g, cloud, tx, ty, tz, roll, pitch, yaw):
    c = _f32(cloud).reshape(-1, 3)
    x, y, z = (np.ascontiguousarray(c[:, k]) for k in range(3))
    return float(lib().orc_cloud_weight(g._probp(), g.size, g.min, g.max, g.res, x, y, z, c.shape[0], tx, ty, tz,
                                        roll, pitch, yaw))


def cloud_weight_batch(g, cloud, tx, ty, tz, yaw, roll=0.0, pitch=0.0, nthreads=None):
    c = _f32(cloud).reshape(-1, 3)
    x, y, z = (np.ascontiguousarray(c[:, k]) for k in range(3))
    n = len(tx)
    out = np.empty(n, np.float32)
    lib().orc_cloud_weight_batch(g.prob, g.size, g.min, g.max, g.res, x, y, z, c.shape[0], _f32(tx), _f32(ty),
                                 _f32(tz), _f32(yaw), n, roll, pitch, out, _nthreads(nthreads))
    return out


def cloud_cells(g, cloud, tx, ty, tz, yaw, roll=0.0, pitch=0.0):
    """The cells computeCloudWeight gathers: (n, m) int64, the cell index of
    point k under pose i, or -1 for a point outside the map.  Needs only the
    grid's geometry (g.prob may be None)."""
    c = _f32(cloud).reshape(-1, 3)
    x, y, z = (np.ascontiguousarray(c[:, k]) for k in range(3))
    n = len(tx)
    out = np.empty((n, c.shape[0]), np.int64)
    lib().orc_cloud_cells(g.size, g.min, g.max, g.res, x, y, z, c.shape[0], _f32(tx), _f32(ty), _f32(tz), _f32(yaw),
                          n, roll, pitch, out.reshape(-1))
    return out


def is_into_map(g, x, y, z, open_=True):
    return bool(lib().orc_is_into_map(g.size, g.min, g.max, 1 if open_ else 0, x, y, z))


def range_weight(x, y, z, ranges, sigma):
    r = _f32(ranges).reshape(-1, 4)
    return float(lib().orc_range_weight(x, y, z, r.reshape(-1), r.shape[0], float(sigma)))


def pf_init(n, pose, devs, z4):
    """Returns (particles (n,7) float32, mean (4,))."""
    p = np.zeros((n, 7), np.float32)
    mean = np.zeros(4, np.float32)
    lib().orc_pf_init(p.reshape(-1), int(n), *[float(v) for v in pose], *[float(v) for v in devs],
                      _f32(z4).reshape(-1), mean)
    return p, mean


def pf_predict(particles, mods, deltas, z4):
    p = _f32(particles).copy()
    lib().orc_pf_predict(p.reshape(-1), p.shape[0], _f64(mods), _f64(deltas), _f32(z4).reshape(-1))
    return p


def pf_update(particles, g, cloud, ranges, alpha, sigma, roll=0.0, pitch=0.0, nthreads=None):
    """Returns (particles, mean)."""
    p = _f32(particles).copy()
    c = _f32(cloud).reshape(-1, 3)
    x, y, z = (np.ascontiguousarray(c[:, k]) for k in range(3))
    r = _f32(ranges).reshape(-1, 4) if ranges is not None else np.zeros((0, 4), np.float32)
    mean = np.zeros(4, np.float32)
    lib().orc_pf_update(p.reshape(-1), p.shape[0], g._probp(), g.size, g.min, g.max, g.res, x, y, z, c.shape[0],
                        r.ctypes.data if r.shape[0] else None, r.shape[0], float(alpha), float(sigma),
                        float(roll), float(pitch), mean, _nthreads(nthreads))
    return p, mean


def resample_indices(w, u01, mode=0):
    w = _f32(w)
    out = np.empty(w.shape[0], np.uint32)
    lib().orc_resample_indices(w, w.shape[0], float(u01), int(mode), out)
    return out


def pf_resample(particles, u01, mode=0):
    """Returns (new particles, indices)."""
    p = _f32(particles).copy()
    idx = np.empty(p.shape[0], np.uint32)
    lib().orc_pf_resample(p.reshape(-1), p.shape[0], float(u01), int(mode), idx.ctypes.data)
    return p, idx


def slice_int8(prob_plane):
    pl = _f32(prob_plane).reshape(-1)
    out = np.empty(pl.shape[0], np.int8)
    lib().orc_slice_int8(pl, pl.shape[0], out)
    return out


def voxel_grid(points, leaf, order_mode=0):
    """pcl::VoxelGrid<PointXYZ> with leaf size `leaf` (Node.cpp:130-141): the
    centroid of each occupied voxel, in voxel-index order, (k, 3) float32.
    order_mode 0: PCL's std::sort order inside a voxel; 1: point-index order."""
    p = _f32(points).reshape(-1, 3)
    out = np.empty(max(1, p.shape[0]) * 3, np.float32)
    k = lib().orc_voxel_grid(p.reshape(-1), p.shape[0], float(np.float32(leaf)), int(order_mode), out)
    return out[:3 * k].reshape(-1, 3)
