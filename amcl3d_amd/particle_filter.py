"""ParticleFilter: device-resident particles with the reference's stages.

Mirrors ``amcl3d::ParticleFilter`` (amcl3d/src/ParticleFilter.h:67-213):
init / predict / update / resample / getMean / isInitialized.  The random
draws that the reference takes from std::mt19937 (ParticleFilter.cpp:246-256)
are arguments: ``z4`` holds 4 standard normals per particle (x, y, z, yaw)
and ``u01`` the uniform of resample.  Particles are read as the reference's
struct Particle {x, y, z, a, w, wp, wr}.
"""
import ctypes

import numpy as np

from . import _lib

PARTICLE_FIELDS = ("x", "y", "z", "a", "w", "wp", "wr")


class ParticleFilter:
    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("a3d_pf_create", ctypes.byref(h))
        self._h = h.value

    def close(self):
        if self._h:
            _lib.call("a3d_pf_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_pf_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def isInitialized(self):
        v = ctypes.c_int()
        _lib.call("a3d_pf_is_initialized", self._h, ctypes.byref(v))
        return bool(v.value)

    def size(self):
        v = ctypes.c_size_t()
        _lib.call("a3d_pf_size", self._h, ctypes.byref(v))
        return v.value

    def init(self, num_particles, x_init, y_init, z_init, a_init, x_dev, y_dev, z_dev, a_dev, z4):
        """ParticleFilter::init (ParticleFilter.cpp:46-95)."""
        z = _lib.f32(z4).reshape(-1)
        _lib.call("a3d_pf_init", self._h, int(num_particles), float(x_init), float(y_init), float(z_init),
                  float(a_init), float(x_dev), float(y_dev), float(z_dev), float(a_dev), _lib.ptr(z))

    def predict(self, odom_x_mod, odom_y_mod, odom_z_mod, odom_a_mod, delta_x, delta_y, delta_z, delta_a, z4):
        """ParticleFilter::predict (ParticleFilter.cpp:97-119)."""
        z = _lib.f32(z4).reshape(-1)
        _lib.call("a3d_pf_predict", self._h, float(odom_x_mod), float(odom_y_mod), float(odom_z_mod),
                  float(odom_a_mod), float(delta_x), float(delta_y), float(delta_z), float(delta_a), _lib.ptr(z))

    def update(self, grid3d, cloud, range_data, alpha, sigma, roll=0.0, pitch=0.0):
        """ParticleFilter::update (ParticleFilter.cpp:121-196).  range_data: (B, 4)
        rows {r, ax, ay, az} or None."""
        x, y, z = _lib.cloud_soa(cloud)
        r = np.zeros((0, 4), np.float32) if range_data is None else _lib.f32(range_data).reshape(-1, 4)
        gh = grid3d.handle if (grid3d is not None and grid3d.is_open) else None
        _lib.call("a3d_pf_update", self._h, gh, _lib.ptr(x), _lib.ptr(y), _lib.ptr(z), x.shape[0],
                  _lib.ptr(r) if r.shape[0] else None, r.shape[0], float(alpha), float(sigma), float(roll),
                  float(pitch))

    def resample(self, u01, mode=_lib.RESAMPLE_REFERENCE, return_indices=False):
        """ParticleFilter::resample (ParticleFilter.cpp:198-222)."""
        idx = np.empty(self.size(), np.uint32) if return_indices else None
        _lib.call("a3d_pf_resample", self._h, float(u01), int(mode), _lib.ptr(idx))
        return idx

    def getMean(self):
        """ParticleFilter::getMean (ParticleFilter.h:94-97): (x, y, z, a)."""
        m = np.zeros(4, np.float32)
        _lib.call("a3d_pf_get_mean", self._h, _lib.ptr(m))
        return m

    def particles(self):
        """(N, 7) float32 rows {x, y, z, a, w, wp, wr}."""
        n = self.size()
        out = np.empty((n, 7), np.float32)
        _lib.call("a3d_pf_get_particles", self._h, _lib.ptr(out), n)
        return out

    def set_particles(self, p7):
        p = _lib.f32(p7).reshape(-1, 7)
        _lib.call("a3d_pf_set_particles", self._h, _lib.ptr(p), p.shape[0])

    # -- resident-input (device buffer) forms, graph capture ---------------
    def predict_dev(self, mods, deltas, d_z4):
        _lib.call("a3d_pf_predict_dev", self._h, _lib.f64(mods), _lib.f64(deltas), d_z4.ptr)

    def update_dev(self, grid3d, d_x, d_y, d_z, m, d_ranges, nb, alpha, sigma, roll=0.0, pitch=0.0):
        _lib.call("a3d_pf_update_dev", self._h, grid3d.handle, d_x, d_y, d_z, int(m), d_ranges, int(nb),
                  float(alpha), float(sigma), float(roll), float(pitch))

    def resample_dev(self, u01, mode=_lib.RESAMPLE_REFERENCE):
        _lib.call("a3d_pf_resample_dev", self._h, float(u01), int(mode))

    def synchronize(self):
        _lib.call("a3d_pf_synchronize", self._h)

    def capture_begin(self):
        _lib.call("a3d_pf_capture_begin", self._h)

    def capture_end(self):
        g = ctypes.c_void_p()
        _lib.call("a3d_pf_capture_end", self._h, ctypes.byref(g))
        return Graph(g.value)

    def timing(self, reset=False):
        """(average cloud-weight kernel ms, launches, evals) of _dev updates."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        e = ctypes.c_uint64()
        _lib.call("a3d_pf_timing", self._h, 1 if reset else 0, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(e))
        return ms.value, n.value, e.value


class Graph:
    """A captured chain of _dev calls (hipGraph)."""

    def __init__(self, h):
        self._h = h

    def launch(self):
        _lib.call("a3d_graph_launch", self._h)

    def close(self):
        if self._h:
            _lib.call("a3d_graph_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_graph_destroy(self._h)
        except Exception:
            pass
