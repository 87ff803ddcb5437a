"""Fleet: many independent filters on one resident grid, updated per launch.

Each filter is a ParticleFilter (amcl3d/src/ParticleFilter.h:67-213); the
fleet updates all of them with one persistent cloud-weight launch over a
work list of particle blocks, longest first.  Results equal the single-filter
path's bit for bit.
"""
import ctypes

import numpy as np

from . import _lib

PERSISTENT = 1     # one persistent launch, atomic work list
ONE_PER_ITEM = 0   # one launch, one wavefront per work item
PLAIN_LOOP = -1    # the plain single-filter kernel launched once per filter


class Fleet:
    def __init__(self, counts):
        self.counts = _lib.u32(counts)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts.astype(np.int64))])
        h = ctypes.c_void_p()
        _lib.call("a3d_fleet_create", len(self.counts), _lib.ptr(self.counts), ctypes.byref(h))
        self._h = h.value

    def close(self):
        if self._h:
            _lib.call("a3d_fleet_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_fleet_destroy(self._h)
        except Exception:
            pass

    @property
    def total(self):
        return int(self.offsets[-1])

    def set_particles(self, p7):
        _lib.call("a3d_fleet_set_particles", self._h, _lib.f32(p7).reshape(-1, 7))

    def particles(self):
        out = np.empty((self.total, 7), np.float32)
        _lib.call("a3d_fleet_get_particles", self._h, _lib.ptr(out))
        return out

    def means(self):
        out = np.empty((len(self.counts), 4), np.float32)
        _lib.call("a3d_fleet_get_means", self._h, _lib.ptr(out))
        return out

    def stage_update(self, clouds, ranges, alpha, sigma, roll=None, pitch=None):
        """clouds: list of (m_f, 3) arrays; ranges: list of (b_f, 4) arrays."""
        F = len(self.counts)
        m = _lib.u32([c.shape[0] for c in clouds])
        nb = _lib.u32([0 if r is None else r.shape[0] for r in ranges])
        cl = np.concatenate([_lib.f32(c).reshape(-1, 3) for c in clouds]) if m.sum() else np.zeros((0, 3), np.float32)
        x, y, z = _lib.cloud_soa(cl)
        rr = [r for r in ranges if r is not None and r.shape[0]]
        R = _lib.f32(np.concatenate(rr)) if rr else np.zeros((0, 4), np.float32)
        roll = np.zeros(F) if roll is None else roll
        pitch = np.zeros(F) if pitch is None else pitch
        _lib.call("a3d_fleet_stage_update", self._h, _lib.ptr(x), _lib.ptr(y), _lib.ptr(z), _lib.ptr(m),
                  _lib.ptr(R) if R.shape[0] else None, _lib.ptr(nb), _lib.f64(np.broadcast_to(alpha, F)),
                  _lib.f64(np.broadcast_to(sigma, F)), _lib.f64(roll), _lib.f64(pitch))
        self.evals = int(np.sum(self.counts.astype(np.int64) * m.astype(np.int64)))

    def update_staged(self, grid, mode=PERSISTENT):
        _lib.call("a3d_fleet_update_staged", self._h, grid.handle, int(mode))

    def synchronize(self):
        _lib.call("a3d_fleet_synchronize", self._h)

    def resample(self, u01, mode=_lib.RESAMPLE_REFERENCE):
        _lib.call("a3d_fleet_resample", self._h, _lib.f32(u01), int(mode))

    def predict(self, mods, deltas, z4):
        """mods, deltas: (F, 4) per filter; z4: (total, 4) normals, filter by filter."""
        _lib.call("a3d_fleet_predict", self._h, _lib.f64(np.ascontiguousarray(mods).reshape(-1)),
                  _lib.f64(np.ascontiguousarray(deltas).reshape(-1)), _lib.f32(np.ascontiguousarray(z4).reshape(-1)))

    # device-parameter forms (capturable): see a3d_fleet_set_layout
    def set_layout(self, m, nb):
        self.layout_m = _lib.u32(m)
        self.layout_nb = _lib.u32(nb)
        _lib.call("a3d_fleet_set_layout", self._h, self.layout_m, self.layout_nb)
        self.evals = int(np.sum(self.counts.astype(np.int64) * self.layout_m.astype(np.int64)))

    def predict_dev(self, d_mods, d_deltas, d_z4):
        """Device pointers: 4 doubles per filter (twice), 4 normals per particle."""
        _lib.call("a3d_fleet_predict_dev", self._h, d_mods, d_deltas, d_z4)

    def update_dev(self, grid, d_x, d_y, d_z, d_ranges, d_params):
        """Device pointers: clouds and ranges concatenated in the layout's
        filter order; params {alpha, sigma, roll, pitch} doubles per filter."""
        _lib.call("a3d_fleet_update_dev", self._h, grid.handle, d_x, d_y, d_z, d_ranges, d_params)

    def resample_dev(self, d_u01, mode=_lib.RESAMPLE_REFERENCE):
        _lib.call("a3d_fleet_resample_dev", self._h, d_u01, int(mode))

    def capture_begin(self):
        _lib.call("a3d_fleet_capture_begin", self._h)

    def capture_end(self):
        from .particle_filter import Graph

        g = ctypes.c_void_p()
        _lib.call("a3d_fleet_capture_end", self._h, ctypes.byref(g))
        return Graph(g.value)

    def timing(self, reset=False):
        ms, n, e = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        _lib.call("a3d_fleet_timing", self._h, 1 if reset else 0, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(e))
        return ms.value, n.value, e.value
