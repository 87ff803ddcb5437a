"""Serving: many localisation sessions with device-resident state and
continuous batching (the request loop runs in the library's worker thread).

A session is a ParticleFilter (amcl3d/src/ParticleFilter.h:67-213) that
lives on the device: it is initialised from a seed (ParticleFilter::init with
the session's step-0 draws) and every request runs predict -> update
(-> resample) with the session's next draws.  Draws are Philox4x32-10 keyed by
the seed and indexed by the session's step, so results do not depend on
which requests share a batch.  ``session_draws`` returns those draws, which
lets the plain ParticleFilter reproduce a session step for step.
"""
import ctypes

import numpy as np

from . import _lib


def session_draws(seed, step, n):
    """(z4 (n, 4) float32, u01 float) of a session's step (0 = init)."""
    z4 = np.empty((n, 4), np.float32)
    u = np.empty(1, np.float32)
    _lib.call("a3d_session_draws", int(seed), int(step), int(n), _lib.ptr(z4), _lib.ptr(u))
    return z4, float(u[0])


class Server:
    def __init__(self, hbm_budget_bytes=64 << 30, max_batch_evals=8_000_000_000, max_batch_sessions=512,
                 max_inflight=3, latest_only=False):
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(max_batch_evals), int(max_batch_sessions),
                                int(max_inflight), 1 if latest_only else 0)
        h = ctypes.c_void_p()
        _lib.call("a3d_server_create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h.value
        self._grids = []
        self._n = {}

    def close(self):
        if self._h:
            _lib.call("a3d_server_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_server_destroy(self._h)
        except Exception:
            pass

    def add_map(self, grid):
        mid = ctypes.c_int()
        _lib.call("a3d_server_add_map", self._h, grid.handle, ctypes.byref(mid))
        self._grids.append(grid)  # the server does not own the grid; keep it alive
        return mid.value

    def open_session(self, map_id, num_particles, pose, devs, seed):
        sid = ctypes.c_int()
        _lib.call("a3d_server_open_session", self._h, int(map_id), int(num_particles),
                  _lib.f32(pose), _lib.f32(devs), int(seed), ctypes.byref(sid))
        self._n[sid.value] = int(num_particles)
        return sid.value

    def open_session_particles(self, map_id, particles, seed):
        """A session from given (n, 7) particles (x, y, z, a, w, wp, wr)."""
        p = _lib.f32(particles).reshape(-1, 7)
        sid = ctypes.c_int()
        _lib.call("a3d_server_open_session_particles", self._h, int(map_id), p, p.shape[0], int(seed),
                  ctypes.byref(sid))
        self._n[sid.value] = int(p.shape[0])
        return sid.value

    def close_session(self, sid):
        _lib.call("a3d_server_close_session", self._h, int(sid))

    def submit(self, sid, mods, deltas, cloud_xyz, ranges, alpha, sigma, roll=0.0, pitch=0.0, resample=True):
        """cloud_xyz: tuple of three float32 arrays (x, y, z); ranges: (B, 4) or None."""
        x, y, z = cloud_xyz
        r = None if ranges is None else _lib.f32(ranges).reshape(-1, 4)
        nb = 0 if r is None else r.shape[0]
        t = ctypes.c_uint64()
        _lib.call("a3d_server_submit", self._h, int(sid), _lib.f64(mods), _lib.f64(deltas),
                  _lib.ptr(x), _lib.ptr(y), _lib.ptr(z), x.shape[0], _lib.ptr(r) if nb else None, nb, float(alpha),
                  float(sigma), float(roll), float(pitch), 1 if resample else 0, ctypes.byref(t))
        return t.value

    def poll(self, max_results=4096):
        buf = (_lib.Result * max_results)()
        n = ctypes.c_size_t()
        _lib.call("a3d_server_poll", self._h, buf, max_results, ctypes.byref(n))
        return [buf[i] for i in range(n.value)]

    def drain(self):
        _lib.call("a3d_server_drain", self._h)

    def particles(self, sid):
        n = self._n[sid]
        out = np.empty((n, 7), np.float32)
        _lib.call("a3d_server_get_particles", self._h, int(sid), _lib.ptr(out), n)
        return out

    def stats(self):
        r, b, e, ms = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        _lib.call("a3d_server_stats", self._h, ctypes.byref(r), ctypes.byref(b), ctypes.byref(e), ctypes.byref(ms))
        return {"requests": r.value, "batches": b.value, "evals": e.value, "busy_ms": ms.value}
