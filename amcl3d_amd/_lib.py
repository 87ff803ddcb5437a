"""ctypes binding of libamcl3d_gpu.so (the C ABI in include/amcl3d_gpu.h).

The library is the only implementation: there is no CPU fallback.  Loading
fails with ``LibraryMissing`` when the shared object has not been built
(``python -c "import __graft_entry__ as g; g.build()"`` builds it).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libamcl3d_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "amcl3d_gpu.h")

A3D_OK = 0
A3D_ERR_ARG = 1
A3D_ERR_HIP = 2
A3D_ERR_STATE = 3
A3D_ERR_NOMEM = 4
A3D_ERR_IO = 5
A3D_ERR_BUSY = 6

RESAMPLE_REFERENCE = 0
RESAMPLE_FP64 = 1

_ERRNAMES = {1: "ARG", 2: "HIP", 3: "STATE", 4: "NOMEM", 5: "IO", 6: "BUSY"}


class LibraryMissing(RuntimeError):
    pass


class Amcl3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("A3D_ERR_%s: %s" % (_ERRNAMES.get(code, code), msg))
        self.code = code


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch_evals", ctypes.c_uint64),
                ("max_batch_sessions", ctypes.c_int), ("max_inflight", ctypes.c_int), ("latest_only", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("session", ctypes.c_int), ("mean", ctypes.c_float * 4),
                ("latency_ms", ctypes.c_double), ("submit_ms", ctypes.c_double), ("evals", ctypes.c_uint64),
                ("deltas", ctypes.c_double * 4), ("superseded", ctypes.c_int)]


class NodeParams(ctypes.Structure):
    _fields_ = [("num_particles", ctypes.c_int), ("resample_interval", ctypes.c_int),
                ("set_initial_pose", ctypes.c_int)] + \
               [(k, ctypes.c_double) for k in ("init_x", "init_y", "init_z", "init_a", "init_x_dev", "init_y_dev",
                                               "init_z_dev", "init_a_dev", "sensor_range", "voxel_size",
                                               "odom_x_mod", "odom_y_mod", "odom_z_mod", "odom_a_mod",
                                               "update_rate", "d_th", "a_th", "take_off_height", "alpha")]


class NodeOutput(ctypes.Structure):
    _fields_ = [("published", ctypes.c_int), ("amcl_out", ctypes.c_int), ("has_takenoff", ctypes.c_int),
                ("base_t", ctypes.c_double * 3), ("base_q", ctypes.c_double * 4),
                ("odom_t", ctypes.c_double * 3), ("odom_q", ctypes.c_double * 4)]


_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_u64 = ctypes.c_uint64
_d = ctypes.c_double
_f = ctypes.c_float
_p = ctypes.POINTER

# name -> argument types (every function returns int unless listed in _RESTYPES)
SIGNATURES = {
    "a3d_last_error": [],
    "a3d_version": [],
    "a3d_set_device": [_i],
    "a3d_device_info": [ctypes.c_char_p, _sz, _p(_i), _p(_sz)],
    "a3d_device_synchronize": [],
    "a3d_buffer_alloc": [_sz, _p(_vp)],
    "a3d_buffer_free": [_vp],
    "a3d_buffer_upload": [_vp, _vp, _sz],
    "a3d_buffer_download": [_vp, _vp, _sz],
    "a3d_buffer_copy": [_vp, _vp, _sz],
    "a3d_grid_create_from_points": [_vp, _sz, _vp, _vp, _d, _d, _i, _p(_vp)],
    "a3d_grid_create_from_cells": [_vp, _vp, _vp, _vp, _vp, _d, _d, _p(_vp)],
    "a3d_grid_load": [ctypes.c_char_p, _vp, _vp, _d, _d, _p(_vp)],
    "a3d_grid_save": [_vp, ctypes.c_char_p],
    "a3d_grid_destroy": [_vp],
    "a3d_grid_get_info": [_vp, _vp, _vp, _vp, _p(_d), _p(_d)],
    "a3d_grid_read_cells": [_vp, _u64, _u64, _vp, _vp],
    "a3d_grid_is_into_map": [_vp, _f, _f, _f, _p(_i)],
    "a3d_grid_cloud_weight": [_vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _sz, _f, _f, _vp],
    "a3d_octree_read": [ctypes.c_char_p, _p(_vp)],
    "a3d_octree_get_info": [_vp, _vp, _vp, _p(_d), _p(_u64), _p(_u64)],
    "a3d_octree_get_points": [_vp, _u64, _u64, _vp],
    "a3d_octree_destroy": [_vp],
    "a3d_grid_open": [ctypes.c_char_p, _d, _i, _p(_vp)],
    "a3d_voxel_grid": [_vp, _sz, _sz, _f, _vp, _vp, _vp, _p(_sz)],
    "a3d_voxel_grid_dev": [_vp, _sz, _sz, _f, _vp, _vp, _vp, _p(_sz)],
    "a3d_pf_create": [_p(_vp)],
    "a3d_pf_destroy": [_vp],
    "a3d_pf_init": [_vp, _i, _f, _f, _f, _f, _f, _f, _f, _f, _vp],
    "a3d_pf_predict": [_vp, _d, _d, _d, _d, _d, _d, _d, _d, _vp],
    "a3d_pf_update": [_vp, _vp, _vp, _vp, _vp, _sz, _vp, _sz, _d, _d, _d, _d],
    "a3d_pf_resample": [_vp, _f, _i, _vp],
    "a3d_pf_get_mean": [_vp, _vp],
    "a3d_pf_is_initialized": [_vp, _p(_i)],
    "a3d_pf_size": [_vp, _p(_sz)],
    "a3d_pf_get_particles": [_vp, _vp, _sz],
    "a3d_pf_set_particles": [_vp, _vp, _sz],
    "a3d_pf_predict_dev": [_vp, _vp, _vp, _vp],
    "a3d_pf_update_dev": [_vp, _vp, _vp, _vp, _vp, _sz, _vp, _sz, _d, _d, _d, _d],
    "a3d_pf_resample_dev": [_vp, _f, _i],
    "a3d_pf_update_noalloc": [_vp, _vp, _sz, _vp, _vp, _vp, _sz, _vp, _sz, _d, _d, _d, _d, _vp, _vp],
    "a3d_pf_synchronize": [_vp],
    "a3d_pf_capture_begin": [_vp],
    "a3d_pf_capture_end": [_vp, _p(_vp)],
    "a3d_graph_launch": [_vp],
    "a3d_graph_destroy": [_vp],
    "a3d_pf_timing": [_vp, _i, _p(_d), _p(_u64), _p(_u64)],
    "a3d_set_tile": [_i],
    "a3d_get_tile": [_p(_i)],
    "a3d_set_deal": [_i],
    "a3d_get_deal": [_p(_i)],
    "a3d_set_cloud_tile": [_i],
    "a3d_get_cloud_tile": [_p(_i)],
    "a3d_set_mapping": [_i],
    "a3d_get_mapping": [_p(_i)],
    "a3d_set_order": [_i],
    "a3d_get_order": [_p(_i)],
    "a3d_fleet_create": [_sz, _vp, _p(_vp)],
    "a3d_fleet_destroy": [_vp],
    "a3d_fleet_set_particles": [_vp, _vp],
    "a3d_fleet_get_particles": [_vp, _vp],
    "a3d_fleet_init": [_vp, _vp, _vp, _vp],
    "a3d_fleet_predict": [_vp, _vp, _vp, _vp],
    "a3d_fleet_update": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_fleet_resample": [_vp, _vp, _i],
    "a3d_fleet_get_means": [_vp, _vp],
    "a3d_fleet_stage_update": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_fleet_update_staged": [_vp, _vp, _i],
    "a3d_fleet_synchronize": [_vp],
    "a3d_fleet_timing": [_vp, _i, _p(_d), _p(_u64), _p(_u64)],
    "a3d_fleet_set_layout": [_vp, _vp, _vp],
    "a3d_fleet_predict_dev": [_vp, _vp, _vp, _vp],
    "a3d_fleet_update_dev": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_fleet_resample_dev": [_vp, _vp, _i],
    "a3d_fleet_capture_begin": [_vp],
    "a3d_fleet_capture_end": [_vp, _p(_vp)],
    "a3d_server_create": [_p(ServerConfig), _p(_vp)],
    "a3d_server_destroy": [_vp],
    "a3d_server_add_map": [_vp, _vp, _p(_i)],
    "a3d_server_open_session": [_vp, _i, _i, _vp, _vp, _u64, _p(_i)],
    "a3d_server_open_session_particles": [_vp, _i, _vp, _sz, _u64, _p(_i)],
    "a3d_server_close_session": [_vp, _i],
    "a3d_server_submit": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _sz, _d, _d, _d, _d, _i, _p(_u64)],
    "a3d_server_poll": [_vp, _p(Result), _sz, _p(_sz)],
    "a3d_server_drain": [_vp],
    "a3d_server_get_particles": [_vp, _i, _vp, _sz],
    "a3d_server_stats": [_vp, _p(_u64), _p(_u64), _p(_u64), _p(_d)],
    "a3d_session_draws": [_u64, _u64, _i, _vp, _vp],
    "a3d_node_create": [_p(NodeParams), _vp, _u64, _p(_vp)],
    "a3d_node_destroy": [_vp],
    "a3d_node_odometry": [_vp, _vp, _vp, _p(NodeOutput)],
    "a3d_node_range": [_vp, _f, _d, _d, _d],
    "a3d_node_point_cloud": [_vp, ctypes.c_int64, _vp, _sz, _sz, _p(_i)],
    "a3d_node_set_initial_pose": [_vp, _vp, _vp, _f, _f, _f, _f],
    "a3d_node_get_state": [_vp, _vp, _vp, _p(_d), _p(_d), _p(_i), _p(_i), _p(_u64)],
    "a3d_node_get_filter": [_vp, _p(_vp)],
}
_RESTYPES = {"a3d_last_error": ctypes.c_char_p, "a3d_version": ctypes.c_char_p}

_LIB = None


def load():
    """Loads the HIP library (once).  Raises LibraryMissing if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing("libamcl3d_gpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def check(rc):
    if rc != A3D_OK:
        raise Amcl3dError(rc, load().a3d_last_error().decode(errors="replace"))


def call(name, *args):
    """Calls a C-ABI function.  numpy arrays may be passed directly: they are
    turned into data pointers here, while this frame keeps them alive (a
    pointer taken from a temporary array would dangle once it is freed)."""
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            if not a.flags["C_CONTIGUOUS"]:
                raise ValueError("%s: array argument must be C-contiguous" % name)
            conv.append(a.ctypes.data)
        else:
            conv.append(a)
    check(getattr(load(), name)(*conv))
    del args


def ptr(a):
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def cloud_soa(cloud):
    """(m, 3) points -> three contiguous float32 arrays x, y, z."""
    c = f32(cloud).reshape(-1, 3)
    return tuple(np.ascontiguousarray(c[:, k]) for k in range(3))


def device_info():
    name = ctypes.create_string_buffer(256)
    cus = _i()
    mem = _sz()
    call("a3d_device_info", name, 256, ctypes.byref(cus), ctypes.byref(mem))
    return {"name": name.value.decode(), "cu_count": cus.value, "hbm_bytes": mem.value}


def synchronize():
    call("a3d_device_synchronize")


def set_tile(particles_per_wave):
    call("a3d_set_tile", int(particles_per_wave))


def get_tile():
    v = _i()
    call("a3d_get_tile", ctypes.byref(v))
    return v.value


def set_cloud_tile(points):
    call("a3d_set_cloud_tile", int(points))


def set_deal(run):
    call("a3d_set_deal", int(run))


def get_deal():
    v = _i()
    call("a3d_get_deal", ctypes.byref(v))
    return v.value


def get_cloud_tile():
    v = _i()
    call("a3d_get_cloud_tile", ctypes.byref(v))
    return v.value


MAP_PARTICLE_LANES = 0
MAP_POINT_LANES = 1


def set_mapping(mapping):
    call("a3d_set_mapping", int(mapping))


def get_mapping():
    v = _i()
    call("a3d_get_mapping", ctypes.byref(v))
    return v.value


ORDER_INDEX = 0
ORDER_SPATIAL = 1


def set_order(order):
    call("a3d_set_order", int(order))


def get_order():
    v = _i()
    call("a3d_get_order", ctypes.byref(v))
    return v.value


class DeviceBuffer:
    """Device memory owned by the library (inputs kept resident in HBM)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        call("a3d_buffer_alloc", self.nbytes, ctypes.byref(p))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 4))
        if a.nbytes:
            call("a3d_buffer_upload", b.ptr, ptr(a), a.nbytes)
        return b

    def download(self, dtype, count, offset_bytes=0):
        out = np.empty(count, dtype=dtype)
        call("a3d_buffer_download", ptr(out), self.ptr + offset_bytes, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            call("a3d_buffer_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _LIB is not None:
                _LIB.a3d_buffer_free(self.ptr)
        except Exception:
            pass


def voxel_grid(points, leaf):
    """pcl::VoxelGrid<PointXYZ> on the device (Node.cpp:130-141): (k, 3) float32
    centroids in voxel order.  points: (n, 3) or (n, 4) float32 (PointXYZ)."""
    p = f32(points)
    p = p.reshape(p.shape[0], -1) if p.size else p.reshape(0, 3)
    n = p.shape[0]
    out = np.empty((3, max(n, 1)), np.float32)
    k = ctypes.c_size_t()
    call("a3d_voxel_grid", p, n, p.shape[1] * 4, float(leaf), out[0], out[1], out[2], ctypes.byref(k))
    return np.ascontiguousarray(out[:, :k.value].T)
