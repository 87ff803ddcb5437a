"""The C-ABI library on the CPU: it loads and exports every declared symbol.

No compute call is made here; the GPU tests call through the same ABI.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from amcl3d_amd import _lib


def _declared():
    text = open(_lib.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(a3d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("a3d_grid_create_from_points", "a3d_grid_cloud_weight", "a3d_grid_is_into_map", "a3d_pf_init",
                 "a3d_pf_predict", "a3d_pf_update", "a3d_pf_resample", "a3d_pf_get_mean", "a3d_fleet_update",
                 "a3d_server_submit"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes binding covers the whole header
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_no_gpu_calls_fail_or_validate():
    L = _lib.load()
    assert b"gfx950" in L.a3d_version()
    inside = ctypes.c_int(7)
    # no map: isIntoMap is false (Grid3d.cpp:203-204)
    assert L.a3d_grid_is_into_map(None, 1.0, 1.0, 1.0, ctypes.byref(inside)) == _lib.A3D_OK
    assert inside.value == 0
    assert L.a3d_set_tile(24) == _lib.A3D_ERR_ARG
    assert b"16, 32 or 64" in L.a3d_last_error()
    assert L.a3d_set_tile(32) == _lib.A3D_OK
    # no map: computeCloudWeight is 0 (Grid3d.cpp:136-137)
    out = np.full(3, 9.0, np.float32)
    t = np.zeros(3, np.float32)
    assert L.a3d_grid_cloud_weight(None, None, None, None, 0, _lib.ptr(t), _lib.ptr(t), _lib.ptr(t), _lib.ptr(t), 3,
                                   0.0, 0.0, _lib.ptr(out)) == _lib.A3D_OK
    assert np.all(out == 0)
    assert L.a3d_pf_init(None, 10, 0, 0, 0, 0, 0.1, 0.1, 0.1, 0.1, None) == _lib.A3D_ERR_ARG
    assert L.a3d_grid_create_from_points(None, 5, None, None, 0.1, 0.1, 0, None) == _lib.A3D_ERR_ARG
    assert L.a3d_fleet_create(0, None, None) == _lib.A3D_ERR_ARG
    assert L.a3d_session_draws(1, 0, 0, None, None) == _lib.A3D_ERR_ARG


def test_bad_grid_arguments_are_rejected_before_the_gpu():
    g = ctypes.c_void_p()
    mn = np.zeros(3)
    mx = np.array([1.0, 1.0, 0.0])  # empty z extent
    pts = np.zeros((1, 3), np.float32)
    rc = _lib.load().a3d_grid_create_from_points(_lib.ptr(pts), 1, _lib.ptr(mn), _lib.ptr(mx), 0.1, 0.05, 0,
                                                 ctypes.byref(g))
    assert rc == _lib.A3D_ERR_ARG
    mx = np.ones(3)
    rc = _lib.load().a3d_grid_create_from_points(_lib.ptr(pts), 1, _lib.ptr(mn), _lib.ptr(mx), -0.1, 0.05, 0,
                                                 ctypes.byref(g))
    assert rc == _lib.A3D_ERR_ARG


def test_product_path_has_no_cpu_fallback():
    # the package never imports the oracle
    import amcl3d_amd
    import pathlib

    src = "".join(p.read_text() for p in pathlib.Path(amcl3d_amd.__file__).parent.glob("*.py"))
    assert "oracle" not in src.replace("oracle is", "")
    with pytest.raises(_lib.LibraryMissing):
        saved = _lib.LIB_PATH, _lib._LIB
        try:
            _lib.LIB_PATH, _lib._LIB = "/nonexistent/libamcl3d_gpu.so", None
            _lib.load()
        finally:
            _lib.LIB_PATH, _lib._LIB = saved


def test_cpp_layer_soft_failures_without_a_map():
    """Grid3d without an open map: computeCloudWeight 0 and isIntoMap false
    (Grid3d.cpp:136-137, 203-204), through the C++ class layer."""
    import subprocess

    shim = os.path.join(os.path.dirname(_lib.LIB_PATH), "shim_check")
    r = subprocess.run([shim, "unopened"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert '"weight": 0' in r.stdout and '"into_map": 0' in r.stdout


def test_kernel_shapes_round_trip_and_refuse_bad_values():
    """Every shape the autotuner tries (mapping, particles per wavefront,
    order, cloud tile, XCD dealing) is set and read back as given, without a
    GPU; out-of-range values are refused with a message and change nothing."""
    from amcl3d_amd import autotune

    L = _lib.load()
    before = autotune.get_shape()
    try:
        names = set()
        for sh in autotune.FLEET_CANDIDATES:
            autotune.set_shape(sh)
            assert autotune.get_shape() == sh
            names.add(autotune.shape_name(sh))
        assert len(names) == len(autotune.FLEET_CANDIDATES)  # names tell shapes apart
        assert autotune.shape_name((0, 16, 1, 256, 4)) == "map0-ppw16-spatial-tile256-deal4"
        autotune.set_shape((0, 16, 1, 128, 0))
        for call, bad in ((L.a3d_set_cloud_tile, 48), (L.a3d_set_cloud_tile, 512), (L.a3d_set_deal, -1),
                          (L.a3d_set_deal, 5000), (L.a3d_set_tile, 8), (L.a3d_set_mapping, 2)):
            assert call(bad) == _lib.A3D_ERR_ARG
            assert L.a3d_last_error()
        assert autotune.get_shape() == (0, 16, 1, 128, 0)
    finally:
        autotune.set_shape(before)
