"""GPU: the header-only C++ layer (include/amcl3d_gpu.hpp), driven as the
reference's node drives Grid3d and ParticleFilter, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from amcl3d_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32
SHIM = os.path.join(os.path.dirname(_lib.LIB_PATH), "shim_check")


@pytest.mark.parametrize("neg", [False, True], ids=["n", "minus-n"])
def test_cpp_layer_full_step(gpu, tmp_path, neg):
    """init -> predict -> update -> resample through the C++ layer; with
    minus-n, init is given -n particles, which the reference resizes to
    abs(n) (ParticleFilter.cpp:51), so the results are the same."""
    c = W.room_case(4, n=700, m=1500)
    bm, s = c["map"], c["step"]
    pts = bm.points().astype(F32)
    n = 700
    rng = np.random.default_rng(9)
    z_init = rng.standard_normal((n, 4)).astype(F32)
    pose = np.asarray(c["pose"], F32)
    devs = np.array([0.3, 0.3, 0.1, 0.2], F32)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([pts.shape[0], s.cloud.shape[0], s.ranges.shape[0], n], np.uint32).tofile(f)
        np.concatenate([bm.min, bm.max, [bm.res, c["sensor_dev"]], s.mods, s.deltas, [s.alpha, s.sigma]]
                       ).astype(np.float64).tofile(f)
        np.concatenate([[s.u01], pose, devs]).astype(F32).tofile(f)
        for a in (pts, s.cloud, s.ranges, z_init, s.z4):
            np.ascontiguousarray(a, F32).tofile(f)
    out = tmp_path / "out.bin"
    print("shim_check: C++ layer, room seed 4, n=700 m=1500, negative count", neg, flush=True)
    r = subprocess.run([SHIM, str(inp), str(out)] + (["neg"] if neg else []), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, F32)
    k = 0

    def take(cnt):
        nonlocal k
        v = got[k:k + cnt]
        k += cnt
        return v

    p_init, m_init = take(7 * n).reshape(n, 7), take(4)
    p_upd, m_upd = take(7 * n).reshape(n, 7), take(4)
    p_res = take(7 * n).reshape(n, 7)
    w_mean, into = take(1)[0], take(1)[0]
    assert k == got.size

    _, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, c["sensor_dev"], exp_mode=1)
    og = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    op, om = oracle.pf_init(n, pose, devs, z_init)
    assert np.array_equal(p_init.view(np.int32), op.view(np.int32))
    assert np.array_equal(m_init.view(np.int32), om.view(np.int32))
    op = oracle.pf_predict(op, s.mods, s.deltas, s.z4)
    op, om = oracle.pf_update(op, og, s.cloud, s.ranges, s.alpha, s.sigma)
    assert np.array_equal(p_upd.view(np.int32), op.view(np.int32))
    assert np.array_equal(m_upd.view(np.int32), om.view(np.int32))
    oq, _ = oracle.pf_resample(op, s.u01, 0)
    assert np.array_equal(p_res.view(np.int32), oq.view(np.int32))
    ow = oracle.cloud_weight(og, s.cloud, om[0], om[1], om[2], 0.0, 0.0, om[3])
    assert np.float32(w_mean) == np.float32(ow)
    assert bool(into) == bool(oracle.is_into_map(og, om[0], om[1], om[2]))


def test_cpp_layer_opens_a_map_file(gpu, tmp_path):
    """Grid3d::open(map_path, sensor_dev) through the C++ layer, as the node
    calls it (Node.cpp:41): the .bt read, computeGrid on the device, the
    cache written; the cells against the oracle chain.  A missing map makes
    open return false (Grid3dTest.cpp:29-42)."""
    from amcl3d_amd import octree as ot
    from oracle import octree as oo

    rng = np.random.default_rng(12)
    pts = np.concatenate([rng.uniform([0, 0, 0], [6, 4, 0.1], (3000, 3)),
                          rng.uniform([0, 0, 0], [0.1, 4, 2.5], (1500, 3)),
                          rng.uniform([2, 1, 0.5], [3, 2, 1.5], (800, 3))])
    path = str(tmp_path / "site.bt")
    ot.write_points(path, pts, 0.05)
    out = tmp_path / "grid.bin"
    print("shim_check map", path, flush=True)
    r = subprocess.run([SHIM, "map", path, "0.05", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(out, "rb").read()
    size = np.frombuffer(raw[:12], np.uint32)
    mn, mx = np.frombuffer(raw[12:36], np.float64), np.frombuffer(raw[36:60], np.float64)
    prob = np.frombuffer(raw[60:], np.float32)
    opts, omn, omx, ores = oo.compute_point_cloud(oo.open_octree(path))
    assert list(mn) == omn and list(mx) == omx
    _, oprob = oracle.compute_grid(np.asarray(opts, F32), omn, omx, ores, 0.05, exp_mode=1)
    assert tuple(size) == oprob.shape[::-1]
    assert np.array_equal(prob.view(np.int32), oprob.reshape(-1).view(np.int32))
    assert os.path.exists(str(tmp_path / "site.grid"))
    r = subprocess.run([SHIM, "map", str(tmp_path / "missing.bt"), "0.05", str(out)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 3 and "Cannot find file" in r.stdout
