"""Grid3d::open(map_path, sensor_dev) on the GPU (amcl3d/src/Grid3d.cpp:25-78):
the OcTree file read by the library, then the .grid cache or computeGrid on
the device, against the oracle chain (oracle/octree.py, then the oracle's
computeGrid) on the same file; the reference's own known answer (KAT-1,
Grid3dTest.cpp:118-169) through the same route.
"""
import os
import struct
import sys

import numpy as np
import pytest

import oracle
from amcl3d_amd import octree as ot
from amcl3d_amd import workloads as W
from oracle import octree as oo

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_grid(path, sd):
    pts, mn, mx, res = oo.compute_point_cloud(oo.open_octree(path))
    dist, prob = oracle.compute_grid(np.asarray(pts, np.float32), mn, mx, res, sd, exp_mode=1)
    return np.asarray(mn), np.asarray(mx), res, dist.reshape(-1), prob.reshape(-1)


@pytest.fixture(scope="module")
def room_bt(tmp_path_factory):
    bm = W.room_map(1)
    d = tmp_path_factory.mktemp("maps")
    path = str(d / "room.bt")
    ot.write_points(path, bm.points(), bm.res)
    return path


def test_open_bt_builds_the_oracle_grid_and_writes_the_cache(gpu, room_bt):
    sd = 0.1
    cache = room_bt[:-3] + ".grid"
    if os.path.exists(cache):
        os.remove(cache)
    _say("open", room_bt)
    g = gpu.Grid3d()
    assert g.open(room_bt, sd, keep_dist=True)
    mn, mx, res, odist, oprob = _oracle_grid(room_bt, sd)
    assert np.array_equal(g.min, mn) and np.array_equal(g.max, mx) and g.res == res
    prob, dist = g.read_cells(dist=True)
    assert np.array_equal(dist, odist)
    assert np.array_equal(prob.view(np.int32), oprob.view(np.int32))
    # Grid3d::saveGrid's file: u32 sizes, f64 sensor_dev, {dist, prob} per cell
    with open(cache, "rb") as f:
        head = f.read(20)
        cells = np.frombuffer(f.read(), np.float32).reshape(-1, 2)
    assert struct.unpack("<3Id", head) == (*(int(v) for v in g.size), sd)
    assert np.array_equal(cells[:, 0], odist) and np.array_equal(cells[:, 1].view(np.int32), oprob.view(np.int32))


def test_open_uses_the_cache_then_rebuilds_for_another_sensor_dev(gpu, room_bt):
    sd = 0.1
    cache = room_bt[:-3] + ".grid"
    g = gpu.Grid3d()
    assert g.open(room_bt, sd)  # writes the cache if the test above did not
    n = g.cells
    # a cache with the same header and other cells is what the next open returns
    with open(cache, "r+b") as f:
        f.seek(20)
        f.write(np.tile(np.float32([1.0, 0.5]), n).tobytes())
    g2 = gpu.Grid3d()
    assert g2.open(room_bt, sd)
    assert np.all(g2.read_cells() == np.float32(0.5))
    # another sensor_dev: loadGrid refuses the cache (Grid3d.cpp:256-260), the
    # grid is computed and the cache rewritten
    sd2 = 0.2
    g3 = gpu.Grid3d()
    assert g3.open(room_bt, sd2)
    _, _, _, _, oprob2 = _oracle_grid(room_bt, sd2)
    assert np.array_equal(g3.read_cells().view(np.int32), oprob2.view(np.int32))
    with open(cache, "rb") as f:
        assert struct.unpack("<3Id", f.read(20))[3] == sd2
    assert g3.sensor_dev == sd2


def test_open_ot_matches_bt(gpu, tmp_path):
    rng = np.random.default_rng(5)
    pts = rng.uniform([-3, -2, 0], [3, 2, 2.5], (4000, 3))
    a = str(tmp_path / "m.bt")
    b = str(tmp_path / "m.ot")
    ot.write_points(a, pts, 0.05)
    ot.write_points(b, pts, 0.05)
    ga, gb = gpu.Grid3d(), gpu.Grid3d()
    assert ga.open(a, 0.05) and gb.open(b, 0.05)
    assert np.array_equal(ga.size, gb.size) and np.array_equal(ga.min, gb.min)
    pa, pb = ga.read_cells(), gb.read_cells()
    assert np.array_equal(pa.view(np.int32), pb.view(np.int32))
    _, _, _, _, oprob = _oracle_grid(b, 0.05)
    assert np.array_equal(pb.view(np.int32), oprob.view(np.int32))


def test_kat_through_the_octree_route(gpu, kat, tmp_path):
    """Grid3dTest.computeCloudWeightTest: open the map file, then
    computeCloudWeight at the known pose = 3.8109049797058105 +- 1e-4.  The
    map file is rebuilt from the reference's occupied cloud (leaf order and
    depths recovered from it, as the CPU tests check)."""
    from test_octree import BT_MIN, _lattice_leaves

    depth, idx = _lattice_leaves(kat["map_cloud"], np.float32(BT_MIN).astype(np.float64))
    path = str(tmp_path / "mapfile_complete.bt")
    ot.write_octree(path, depth, idx, np.ones(len(depth), bool), kat["map_res"])
    g = gpu.Grid3d()
    p = kat["kat_pose"]
    assert g.computeCloudWeight(kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], p["roll"], p["pitch"],
                                p["yaw"]) == 0  # not open
    _say("kat: open", path)
    assert g.open(path, kat["sensor_dev"])
    # octomap's bounds arithmetic puts max y one ulp above 9.7, so the grid
    # has 385 rows here where the reference's nav slice shows 384
    _, mn, mx, res = oo.compute_point_cloud(oo.open_octree(path))
    assert list(g.size) == [int(np.ceil((mx[a] - mn[a]) / res)) for a in range(3)] == [522, 385, 153]
    w = g.computeCloudWeight(kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], p["roll"], p["pitch"], p["yaw"])
    prob = g.read_cells()
    og = oracle.GridMeta(g.min, g.max, g.res, prob)
    ow = oracle.cloud_weight(og, kat["sensor_cloud"], p["tx"], p["ty"], p["tz"], p["roll"], p["pitch"], p["yaw"])
    _say("KAT-1 through the octree route:", w, "reference 3.8109049797058105")
    assert w == ow
    assert abs(w - kat["kat_weight"]) <= kat["kat_tolerance"]


def test_open_failures_return_false(gpu, tmp_path):
    """Grid3dTest.cpp:29-42: missing and empty maps, a bare ".bt"."""
    g = gpu.Grid3d()
    assert not g.open(str(tmp_path / "mapfile_null.bt"), 0.05)
    assert "Cannot find file" in g.last_error
    assert not g.open(".bt", 0.05)
    p = str(tmp_path / "mapfile_null.ot")
    open(p, "wb").close()
    assert not g.open(p, 0.05)
    assert "OcTree cannot be created from file" in g.last_error
    p = str(tmp_path / "root_only.bt")
    with open(p, "wb") as f:
        f.write(b"# Octomap OcTree binary file\nid OcTree\nsize 1\nres 0.05\ndata\n\0\0")
    assert not g.open(p, 0.05)
    assert "OcTree is empty" in g.last_error
    assert not g.is_open


def test_cache_name_follows_the_reference(gpu, tmp_path):
    """The cache is the map path up to the FIRST ".bt" in it, plus ".grid"
    (Grid3d.cpp:48-52)."""
    d = tmp_path / "site.bt.d"
    d.mkdir()
    path = str(d / "map.bt")
    rng = np.random.default_rng(2)
    ot.write_points(path, rng.uniform(0, 2, (500, 3)), 0.1)
    g = gpu.Grid3d()
    assert g.open(path, 0.1)
    assert os.path.exists(str(tmp_path / "site.grid"))
    assert not os.path.exists(str(d / "map.grid"))
