"""OcTree map files (.bt / .ot): openOcTree + computePointCloud
(amcl3d/src/PointCloudTools.cpp:26-82), the library's host reader against the
oracle (oracle/octree.py) and the oracle against the reference's own data.

These run on the CPU: the reader is host code in the library and makes no
GPU call.  Grid3d::open on top of it (the device grid and the .grid cache) is
in tests/test_gpu_octree.py.
"""
import math
import os
import struct

import numpy as np
import pytest

from amcl3d_amd import octree as ot
from oracle import octree as oo

RES = 0.05
# PointCloudToolsTest.cpp:42-53,166-173 (mapfile_complete.bt)
BT_MIN = (-17.35, -9.5, -1.4)
BT_MAX = (8.75, 9.7, 6.25)


def _ulps(a, b):
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def _double_eq(a, b):
    """gtest's ASSERT_DOUBLE_EQ: within 4 units in the last place."""
    return _ulps(float(a), float(b)) <= 4


def _lattice_leaves(xyz, shift):
    """Depth and index of the OcTree leaf centred at each point of a cloud
    given as float(centre) - float(shift): the centre's key-space coordinate
    2*(c/res + 2^15) is (2i+1) * 2^(16-d), exact on the lattice."""
    c = xyz.astype(np.float64) + np.asarray(shift, np.float64)
    u2f = 2.0 * (c / RES + 32768.0)
    u2 = np.rint(u2f).astype(np.int64)
    assert np.abs(u2f - u2).max() < 1e-2, "points off the OcTree lattice"
    k = np.zeros_like(u2)
    a = u2.copy()
    while True:
        even = (a % 2 == 0) & (a > 0)
        if not even.any():
            break
        a[even] //= 2
        k[even] += 1
    assert np.all(k[:, 0] == k[:, 1]) and np.all(k[:, 0] == k[:, 2]), "axes disagree on the leaf depth"
    depth = 16 - k[:, 0]
    idx = (a - 1) // 2  # a = 2i + 1
    return depth, idx


@pytest.fixture(scope="module")
def map_bt(kat, tmp_path_factory):
    """The test map of the reference (mappointcloud_msg.bin's occupied
    centres, stored as float(centre) - float(min)) as a .bt of its occupied
    leaves: 179,415 depth-16 leaves and 136 pruned depth-15 ones."""
    shift = np.float32(BT_MIN).astype(np.float64)
    depth, idx = _lattice_leaves(kat["map_cloud"], shift)
    assert len(depth) == 179551
    assert set(np.unique(depth)) == {15, 16}
    path = str(tmp_path_factory.mktemp("octree") / "mapfile_complete.bt")
    ot.write_octree(path, depth, idx, np.ones(len(depth), bool), RES)
    return path


def test_oracle_reproduces_the_reference_map_cloud(kat, map_bt):
    """KAT: leaf order, centres and count of the reference's occupied cloud,
    and the bounds and resolution its tests assert for mapfile_complete.bt."""
    pts, mn, mx, res = oo.compute_point_cloud(oo.open_octree(map_bt))
    assert len(pts) == 179551  # PointCloudToolsTest.cpp:166
    got = (np.asarray(pts, np.float32) - np.float32(BT_MIN)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), kat["map_cloud"].view(np.uint32))
    assert all(_double_eq(a, b) for a, b in zip(mn, BT_MIN))
    assert all(_double_eq(a, b) for a, b in zip(mx, BT_MAX))
    assert res == 0.05


def test_library_reads_the_reference_map_like_the_oracle(kat, map_bt):
    t = ot.openOcTree(map_bt)
    info = ot.computePointCloud(t)
    assert info.cloud.shape == (179551, 3)
    got = (info.cloud - np.float32(BT_MIN)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), kat["map_cloud"].view(np.uint32))
    pts, mn, mx, res = oo.compute_point_cloud(oo.open_octree(map_bt))
    assert np.array_equal(info.cloud, np.asarray(pts, np.float32))
    assert list(info.min) == mn and list(info.max) == mx and info.octo_resol == res
    assert t.size() == oo.open_octree(map_bt).size()
    for a, b in zip(t.getMetricMin() + t.getMetricMax(), BT_MIN + BT_MAX):
        assert _double_eq(a, b)


@pytest.mark.parametrize("depth", list(range(0, 17)))
def test_single_leaf_centre_and_bounds_closed_form(tmp_path, depth):
    """One occupied leaf at depth d, index i: centre (i - 2^(d-1) + 0.5) *
    res * 2^(16-d), cube [centre - side/2, centre + side/2]; the root is
    centred on the origin."""
    rng = np.random.default_rng(depth)
    i = rng.integers(0, 1 << depth, 3) if depth else np.zeros(3, np.int64)
    side = RES * 2.0 ** (16 - depth)
    for ext in (".ot", ".bt") if depth else (".ot",):
        path = str(tmp_path / ("one" + ext))
        ot.write_octree(path, [depth], [i], [True], RES)
        t = ot.openOcTree(path)
        assert t.size() == depth + 1
        want = [((int(v) - 2 ** (depth - 1) + 0.5) * side) if depth else 0.0 for v in i]
        got = t.points()
        assert got.shape == (1, 3)
        assert np.array_equal(got[0], np.float32(want))
        for a in range(3):
            assert math.isclose(t.min[a], want[a] - side / 2, rel_tol=0, abs_tol=1e-9 * side)
            assert math.isclose(t.max[a], want[a] + side / 2, rel_tol=0, abs_tol=1e-9 * side)
        p, mn, mx, _ = oo.compute_point_cloud(oo.open_octree(path)) if depth else (None, None, None, None)
        if depth:
            assert np.array_equal(np.asarray(p, np.float32), got)
            assert list(t.min) == mn and list(t.max) == mx


def _random_tree(rng, nmax=3000, max_depth=16):
    """Random non-overlapping leaves: a random subdivision from the root with
    pruned (shallow) leaves, free and occupied leaves and absent children."""
    depth, idx, occ = [], [], []
    stack = [(0, (0, 0, 0))]
    while stack and len(depth) < nmax:
        d, (i, j, k) = stack.pop()
        for c in range(8):
            r = rng.random()
            if r < 0.35:
                continue  # absent
            ci = (2 * i + (c & 1), 2 * j + ((c >> 1) & 1), 2 * k + ((c >> 2) & 1))
            if d + 1 < max_depth and (r < 0.55 + 0.04 * d or d + 1 < 3):
                stack.append((d + 1, ci))
            else:
                depth.append(d + 1)
                idx.append(ci)
                occ.append(rng.random() < 0.6)
    return np.array(depth), np.array(idx), np.array(occ)


def _compare(path):
    t = ot.openOcTree(path)
    ref = oo.open_octree(path)
    assert t.size() == ref.size()
    pts, mn, mx, res = oo.compute_point_cloud(ref)
    assert np.array_equal(t.points(), np.asarray(pts, np.float32).reshape(-1, 3))
    assert list(t.min) == mn and list(t.max) == mx and t.resolution == res
    return t


@pytest.mark.parametrize("seed", range(6))
def test_random_trees_bt_and_ot_agree_with_the_oracle(tmp_path, seed):
    rng = np.random.default_rng(100 + seed)
    depth, idx, occ = _random_tree(rng)
    res = [0.05, 0.1, 0.2, 0.025, 1.0 / 3.0, 0.07][seed]
    for ext in (".bt", ".ot"):
        path = str(tmp_path / ("t" + ext))
        n = ot.write_octree(path, depth, idx, occ, res)
        t = _compare(path)
        assert t.size() == n
        assert t.occupied == int(occ.sum())


def test_ot_log_odds_threshold(tmp_path):
    """.ot leaves are occupied when log-odds >= 0 (NaN is not, -0.0 is)."""
    vals = np.array([0.0, -0.0, 1e-45, -1e-45, np.nan, np.inf, -np.inf, 3.5, -2.0, 1e-30, -1e-30], np.float32)
    n = len(vals)
    idx = np.stack([np.arange(n), np.zeros(n, np.int64), np.zeros(n, np.int64)], 1)
    path = str(tmp_path / "v.ot")
    ot.write_octree(path, np.full(n, 16), idx, np.zeros(n, bool), RES, values=vals)
    t = _compare(path)
    occ = ~np.isnan(vals) & (vals >= 0)
    assert t.occupied == int(occ.sum())


def test_bt_and_ot_of_the_same_tree_match(tmp_path):
    rng = np.random.default_rng(7)
    depth, idx, occ = _random_tree(rng, nmax=5000)
    a = str(tmp_path / "m.bt")
    b = str(tmp_path / "m.ot")
    ot.write_octree(a, depth, idx, occ, RES)
    ot.write_octree(b, depth, idx, occ, RES)
    ta, tb = ot.openOcTree(a), ot.openOcTree(b)
    assert ta.size() == tb.size()
    assert np.array_equal(ta.points(), tb.points())
    assert np.array_equal(ta.min, tb.min) and np.array_equal(ta.max, tb.max)


def test_write_points_roundtrip(tmp_path):
    """write_points stores one depth-16 leaf per distinct cell; reading it
    back gives the cell centres in leaf order."""
    rng = np.random.default_rng(3)
    p = rng.uniform(-5, 5, (20000, 3))
    path = str(tmp_path / "p.bt")
    ot.write_points(path, p, 0.1)
    t = _compare(path)
    key = np.unique(np.floor(p / 0.1).astype(np.int64), axis=0)
    assert t.occupied == len(key)
    c = t.points().astype(np.float64)
    assert np.array_equal(np.unique(np.floor(c / 0.1).astype(np.int64), axis=0), key)


# ---------------------------------------------------------------- byte level
def _bt(body, size, res="0.05", extra=b""):
    return (b"# Octomap OcTree binary file\n" + extra + b"id OcTree\nsize %d\nres %s\ndata\n" % (size, res.encode())
            + body)


def _write(tmp_path, name, data):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(data)
    return p


def _chain_bt(depth_of_leaf, kind=2):
    """A chain of nodes down to one child 0 at the given depth: bytes of
    every node with children (child 0 marked 11 above the leaf)."""
    body = b""
    for _ in range(depth_of_leaf - 1):
        body += bytes([3, 0])
    body += bytes([kind, 0])
    return body


def test_bt_childless_inner_node_is_a_free_leaf(tmp_path):
    # root: child 0 marked 11 with no children of its own, child 1 occupied leaf
    body = bytes([0b1011, 0]) + bytes([0, 0])
    p = _write(tmp_path, "c.bt", _bt(body, 3))
    t = _compare(p)
    assert t.size() == 3 and t.occupied == 1


def test_bt_header_variants(tmp_path):
    body = _chain_bt(16)
    base = _compare(_write(tmp_path, "a.bt", _bt(body, 17)))
    # comments, unknown keys, CR LF line ends, text after "data", the deprecated id
    hdr = (b"# Octomap OcTree binary file\r\n# a comment\r\nid 1\r\nfoo bar baz\r\nsize 17\r\nres 0.05\r\n"
           b"data trailing words\r\n")
    t = _compare(_write(tmp_path, "b.bt", hdr + body))
    assert np.array_equal(t.points(), base.points())
    # any tree id: the binary format is the same for every occupancy tree
    t = _compare(_write(tmp_path, "c.bt", _bt(body, 17).replace(b"id OcTree", b"id ColorOcTree")))
    assert t.occupied == 1
    # the legacy header: int32 3, f64 res, u32 size
    t = _compare(_write(tmp_path, "d.bt", struct.pack("<idI", 3, 0.05, 17) + body))
    assert np.array_equal(t.points(), base.points())


ERRORS_BT = {
    "empty": b"",
    "size_mismatch": _bt(_chain_bt(16), 16),
    "truncated": _bt(_chain_bt(16)[:-1], 17),
    "below_depth_16": _bt(_chain_bt(17), 18),
    "res_zero": _bt(_chain_bt(3), 4, res="0"),
    "res_negative": _bt(_chain_bt(3), 4, res="-0.05"),
    "no_data_line": b"# Octomap OcTree binary file\nid OcTree\nsize 4\nres 0.05\n",
    "no_id": b"# Octomap OcTree binary file\nsize 4\nres 0.05\ndata\n" + _chain_bt(3),
    "legacy_wrong_type": struct.pack("<idI", 2, 0.05, 4) + _chain_bt(3),
    "legacy_bad_res": struct.pack("<idI", 3, -1.0, 4) + _chain_bt(3),
    "ot_header_in_bt": b"# Octomap OcTree file\nid OcTree\nsize 4\nres 0.05\ndata\n" + _chain_bt(3),
}


@pytest.mark.parametrize("case", sorted(ERRORS_BT))
def test_bt_errors(tmp_path, case):
    p = _write(tmp_path, "e.bt", ERRORS_BT[case])
    with pytest.raises(RuntimeError, match="OcTree cannot be read"):
        ot.openOcTree(p)
    with pytest.raises(RuntimeError, match="OcTree cannot be read"):
        oo.open_octree(p)


def _ot(nodes, size=None, ident=b"OcTree", res=b"0.05"):
    body = b"".join(struct.pack("<fB", v, m) for v, m in nodes)
    size = len(nodes) if size is None else size
    return b"# Octomap OcTree file\nid " + ident + b"\nsize %d\nres " % size + res + b"\ndata\n" + body


ERRORS_OT = {
    "empty": b"",
    "color_tree": _ot([(1.0, 1), (1.0, 0)], ident=b"ColorOcTree"),
    "truncated": _ot([(1.0, 1)]),
    "below_depth_16": _ot([(1.0, 1)] * 17 + [(1.0, 0)]),
    "res_zero": _ot([(1.0, 1), (1.0, 0)], res=b"0"),
    "bt_header_in_ot": _bt(_chain_bt(3), 4),
}


@pytest.mark.parametrize("case", sorted(ERRORS_OT))
def test_ot_errors(tmp_path, case):
    p = _write(tmp_path, "e.ot", ERRORS_OT[case])
    with pytest.raises(RuntimeError, match="OcTree cannot be created from file"):
        ot.openOcTree(p)
    with pytest.raises(RuntimeError, match="OcTree cannot be created from file"):
        oo.open_octree(p)


def test_ot_size_is_not_checked(tmp_path):
    """AbstractOcTree::read reads the tree whatever size the header states."""
    t = _compare(_write(tmp_path, "s.ot", _ot([(1.0, 0b10000001), (1.0, 0), (-1.0, 0)], size=99)))
    assert t.size() == 3 and t.occupied == 1


def test_open_errors_of_the_reference_tests(tmp_path):
    """PointCloudToolsTest.cpp:59-103 and Grid3dTest.cpp:29-42."""
    for name in ("unknown_file.bt", "unknown_file.ot", "unknown_file.unk", ".bt"):
        with pytest.raises(RuntimeError, match="Cannot find file"):
            ot.openOcTree(str(tmp_path / name) if name != ".bt" else name)
    p = _write(tmp_path, "mapfile_unknown_extension.unk", b"")
    with pytest.raises(RuntimeError, match="OcTree cannot be created from file"):
        ot.openOcTree(p)
    d = tmp_path / "dir.bt"
    d.mkdir()
    with pytest.raises(RuntimeError, match="OcTree cannot be read"):
        ot.openOcTree(str(d))  # exists, but is not a readable file
    with pytest.raises(RuntimeError, match="OcTree cannot be read"):
        oo.open_octree(str(d))


def test_empty_trees(tmp_path):
    """size 0, and a root without children: openOcTree succeeds,
    computePointCloud raises "OcTree is empty" (PointCloudTools.cpp:56-58)."""
    for p in (_write(tmp_path, "z.bt", _bt(b"", 0)), _write(tmp_path, "r.bt", _bt(bytes([0, 0]), 1)),
              _write(tmp_path, "z.ot", _ot([], size=0)), _write(tmp_path, "r.ot", _ot([(1.0, 0)]))):
        t = ot.openOcTree(p)
        assert t.size() <= 1
        with pytest.raises(RuntimeError, match="OcTree is empty"):
            ot.computePointCloud(t)
        with pytest.raises(RuntimeError, match="OcTree is empty"):
            oo.compute_point_cloud(oo.open_octree(p))
    with pytest.raises(RuntimeError, match="OcTree is NULL"):
        ot.computePointCloud(None)


def test_planted_mistakes_are_caught(kat, map_bt, monkeypatch):
    """The fixture check fails for a plausible mistake in the centre formula
    or the child order (planted in the oracle)."""
    pts, _, _, _ = oo.compute_point_cloud(oo.open_octree(map_bt))
    base = np.asarray(pts, np.float32)
    # centre at the cube's corner instead of its middle
    monkeypatch.setattr(oo.Tree, "key_to_coord",
                        lambda self, key, depth: float(key - oo.TREE_MAX_VAL) * self.res)
    bad, _, _, _ = oo.compute_point_cloud(oo.open_octree(map_bt))
    assert not np.array_equal(np.asarray(bad, np.float32), base)
    monkeypatch.undo()
    # children visited 7..0
    orig = oo.Tree.leaves

    def reversed_leaves(self):
        out = list(orig(self))
        return iter(out[::-1])

    monkeypatch.setattr(oo.Tree, "leaves", reversed_leaves)
    bad, _, _, _ = oo.compute_point_cloud(oo.open_octree(map_bt))
    got = (np.asarray(bad, np.float32) - np.float32(BT_MIN)).astype(np.float32)
    assert not np.array_equal(got.view(np.uint32), kat["map_cloud"].view(np.uint32))


@pytest.mark.parametrize("seed", range(4))
def test_one_pass_bt_writer_matches_the_general_writer(tmp_path, seed):
    """write_points' one-pass .bt writer gives the bytes of write_octree for
    the same depth-16 leaves."""
    rng = np.random.default_rng(seed)
    n = [1, 7, 3000, 40000][seed]
    p = rng.uniform(-40, 40, (n, 3)) * rng.uniform(0.05, 1.0)
    a = str(tmp_path / "a.bt")
    b = str(tmp_path / "b.bt")
    ot.write_points(a, p, 0.05)
    key = np.unique(np.floor(p / 0.05).astype(np.int64) + (1 << 15), axis=0)
    ot.write_octree(b, np.full(len(key), 16), key, np.ones(len(key), bool), 0.05)
    assert open(a, "rb").read() == open(b, "rb").read()
