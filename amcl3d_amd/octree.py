"""OcTree map files: the reader the reference's map loading uses, and a writer.

``openOcTree`` / ``computePointCloud`` mirror amcl3d/src/PointCloudTools.cpp:
26-82 over the C ABI (the parse is host C++ in the library): they read an
octomap OcTree file, binary ``.bt`` or full ``.ot``, into the centres of its
occupied leaves (floats, in octomap's leaf-iterator order) and the metric
bounds and resolution.  Failures raise ``RuntimeError`` (``Amcl3dError``) with
the reference's messages.

``write_octree`` / ``write_points`` write such files from leaves given by
depth and index (numpy, no per-node Python loop), so point maps can be
stored in the format the reference loads.  Layout of both formats, as
octomap 1.x defines them:

* ``.bt``: text header up to a ``data`` line, then every node that has
  children, parents first and children in order, as two bytes; bit pair
  (2c, 2c+1) of byte c // 4 describes child c: 10 free leaf, 01 occupied leaf,
  11 node with children, 00 absent.
* ``.ot``: the same header, then every node parents first: float32 log-odds
  and a byte with bit c set when child c exists.
* Child c takes the upper half of x for bit 0 of c, of y for bit 1, of z for
  bit 2; the node at depth d with index i along an axis is centred at
  ``(i - 2**(d-1) + 0.5) * res * 2**(16-d)``.
"""
import ctypes

import numpy as np

from . import _lib

TREE_DEPTH = 16
_OCC_LOGODDS = np.float32(3.5)   # octomap's default clamping maximum
_FREE_LOGODDS = np.float32(-2.0)  # and minimum


class OcTree:
    """An OcTree file read by the library (openOcTree's result)."""

    def __init__(self, h):
        self._h = h
        mn = np.zeros(3)
        mx = np.zeros(3)
        res = ctypes.c_double()
        nodes = ctypes.c_uint64()
        occ = ctypes.c_uint64()
        _lib.call("a3d_octree_get_info", h, _lib.ptr(mn), _lib.ptr(mx), ctypes.byref(res), ctypes.byref(nodes),
                  ctypes.byref(occ))
        self.min, self.max, self.resolution = mn, mx, res.value
        self.nodes, self.occupied = nodes.value, occ.value

    def size(self):
        """OcTree::size(): the number of nodes."""
        return self.nodes

    def getMetricMin(self):
        return tuple(float(v) for v in self.min)

    def getMetricMax(self):
        return tuple(float(v) for v in self.max)

    def getResolution(self):
        return self.resolution

    def points(self):
        """Occupied leaf centres, (occupied, 3) float32 in leaf order."""
        out = np.empty((self.occupied, 3), np.float32)
        _lib.call("a3d_octree_get_points", self._h, 0, self.occupied, _lib.ptr(out))
        return out

    def close(self):
        if self._h:
            _lib.call("a3d_octree_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.a3d_octree_destroy(self._h)
        except Exception:
            pass


class PointCloudInfo:
    """computePointCloud's result (PointCloudTools.h:56-68)."""

    def __init__(self, cloud, mn, mx, res):
        self.cloud = cloud
        self.octo_min_x, self.octo_min_y, self.octo_min_z = (float(v) for v in mn)
        self.octo_max_x, self.octo_max_y, self.octo_max_z = (float(v) for v in mx)
        self.octo_resol = float(res)

    @property
    def min(self):
        return np.array([self.octo_min_x, self.octo_min_y, self.octo_min_z])

    @property
    def max(self):
        return np.array([self.octo_max_x, self.octo_max_y, self.octo_max_z])


def openOcTree(file_path):
    """openOcTree (PointCloudTools.cpp:26-49)."""
    h = ctypes.c_void_p()
    _lib.call("a3d_octree_read", str(file_path).encode(), ctypes.byref(h))
    return OcTree(h.value)


def computePointCloud(octo_tree):
    """computePointCloud (PointCloudTools.cpp:51-82)."""
    if octo_tree is None:
        raise RuntimeError("OcTree is NULL")
    if octo_tree.size() <= 1:
        raise RuntimeError("OcTree is empty")
    return PointCloudInfo(octo_tree.points(), octo_tree.min, octo_tree.max, octo_tree.resolution)


# ------------------------------------------------------------------ writer
def _codes(depth, idx):
    """Child digits of each node from the root, left-aligned to 16 levels:
    digit l (1-based) sits at bits 3*(16-l) .. 3*(16-l)+2."""
    depth = depth.astype(np.int64)
    idx = idx.astype(np.uint64)
    code = np.zeros(len(depth), np.uint64)
    for lvl in range(1, TREE_DEPTH + 1):
        shift = depth - lvl  # bit of the index that holds this level's digit
        live = shift >= 0
        s = np.where(live, shift, 0).astype(np.uint64)
        digit = ((idx[:, 0] >> s) & 1) | (((idx[:, 1] >> s) & 1) << np.uint64(1)) | \
                (((idx[:, 2] >> s) & 1) << np.uint64(2))
        code |= np.where(live, digit, 0).astype(np.uint64) << np.uint64(3 * (TREE_DEPTH - lvl))
    return code


def _tree(depth, idx, check=True):
    """Nodes of the tree whose leaves are given: (codes, depths) of every
    node, leaves first then the inner nodes, plus the leaf count."""
    depth = np.asarray(depth, np.int64)
    idx = np.asarray(idx, np.int64).reshape(-1, 3)
    if len(depth) == 0:
        raise ValueError("no leaves")
    if depth.min() < 0 or depth.max() > TREE_DEPTH:
        raise ValueError("leaf depth outside 0..16")
    if np.any(idx < 0) or np.any(idx >= (np.int64(1) << depth)[:, None]):
        raise ValueError("leaf index outside its depth's range")
    code = _codes(depth, idx)
    srt = np.argsort(code, kind="stable")
    cs, ds = code[srt], depth[srt]
    inner_c, inner_d = [], []
    for a in range(TREE_DEPTH):
        sh = np.uint64(3 * (TREE_DEPTH - a))
        v = cs[ds > a] >> sh  # sorted, so equal prefixes are adjacent
        if len(v) == 0:
            continue
        first = np.empty(len(v), bool)
        first[0] = True
        np.not_equal(v[1:], v[:-1], out=first[1:])
        inner_c.append(v[first] << sh)
        inner_d.append(np.full(int(first.sum()), a, np.int64))
    ic = np.concatenate(inner_c) if inner_c else np.zeros(0, np.uint64)
    idp = np.concatenate(inner_d) if inner_d else np.zeros(0, np.int64)
    if check:  # leaves must be distinct and must not nest
        key_leaf = (code << np.uint64(5)) | depth.astype(np.uint64)
        key_inner = (ic << np.uint64(5)) | idp.astype(np.uint64)
        if len(np.intersect1d(key_leaf, key_inner)) or len(np.unique(key_leaf)) != len(key_leaf):
            raise ValueError("leaves overlap")
    return np.concatenate([code, ic]), np.concatenate([depth, idp]), len(depth)


def _layout(codes, depths):
    """Pre-order (parents before children, children in child order) of the
    nodes, and for every node at depth >= 1 its index, its parent's pre-order
    position and its child digit."""
    key = (codes << np.uint64(5)) | depths.astype(np.uint64)
    order = np.argsort(key)
    sk = key[order]
    child = np.nonzero(depths >= 1)[0]
    c = codes[child]
    dd = depths[child]
    sh = (3 * (TREE_DEPTH - dd)).astype(np.uint64)
    digit = ((c >> sh) & np.uint64(7)).astype(np.int64)
    pcode = (c >> (sh + np.uint64(3))) << (sh + np.uint64(3))
    ppos = np.searchsorted(sk, (pcode << np.uint64(5)) | (dd - 1).astype(np.uint64))
    return order, child, ppos, digit


def _header(magic, nodes, res):
    return ("%s\n# written by amcl3d_amd\nid OcTree\nsize %d\nres %r\ndata\n" % (magic, nodes, float(res))).encode()


def write_octree(path, depth, idx, occupied, res, values=None, check=True):
    """Writes the OcTree with the given leaves (depth, index (i, j, k) at that
    depth, occupied flag) to ``path``: ``.bt`` (binary) or ``.ot`` (full, leaf
    log-odds ``values`` or the octomap clamping defaults).  Leaves must not
    overlap."""
    path = str(path)
    occupied = np.asarray(occupied, bool)
    codes, depths, nleaf = _tree(depth, idx, check)
    order, child, ppos, digit = _layout(codes, depths)
    pos = np.empty(len(order), np.int64)
    pos[order] = np.arange(len(order))
    if path.endswith(".bt"):
        # two bytes per node with children, in pre-order
        kind = np.full(len(codes), 3, np.int64)
        kind[:nleaf] = np.where(occupied, 2, 1)
        bits = kind[child] << (2 * digit)
        mask = np.bincount(ppos, weights=bits.astype(np.float64), minlength=len(codes)).astype(np.int64)
        inner_pos = np.sort(pos[nleaf:])
        m = mask[inner_pos]
        # a root without children is two zero bytes (it reads back occupied)
        body = np.stack([m & 0xFF, m >> 8], 1).astype(np.uint8).tobytes() if len(m) else b"\0\0"
        head = _header("# Octomap OcTree binary file", len(codes), res)
    elif path.endswith(".ot"):
        val = np.empty(len(codes), np.float32)
        lv = np.where(occupied, _OCC_LOGODDS, _FREE_LOGODDS) if values is None else np.asarray(values, np.float32)
        val[:nleaf] = lv
        # inner nodes carry the largest log-odds below them
        inner_val = np.full(len(codes), -np.finfo(np.float32).max, np.float32)
        inner_val[:nleaf] = lv
        for dd in range(TREE_DEPTH, 0, -1):
            sel = child[depths[child] == dd]
            if len(sel) == 0:
                continue
            pp = ppos[np.searchsorted(child, sel)]
            np.fmax.at(inner_val, order[pp], inner_val[sel])  # NaN children do not count
        val[nleaf:] = inner_val[nleaf:]
        cmask = np.bincount(ppos, weights=(1 << digit).astype(np.float64), minlength=len(codes)).astype(np.uint8)
        rec = np.zeros(len(codes), dtype=[("v", "<f4"), ("m", "u1")])
        rec["v"] = val[order]
        rec["m"] = cmask
        body = rec.tobytes()
        head = _header("# Octomap OcTree file", len(codes), res)
    else:
        raise ValueError("the file name must end in .bt or .ot")
    with open(path, "wb") as f:
        f.write(head)
        f.write(body)
    return len(codes)


def _spread3(v):
    """The 16 bits of v moved to bits 0, 3, 6, ... 45."""
    v = v.astype(np.uint64) & np.uint64(0xFFFF)
    for sh, m in ((16, 0x0000FF0000FF), (8, 0x00F00F00F00F), (4, 0x0C30C30C30C3), (2, 0x249249249249)):
        v = (v | (v << np.uint64(sh))) & np.uint64(m)
    return v


def _write_bt_cells(path, key, res):
    """.bt of occupied depth-16 leaves at the distinct keys (n, 3) in one
    pass over their sorted child-digit codes: leaf i opens the inner nodes
    whose prefixes differ from leaf i-1's, at depths lcp .. 15, so the nodes
    come out in pre-order without a sort of the inner nodes."""
    code = np.unique(_spread3(key[:, 0]) | (_spread3(key[:, 1]) << np.uint64(1)) |
                     (_spread3(key[:, 2]) << np.uint64(2)))
    n = len(code)
    # first level whose digit differs from the previous leaf (1..16)
    diff = np.empty(n, np.uint64)
    diff[0] = 0
    np.bitwise_xor(code[1:], code[:-1], out=diff[1:])
    _, e = np.frexp(diff[1:].astype(np.float64))  # highest set bit = e - 1 (exact below 2^53)
    first = np.empty(n, np.int64)
    first[0] = 0
    first[1:] = TREE_DEPTH - (e - 1) // 3
    count = TREE_DEPTH - first  # inner nodes opened by each leaf, depths first .. 15
    end = np.cumsum(count)
    total = int(end[-1])
    depth = (np.arange(total, dtype=np.int64) - np.repeat(end - count, count) + np.repeat(first, count)).astype(np.int8)
    owner = np.repeat(code, count)
    mask = np.zeros(total, np.float64)
    for a in range(1, TREE_DEPTH):
        at = np.nonzero(depth == a)[0]
        above = np.nonzero(depth == a - 1)[0]
        parent = above[np.searchsorted(above, at) - 1]
        digit = ((owner[at] >> np.uint64(3 * (TREE_DEPTH - a))) & np.uint64(7)).astype(np.int64)
        mask += np.bincount(parent, weights=(3 << (2 * digit)).astype(np.float64), minlength=total)
    del owner
    leaf_digit = (code & np.uint64(7)).astype(np.int64)
    mask += np.bincount(end - 1, weights=(2 << (2 * leaf_digit)).astype(np.float64), minlength=total)
    m = mask.astype(np.int64)
    body = np.stack([m & 0xFF, m >> 8], 1).astype(np.uint8).tobytes()
    with open(path, "wb") as f:
        f.write(_header("# Octomap OcTree binary file", total + n, res))
        f.write(body)
    return total + n


def write_points(path, xyz, res):
    """Writes the occupied cells of a point set as a ``.bt``/``.ot`` OcTree of
    resolution ``res`` (one occupied depth-16 leaf per distinct cell, key
    floor(coordinate / res) + 2^15 per axis as octomap's coordToKey)."""
    p = np.asarray(xyz, np.float64).reshape(-1, 3)
    key = np.floor(p * (1.0 / res)).astype(np.int64) + (1 << 15)
    if np.any(key < 0) or np.any(key >= (1 << 16)):
        raise ValueError("points outside the OcTree's 2^16-cell range")
    if str(path).endswith(".bt"):
        return _write_bt_cells(str(path), key, res)
    packed = np.unique((key[:, 0] << 32) | (key[:, 1] << 16) | key[:, 2])
    idx = np.stack([packed >> 32, (packed >> 16) & 0xFFFF, packed & 0xFFFF], 1)
    return write_octree(path, np.full(len(idx), TREE_DEPTH), idx, np.ones(len(idx), bool), res, check=False)
