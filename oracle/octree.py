"""Oracle for the OcTree map files (test infrastructure only, like the rest of
this package): openOcTree + computePointCloud (amcl3d/src/PointCloudTools.cpp:
26-82) restated in pure Python after octomap 1.x, the library the reference
links (not vendored, absent here), following that library's own structure:

* ``.bt``  OccupancyOcTreeBase::readBinaryNode: two bytes per node with
  children, bitset bits (2i, 2i+1) of child i: (1,0) free leaf, (0,1)
  occupied leaf, (1,1) node with children, read after both bytes, in child
  order; such a node then takes the largest log-odds of its children
  (-FLT_MAX when it has none).  The root starts at the clamping maximum.
  AbstractOccupancyOcTree::readBinary: header line, else the legacy header
  (int32 3, f64 res, u32 size); the node count must equal the header's size.
* ``.ot``  AbstractOcTree::read + OcTreeBaseImpl::readNodesRecurs: f32 value,
  a children byte (bit i: child i exists), children in order; tree type
  OcTree (or the deprecated id "1").
* Leaves are visited with octomap's tree iterator (a stack, children pushed
  7..0, depth limit 16); child keys by computeChildKey from the root key
  2^15; coordinates by keyToCoord (floor formula for inner depths); occupied
  when log-odds >= 0 (isNodeOccupied, threshold 0.5); bounds as calcMinMax.

Pins (tests/test_octree.py): the reference's occupied map cloud
(amcl3d/tests/data/mappointcloud_msg.bin, decoded to
tests/golden/kat_map_cloud.npz: 179,551 centres in leaf order) and the
bounds and resolution PointCloudToolsTest.cpp:42-53,166-173 assert, for a
tree rebuilt from those centres; closed-form centres and bounds of single
leaves at every depth; the error cases of PointCloudToolsTest.cpp:59-103.
The (free, occupied) meaning of the .bt bit pairs is taken from octomap's
source as published; no file written by octomap itself is available to pin
it, so that one assignment is parity unpinned.
"""
import math
import os
import struct

FLT_MAX = 3.4028234663852886e38
TREE_DEPTH = 16
TREE_MAX_VAL = 32768
CLAMP_MAX = struct.unpack("<f", struct.pack("<f", 3.5))[0]
CLAMP_MIN = struct.unpack("<f", struct.pack("<f", -2.0))[0]


class OracleError(RuntimeError):
    pass


class _Node:
    __slots__ = ("v", "ch")

    def __init__(self, v=0.0):
        self.v = v
        self.ch = None


class _Stream:
    def __init__(self, data):
        self.d = data
        self.i = 0

    def read(self, n):
        if self.i + n > len(self.d):
            raise OracleError("truncated")
        b = self.d[self.i:self.i + n]
        self.i += n
        return b

    def getline(self):
        j = self.d.find(b"\n", self.i)
        if j < 0:
            line, self.i = self.d[self.i:], len(self.d)
        else:
            line, self.i = self.d[self.i:j], j + 1
        return line

    def token(self):
        ws = b" \t\n\v\f\r"
        while self.i < len(self.d) and self.d[self.i] in ws:
            self.i += 1
        j = self.i
        while j < len(self.d) and self.d[j] not in ws:
            j += 1
        t, self.i = self.d[self.i:j], j
        return t.decode("latin-1")

    def skip_line(self):
        j = self.d.find(b"\n", self.i)
        self.i = len(self.d) if j < 0 else j + 1


def _read_header(s):
    """AbstractOcTree::readHeader."""
    ident, size, res = "", 0, 0.0
    while True:
        tok = s.token()
        if tok == "":
            raise OracleError("no data line")
        if tok == "data":
            s.skip_line()
            break
        if tok.startswith("#"):
            s.skip_line()
        elif tok == "id":
            ident = s.token()
        elif tok == "res":
            res = float(s.token())
        elif tok == "size":
            v = s.token()
            if not v.isdigit() or int(v) > 0xFFFFFFFF:
                raise OracleError("bad size")
            size = int(v)
        else:
            s.skip_line()
    if ident == "" or not res > 0.0:
        raise OracleError("bad header")
    if ident == "1":
        ident = "OcTree"
    return ident, size, res


def _create(node, i, v):
    if node.ch is None:
        node.ch = [None] * 8
    c = _Node(v)
    node.ch[i] = c
    return c


def _read_binary_node(s, node, depth):
    """OccupancyOcTreeBase::readBinaryNode (depth checked: octomap's trees
    have 16 levels)."""
    b = s.read(2)
    node.v = CLAMP_MAX
    inner = []
    for i in range(8):
        byte = b[i // 4]
        lo = (byte >> (2 * (i % 4))) & 1
        hi = (byte >> (2 * (i % 4) + 1)) & 1
        if lo == 1 and hi == 0:
            _create(node, i, CLAMP_MIN)
        elif lo == 0 and hi == 1:
            _create(node, i, CLAMP_MAX)
        elif lo == 1 and hi == 1:
            inner.append(_create(node, i, -200.0))
    for c in inner:
        if depth + 1 >= TREE_DEPTH:
            raise OracleError("node below depth 16")
        _read_binary_node(s, c, depth + 1)
        c.v = max((g.v for g in c.ch if g is not None), default=-FLT_MAX) if c.ch else -FLT_MAX


def _read_nodes_recurs(s, node, depth):
    """OcTreeBaseImpl::readNodesRecurs."""
    node.v = struct.unpack("<f", s.read(4))[0]
    m = s.read(1)[0]
    if m and depth >= TREE_DEPTH:
        raise OracleError("node below depth 16")
    for i in range(8):
        if (m >> i) & 1:
            _read_nodes_recurs(s, _create(node, i, 0.0), depth + 1)


def _count(node):
    n, st = 0, [node]
    while st:
        x = st.pop()
        n += 1
        if x.ch:
            st.extend(c for c in x.ch if c is not None)
    return n


class Tree:
    def __init__(self, root, res):
        self.root = root
        self.res = res

    def size(self):
        return 0 if self.root is None else _count(self.root)

    def node_size(self, depth):
        return self.res * float(1 << (TREE_DEPTH - depth))

    def key_to_coord(self, key, depth):
        """OcTreeBaseImpl::keyToCoord(key, depth)."""
        if depth == 0:
            return 0.0
        if depth == TREE_DEPTH:
            return (float(key - TREE_MAX_VAL) + 0.5) * self.res
        return (math.floor((float(key) - float(TREE_MAX_VAL)) / float(1 << (TREE_DEPTH - depth))) + 0.5) * \
            self.node_size(depth)

    def leaves(self):
        """The leaf iterator: (node, key, depth) in octomap's order."""
        if self.root is None:
            return
        st = [(self.root, (TREE_MAX_VAL,) * 3, 0)]
        while st:
            node, key, depth = st.pop()
            if node.ch is None or depth == TREE_DEPTH or all(c is None for c in node.ch):
                yield node, key, depth
                continue
            off = TREE_MAX_VAL >> (depth + 1)
            for i in range(7, -1, -1):
                c = node.ch[i]
                if c is None:
                    continue
                ck = tuple(key[a] + off if (i >> a) & 1 else key[a] - off - (0 if off else 1) for a in range(3))
                st.append((c, ck, depth + 1))

    def metric_bounds(self):
        """OcTreeBaseImpl::calcMinMax."""
        if self.root is None:
            return [0.0] * 3, [0.0] * 3
        mn, mx = [1.7976931348623157e308] * 3, [-1.7976931348623157e308] * 3
        for _, key, depth in self.leaves():
            size = self.node_size(depth)
            half = size / 2.0
            for a in range(3):
                x = self.key_to_coord(key[a], depth) - half
                if x < mn[a]:
                    mn[a] = x
                x += size
                if x > mx[a]:
                    mx[a] = x
        return mn, mx


def _read_bt(data):
    s = _Stream(data)
    first = s.getline()
    if first.startswith(b"# Octomap OcTree binary file"):
        _, size, res = _read_header(s)
    else:
        s.i = 0
        if len(data) < 4 or struct.unpack("<i", s.read(4))[0] != 3:
            raise OracleError("not an OcTree")
        res = struct.unpack("<d", s.read(8))[0]
        if not res > 0.0:
            raise OracleError("bad resolution")
        size = struct.unpack("<I", s.read(4))[0]
    t = Tree(None, res)
    if size > 0:
        t.root = _Node()
        _read_binary_node(s, t.root, 0)
    if t.size() != size:
        raise OracleError("size mismatch")
    return t


def _read_ot(data):
    s = _Stream(data)
    if not s.getline().startswith(b"# Octomap OcTree file"):
        raise OracleError("not an OcTree file")
    ident, size, res = _read_header(s)
    if ident != "OcTree":
        raise OracleError("not an OcTree")
    t = Tree(None, res)
    if size > 0:
        t.root = _Node()
        _read_nodes_recurs(s, t.root, 0)
    return t


def open_octree(path):
    """openOcTree: a Tree, or RuntimeError with the reference's messages."""
    path = str(path)
    if not os.path.exists(path):
        raise RuntimeError("Cannot find file " + path)
    try:
        with open(path, "rb") as f:
            data = f.read()
    except OSError:
        data = None
    if path.endswith(".bt"):
        try:
            if data is None:
                raise OracleError("unreadable")
            return _read_bt(data)
        except OracleError:
            raise RuntimeError("OcTree cannot be read")
    if path.endswith(".ot"):
        try:
            if data is None:
                raise OracleError("unreadable")
            return _read_ot(data)
        except OracleError:
            pass
    raise RuntimeError("OcTree cannot be created from file " + path)


def compute_point_cloud(tree):
    """computePointCloud: (float32-rounded centres as a list of (x, y, z),
    min, max, res)."""
    if tree is None:
        raise RuntimeError("OcTree is NULL")
    if tree.size() <= 1:
        raise RuntimeError("OcTree is empty")
    mn, mx = tree.metric_bounds()
    f32 = lambda v: struct.unpack("<f", struct.pack("<f", v))[0]  # noqa: E731
    pts = [tuple(f32(tree.key_to_coord(k, d)) for k in key) for node, key, d in tree.leaves() if node.v >= 0.0]
    return pts, mn, mx, tree.res
