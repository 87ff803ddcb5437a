// OcTree map files read on the host: the occupied leaf centres and metric
// bounds that openOcTree + computePointCloud produce
// (amcl3d/src/PointCloudTools.cpp:26-82), and Grid3d::open on top of them
// (amcl3d/src/Grid3d.cpp:25-78): the .grid cache next to the map, else
// computeGrid on the device, then the cache written.
//
// File formats (those of octomap 1.x, the library the reference links):
//  .bt  a first line starting "# Octomap OcTree binary file", then header
//       lines "id <type>", "size <nodes>", "res <metres>" ("#" comment lines
//       and unknown keys skipped) up to a line "data"; or, with no such first
//       line, the legacy header int32 3, f64 res, u32 size.  Then the tree,
//       parents before children: a node with children is two bytes whose bit
//       pairs (2c, 2c+1) of byte c/4 describe child c as 10 free leaf,
//       01 occupied leaf, 11 node with children, 00 absent; the nodes of the
//       children marked 11 follow, in child order.  The node count read must
//       equal the header's size.
//  .ot  a first line starting "# Octomap OcTree file", the same header (id
//       OcTree), then every node parents first: f32 log-odds, a byte whose
//       bit c is set when child c exists, then those children in order.  A
//       leaf is occupied when its log-odds is >= 0 (occupancy threshold 0.5).
// Child c of a node takes the upper half of x when bit 0 of c is set, of y
// for bit 1, of z for bit 2.  The root (depth 0) is a cube of 2^16 cells of
// side res centred on the origin; the node at depth d with index i along an
// axis (0 <= i < 2^d) is centred there at (i - 2^(d-1) + 0.5) * res * 2^(16-d)
// (0 for the root).  Leaves are the nodes without children, visited depth
// first in child order; the bounds are the extremes of their cubes.
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "a3d_handles.h"

using namespace a3d;

struct a3d_octree {
  double res = 0.0;
  uint64_t nodes = 0;
  uint64_t leaves = 0;
  double min[3] = {0.0, 0.0, 0.0};
  double max[3] = {0.0, 0.0, 0.0};
  std::vector<float> xyz;  // occupied leaf centres, 3 floats each, in leaf order
};

namespace {

constexpr int kTreeDepth = 16;

struct Reader {
  const uint8_t* p;
  const uint8_t* end;
  bool ok = true;

  bool take(void* dst, size_t n) {
    if ((size_t)(end - p) < n) return ok = false;
    std::memcpy(dst, p, n);
    p += n;
    return true;
  }
};

bool is_space(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

// The next whitespace-separated token (empty at the end of the data).
std::string next_token(Reader& r) {
  while (r.p < r.end && is_space(*r.p)) ++r.p;
  const uint8_t* b = r.p;
  while (r.p < r.end && !is_space(*r.p)) ++r.p;
  return std::string(reinterpret_cast<const char*>(b), r.p - b);
}

void skip_line(Reader& r) {
  while (r.p < r.end && *r.p++ != '\n') {
  }
}

// Header lines after the first: keys id, size, res up to "data" (the binary
// data starts after the end of that line).
bool read_header(Reader& r, std::string& id, uint64_t& size, double& res) {
  id.clear();
  size = 0;
  res = 0.0;
  for (;;) {
    const std::string tok = next_token(r);
    if (tok.empty()) return false;  // no "data" line
    if (tok == "data") {
      skip_line(r);
      break;
    }
    if (tok[0] == '#') {
      skip_line(r);
    } else if (tok == "id") {
      id = next_token(r);
    } else if (tok == "res") {
      const std::string v = next_token(r);
      char* e = nullptr;
      errno = 0;
      res = std::strtod(v.c_str(), &e);
      if (v.empty() || *e || errno) return false;
    } else if (tok == "size") {
      const std::string v = next_token(r);
      if (v.empty() || v.size() > 10) return false;
      uint64_t s = 0;
      for (char c : v) {
        if (c < '0' || c > '9') return false;
        s = 10 * s + (uint64_t)(c - '0');
      }
      if (s > 0xffffffffull) return false;
      size = s;
    } else {
      skip_line(r);  // unknown keyword
    }
  }
  if (id.empty() || !(res > 0.0)) return false;
  if (id == "1") id = "OcTree";  // the deprecated id of OcTree files
  return true;
}

bool first_line_is(Reader& r, const char* header) {
  const uint8_t* nl = static_cast<const uint8_t*>(std::memchr(r.p, '\n', r.end - r.p));
  const uint8_t* le = nl ? nl : r.end;
  const size_t n = std::strlen(header);
  if ((size_t)(le - r.p) < n || std::memcmp(r.p, header, n) != 0) return false;
  r.p = nl ? nl + 1 : r.end;
  return true;
}

struct Walk {
  a3d_octree* t;
  double size[kTreeDepth + 1];  // node side by depth: res * 2^(16-d)

  explicit Walk(a3d_octree* tree) : t(tree) {
    for (int d = 0; d <= kTreeDepth; ++d) size[d] = t->res * (double)(1u << (kTreeDepth - d));
    for (int a = 0; a < 3; ++a) {
      t->min[a] = DBL_MAX;
      t->max[a] = -DBL_MAX;
    }
  }

  double centre(int d, uint32_t i) const {
    if (d == 0) return 0.0;
    return ((double)((int64_t)i - (int64_t)(1u << (d - 1))) + 0.5) * size[d];
  }

  void leaf(int d, const uint32_t ix[3], bool occupied) {
    ++t->leaves;
    double c[3];
    for (int a = 0; a < 3; ++a) {
      c[a] = centre(d, ix[a]);
      // the cube's extremes, in the order and arithmetic octomap's bounds use
      double lo = c[a] - size[d] / 2.0;
      if (lo < t->min[a]) t->min[a] = lo;
      lo += size[d];
      if (lo > t->max[a]) t->max[a] = lo;
    }
    if (occupied)
      for (int a = 0; a < 3; ++a) t->xyz.push_back((float)c[a]);
  }

  static void child_index(const uint32_t ix[3], int c, uint32_t out[3]) {
    for (int a = 0; a < 3; ++a) out[a] = 2 * ix[a] + ((c >> a) & 1);
  }

  // .bt: the two bytes of a node with children at depth d, then its subtrees.
  // Returns false on truncated data or a node below depth 16.
  bool bt_node(Reader& r, int d, const uint32_t ix[3], bool& has_children) {
    uint8_t b[2];
    if (!r.take(b, 2)) return false;
    has_children = false;
    for (int c = 0; c < 8; ++c) {
      const int bits = (b[c >> 2] >> (2 * (c & 3))) & 3;
      if (bits == 0) continue;  // absent
      has_children = true;
      ++t->nodes;
      uint32_t cx[3];
      child_index(ix, c, cx);
      if (bits == 3) {
        if (d + 1 >= kTreeDepth) return false;
        bool grand = false;
        if (!bt_node(r, d + 1, cx, grand)) return false;
        // a node marked as having children whose bytes list none is a leaf
        // with the smallest log-odds: free
        if (!grand) leaf(d + 1, cx, false);
      } else {
        leaf(d + 1, cx, bits == 2);  // 01: occupied (bit 2c+1), 10: free (bit 2c)
      }
    }
    return true;
  }

  // .ot: a node at depth d (value, children byte, children).
  bool ot_node(Reader& r, int d, const uint32_t ix[3]) {
    float v;
    uint8_t m;
    if (!r.take(&v, 4) || !r.take(&m, 1)) return false;
    ++t->nodes;
    if (m == 0) {
      leaf(d, ix, v >= 0.0f);
      return true;
    }
    if (d >= kTreeDepth) return false;
    for (int c = 0; c < 8; ++c) {
      if (!((m >> c) & 1)) continue;
      uint32_t cx[3];
      child_index(ix, c, cx);
      if (!ot_node(r, d + 1, cx)) return false;
    }
    return true;
  }
};

bool ends_with(const std::string& s, const char* suffix) {
  const size_t n = std::strlen(suffix);
  return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

bool read_file(const std::string& path, std::vector<uint8_t>& buf) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  buf.clear();
  uint8_t chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

// OcTree::readBinary (.bt).  False: "OcTree cannot be read".
bool parse_bt(const std::vector<uint8_t>& buf, a3d_octree* t) {
  Reader r{buf.data(), buf.data() + buf.size()};
  uint64_t size = 0;
  double res = 0.0;
  if (first_line_is(r, "# Octomap OcTree binary file")) {
    std::string id;
    if (!read_header(r, id, size, res)) return false;
  } else {
    r.p = buf.data();  // the legacy header
    int32_t type = -1;
    uint32_t s32 = 0;
    if (!r.take(&type, 4) || type != 3) return false;
    if (!r.take(&res, 8) || !(res > 0.0)) return false;
    if (!r.take(&s32, 4)) return false;
    size = s32;
  }
  t->res = res;
  Walk w(t);
  if (size > 0) {
    t->nodes = 1;
    const uint32_t root[3] = {0, 0, 0};
    bool has_children = false;
    if (!w.bt_node(r, 0, root, has_children)) return false;
    // a root without children is a leaf, with the log-odds of an occupied one
    if (!has_children) w.leaf(0, root, true);
  }
  return t->nodes == size;
}

// AbstractOcTree::read (.ot) as an OcTree.  False: no tree.
bool parse_ot(const std::vector<uint8_t>& buf, a3d_octree* t) {
  Reader r{buf.data(), buf.data() + buf.size()};
  if (!first_line_is(r, "# Octomap OcTree file")) return false;
  std::string id;
  uint64_t size = 0;
  double res = 0.0;
  if (!read_header(r, id, size, res) || id != "OcTree") return false;
  t->res = res;
  Walk w(t);
  if (size > 0) {
    const uint32_t root[3] = {0, 0, 0};
    if (!w.ot_node(r, 0, root)) return false;
  }
  return true;
}

void reset_bounds_if_empty(a3d_octree* t) {
  if (t->leaves) return;
  for (int a = 0; a < 3; ++a) t->min[a] = t->max[a] = 0.0;
}

}  // namespace

extern "C" {

int a3d_octree_read(const char* path, a3d_octree** out) {
  if (!path || !out) return fail(A3D_ERR_ARG, "a3d_octree_read: null argument");
  *out = nullptr;
  const std::string p(path);
  struct stat st;
  if (::stat(path, &st) != 0) return fail(A3D_ERR_IO, "Cannot find file " + p);
  a3d_octree* t = new a3d_octree();
  std::vector<uint8_t> buf;
  if (ends_with(p, ".bt")) {
    if (!read_file(p, buf) || !parse_bt(buf, t)) {
      delete t;
      return fail(A3D_ERR_IO, "OcTree cannot be read");
    }
  } else if (ends_with(p, ".ot")) {
    if (!read_file(p, buf) || !parse_ot(buf, t)) {
      delete t;
      return fail(A3D_ERR_IO, "OcTree cannot be created from file " + p);
    }
  } else {
    delete t;
    return fail(A3D_ERR_IO, "OcTree cannot be created from file " + p);
  }
  reset_bounds_if_empty(t);
  *out = t;
  return A3D_OK;
}

int a3d_octree_get_info(const a3d_octree* t, double min[3], double max[3], double* res, uint64_t* nodes,
                        uint64_t* occupied) {
  if (!t) return fail(A3D_ERR_ARG, "a3d_octree_get_info: null tree");
  for (int a = 0; a < 3; ++a) {
    if (min) min[a] = t->min[a];
    if (max) max[a] = t->max[a];
  }
  if (res) *res = t->res;
  if (nodes) *nodes = t->nodes;
  if (occupied) *occupied = t->xyz.size() / 3;
  return A3D_OK;
}

int a3d_octree_get_points(const a3d_octree* t, uint64_t first, uint64_t count, float* xyz) {
  if (!t || (count && !xyz)) return fail(A3D_ERR_ARG, "a3d_octree_get_points: null argument");
  const uint64_t n = t->xyz.size() / 3;
  if (first > n || count > n - first) return fail(A3D_ERR_ARG, "a3d_octree_get_points: range past the last point");
  if (count) std::memcpy(xyz, t->xyz.data() + 3 * first, count * 3 * sizeof(float));
  return A3D_OK;
}

int a3d_octree_destroy(a3d_octree* t) {
  delete t;
  return A3D_OK;
}

int a3d_grid_open(const char* map_path, double sensor_dev, int keep_dist, a3d_grid** out) {
  if (!map_path || !out) return fail(A3D_ERR_ARG, "a3d_grid_open: null argument");
  *out = nullptr;
  a3d_octree* t = nullptr;
  int rc = a3d_octree_read(map_path, &t);
  if (rc) return rc;
  if (t->nodes <= 1) {  // PointCloudTools.cpp:56-58
    a3d_octree_destroy(t);
    return fail(A3D_ERR_IO, "OcTree is empty");
  }
  // the cache path as Grid3d.cpp:48-52 forms it: the map path up to the first
  // ".bt" (".ot") in it, then ".grid"
  const std::string p(map_path);
  const std::string grid_path = p.substr(0, p.find(ends_with(p, ".bt") ? ".bt" : ".ot")) + ".grid";
  a3d_grid* G = nullptr;
  rc = a3d_grid_load(grid_path.c_str(), t->min, t->max, t->res, sensor_dev, &G);
  if (rc == A3D_ERR_HIP || rc == A3D_ERR_NOMEM) {
    a3d_octree_destroy(t);
    return rc;
  }
  if (rc != A3D_OK) {
    // no usable cache: computeGrid on the device (the .grid format stores
    // dist, so it is built), then the cache written; a failed write leaves
    // the grid open, as saveGrid's result is not checked (Grid3d.cpp:72-75)
    const size_t n = t->xyz.size() / 3;
    rc = a3d_grid_create_from_points(t->xyz.data(), n, t->min, t->max, t->res, sensor_dev, 1, &G);
    if (rc) {
      a3d_octree_destroy(t);
      return rc;
    }
    (void)a3d_grid_save(G, grid_path.c_str());
  }
  a3d_octree_destroy(t);
  if (!keep_dist && G->dist) {
    (void)hipStreamSynchronize(G->stream);
    (void)hipFree(G->dist);
    G->dist = nullptr;
    G->g.dist = nullptr;
  }
  *out = G;
  return A3D_OK;
}

}  // extern "C"
