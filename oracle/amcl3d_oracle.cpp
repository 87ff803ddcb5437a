// CPU restatement of the amcl3d measurement-update hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the timed CPU
// baseline; the product path never calls it.
//
// It restates, line by line and with the same float/double mixing, the
// reference code in:
//   amcl3d/src/PointCloudTools.cpp:84-149   computeGrid
//   amcl3d/src/Grid3d.cpp:133-208           computeCloudWeight, isIntoMap
//   amcl3d/src/ParticleFilter.cpp:46-244    init, predict, update, resample,
//                                           computeRangeWeight
// Build flags: g++ -O2 -ffp-contract=off (no FMA contraction, no fast-math),
// i.e. the arithmetic of the reference built by GCC on x86-64 with SSE math.
//
// Decisions that the reference leaves open (each one is pinned by a test):
//  * Transcendentals: the reference calls unqualified sin/cos/exp/sqrt on
//    float arguments; built with GCC 5 these resolve to the double versions,
//    so every such call is evaluated in double and rounded once to float
//    (ParticleFilter.cpp:77,110-111,239-240; Grid3d.cpp:139-144).
//  * Random draws: std::mt19937 + normal_distribution<float> are replaced by
//    host-supplied standard normals z; a draw with standard deviation s is
//    (float)(z * (float)s), which is what normal_distribution<float> does
//    with a generated standard normal (ParticleFilter.cpp:246-250).  The
//    uniform draw of resample is host-supplied too (ParticleFilter.cpp:202).
//  * computeGrid's nearest neighbour (PCL KdTreeFLANN, FLANN L2_Simple<float>)
//    is restated as the exact minimum over all map points of the float
//    squared distance ((dx*dx + dy*dy) + dz*dz), dx = s.x - p.x in float;
//    a kd-tree with a conservative pruning rule finds that minimum.
//  * The 250M-cell cap of PointCloudTools.cpp:104-105 is lifted and cell
//    indices are 64-bit.
//  * resample: when the float running sum ends below u (the reference then
//    reads p_[N], ParticleFilter.cpp:212-216) the index is clamped to N-1.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstddef>
#include <algorithm>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct GridView {
  const float* prob;
  uint32_t size[3];
  double min[3];
  double max[3];
  double res;
};

// Grid3d.cpp:139-149 with the GCC-5 overloads: sin/cos in double, each
// rotation entry rounded once to float.
void rotation(float roll, float pitch, float yaw, float r[9]) {
  const double sr = std::sin((double)roll), cr = std::cos((double)roll);
  const double sp = std::sin((double)pitch), cp = std::cos((double)pitch);
  const double sy = std::sin((double)yaw), cy = std::cos((double)yaw);
  r[0] = (float)(cy * cp);
  r[1] = (float)(cy * sp * sr - sy * cr);
  r[2] = (float)(cy * sp * cr + sy * sr);
  r[3] = (float)(sy * cp);
  r[4] = (float)(sy * sp * sr + cy * cr);
  r[5] = (float)(sy * sp * cr - cy * sr);
  r[6] = (float)(-sp);
  r[7] = (float)(cp * sr);
  r[8] = (float)(cp * cr);
}

// The cell each point of the cloud lands in under one pose, as
// Grid3d.cpp:139-196 computes it: visit(k, idx) for every point k inside the
// map, in cloud order.  Shared by cloud_weight and orc_cloud_cells.
template <class Visit>
void cloud_cells(const GridView& g, const float* px, const float* py, const float* pz, size_t m, float tx, float ty,
                 float tz, float roll, float pitch, float yaw, Visit visit) {
  float r[9];
  rotation(roll, pitch, yaw, r);
  const double sx = g.max[0] - g.min[0];
  const double sy = g.max[1] - g.min[1];
  const double sz = g.max[2] - g.min[2];
  const double ox = tx - g.min[0];
  const double oy = ty - g.min[1];
  const double oz = tz - g.min[2];
  const uint64_t step_y = g.size[0];
  const uint64_t step_z = (uint64_t)g.size[0] * g.size[1];
  const uint64_t grid_size = step_z * g.size[2];
  const float error_z = 0;
  for (size_t k = 0; k < m; ++k) {
    const float x = px[k], y = py[k], z = pz[k] + error_z;
    const float nx = (float)((double)(x * r[0] + y * r[1] + z * r[2]) + ox);
    const float ny = (float)((double)(x * r[3] + y * r[4] + z * r[5]) + oy);
    const float nz = (float)((double)(x * r[6] + y * r[7] + z * r[8]) + oz);
    if (nx >= 0 && nx < sx && ny >= 0 && ny < sy && nz >= 0 && nz < sz) {
      const uint32_t ix = (uint32_t)std::floor(nx / g.res);
      const uint32_t iy = (uint32_t)std::floor(ny / g.res);
      const uint32_t iz = (uint32_t)std::floor(nz / g.res);
      if (ix < g.size[0] && iy < g.size[1] && iz < g.size[2]) {
        const uint64_t idx = ix + iy * step_y + iz * step_z;
        if (idx < grid_size) visit(k, idx);
      }
    }
  }
}

// Grid3d.cpp:133-199: the float sum of prob over the points inside the map,
// in cloud order, divided by their count when there are more than 10.
float cloud_weight(const GridView& g, const float* px, const float* py, const float* pz, size_t m, float tx,
                   float ty, float tz, float roll, float pitch, float yaw) {
  if (!g.prob) return 0;
  float weight = 0.f;
  int n = 0;
  cloud_cells(g, px, py, pz, m, tx, ty, tz, roll, pitch, yaw, [&](size_t, uint64_t idx) {
    weight += g.prob[idx];
    n += 1;
  });
  return (n <= 10) ? 0 : weight / n;
}

// Grid3d.cpp:201-208
bool is_into_map(const GridView& g, float x, float y, float z) {
  if (!g.prob) return false;
  return x >= g.min[0] && x < g.max[0] && y >= g.min[1] && y < g.max[1] && z >= g.min[2] && z < g.max[2];
}

// ParticleFilter.cpp:224-244
float range_weight(float x, float y, float z, const float* rng, size_t nb, double sigma) {
  if (nb == 0) return 0;
  float w = 1;
  const float k1 = (float)(1.f / (sigma * std::sqrt(2 * kPi)));
  const float k2 = (float)(0.5f / (sigma * sigma));
  for (size_t i = 0; i < nb; ++i) {
    const float rr = rng[4 * i + 0], ax = rng[4 * i + 1], ay = rng[4 * i + 2], az = rng[4 * i + 3];
    const float r = (float)std::sqrt((double)((x - ax) * (x - ax) + (y - ay) * (y - ay) + (z - az) * (z - az)));
    const float e = -k2 * (r - rr) * (r - rr);
    w = (float)((double)w * ((double)k1 * std::exp((double)e)));
  }
  return w;
}

// normal_distribution<float>(0, s) applied to a standard normal z.
inline float gauss_draw(float z, double sigma) {
  const float s = (float)sigma;
  const float mean = (float)0.0;
  return z * s + mean;
}

struct Particle {
  float x, y, z, a, w, wp, wr;
};
static_assert(sizeof(Particle) == 28, "Particle is 7 floats");

// ----------------------------------------------------------------------------
// Exact float nearest neighbour (restating KdTreeFLANN::nearestKSearch, k=1):
// the minimum over all map points of the FLANN L2_Simple<float> squared
// distance.  A kd-tree prunes a box only when its double-precision lower
// bound, shrunk by 1e-6 relative, still exceeds the best float distance; the
// float distance of any point is at least (1 - 3e-7) times its exact value,
// so no pruned point can be the minimum.
// ----------------------------------------------------------------------------
struct KdTree {
  struct Node {
    float lo[3], hi[3];
    uint32_t begin, end;  // point range
    int32_t left, right;  // children (-1: leaf)
  };
  std::vector<float> px, py, pz;
  std::vector<Node> nodes;

  size_t count() const { return px.size(); }

  int32_t make(std::vector<uint32_t>& ord, const float* xyz, uint32_t b, uint32_t e) {
    Node nd;
    for (int a = 0; a < 3; ++a) {
      nd.lo[a] = INFINITY;
      nd.hi[a] = -INFINITY;
    }
    for (uint32_t i = b; i < e; ++i)
      for (int a = 0; a < 3; ++a) {
        nd.lo[a] = std::min(nd.lo[a], xyz[3 * ord[i] + a]);
        nd.hi[a] = std::max(nd.hi[a], xyz[3 * ord[i] + a]);
      }
    nd.begin = b;
    nd.end = e;
    nd.left = nd.right = -1;
    const int32_t id = (int32_t)nodes.size();
    nodes.push_back(nd);
    if (e - b > 8) {
      int ax = 0;
      for (int a = 1; a < 3; ++a)
        if (nd.hi[a] - nd.lo[a] > nd.hi[ax] - nd.lo[ax]) ax = a;
      const uint32_t mid = b + (e - b) / 2;
      std::nth_element(ord.begin() + b, ord.begin() + mid, ord.begin() + e,
                       [&](uint32_t u, uint32_t v) { return xyz[3 * u + ax] < xyz[3 * v + ax]; });
      const int32_t l = make(ord, xyz, b, mid);
      const int32_t r = make(ord, xyz, mid, e);
      nodes[id].left = l;
      nodes[id].right = r;
    }
    return id;
  }

  void build(const float* xyz, size_t n) {
    std::vector<uint32_t> ord;
    ord.reserve(n);
    for (size_t i = 0; i < n; ++i)
      if (std::isfinite(xyz[3 * i]) && std::isfinite(xyz[3 * i + 1]) && std::isfinite(xyz[3 * i + 2]))
        ord.push_back((uint32_t)i);  // PCL drops invalid points
    if (ord.empty()) return;
    nodes.reserve(2 * ord.size() / 4 + 8);
    make(ord, xyz, 0, (uint32_t)ord.size());
    px.resize(ord.size());
    py.resize(ord.size());
    pz.resize(ord.size());
    for (size_t i = 0; i < ord.size(); ++i) {
      px[i] = xyz[3 * ord[i]];
      py[i] = xyz[3 * ord[i] + 1];
      pz[i] = xyz[3 * ord[i] + 2];
    }
  }

  static double box_lb(const Node& nd, float qx, float qy, float qz) {
    const double q[3] = {qx, qy, qz};
    double s = 0;
    for (int a = 0; a < 3; ++a) {
      double d = 0;
      if (q[a] < nd.lo[a]) d = (double)nd.lo[a] - q[a];
      else if (q[a] > nd.hi[a]) d = q[a] - (double)nd.hi[a];
      s += d * d;
    }
    return s;
  }

  float nearest_d2(float qx, float qy, float qz) const {
    float best = INFINITY;
    if (nodes.empty()) return best;
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node& nd = nodes[stack[--sp]];
      if (box_lb(nd, qx, qy, qz) * (1.0 - 1e-6) > (double)best) continue;
      if (nd.left < 0) {
        for (uint32_t i = nd.begin; i < nd.end; ++i) {
          const float dx = qx - px[i];
          const float dy = qy - py[i];
          const float dz = qz - pz[i];
          float d = 0.f;
          d += dx * dx;
          d += dy * dy;
          d += dz * dz;
          if (d < best) best = d;
        }
        continue;
      }
      const double bl = box_lb(nodes[nd.left], qx, qy, qz);
      const double br = box_lb(nodes[nd.right], qx, qy, qz);
      // push the farther child first so the nearer one is searched first
      if (bl <= br) {
        stack[sp++] = nd.right;
        stack[sp++] = nd.left;
      } else {
        stack[sp++] = nd.left;
        stack[sp++] = nd.right;
      }
    }
    return best;
  }
};

}  // namespace

extern "C" {

// Grid sizes, PointCloudTools.cpp:93-98.
void orc_grid_sizes(const double mn[3], const double mx[3], double res, uint32_t size_out[3]) {
  for (int a = 0; a < 3; ++a) size_out[a] = (uint32_t)std::ceil((mx[a] - mn[a]) / res);
}

// computeGrid, PointCloudTools.cpp:84-149 (cap lifted, 64-bit indices).
// dist/prob: size[0]*size[1]*size[2] floats each, x fastest
// (index = ix + iy*sx + iz*sx*sy, PointCloudTools.cpp:131).
// exp_mode 0: prob uses glibc expf (the reference's call);
// exp_mode 1: prob uses (float)exp((double)x), the correctly rounded expf.
// Returns the number of map points used (finite ones).
int64_t orc_compute_grid(const float* xyz, size_t n, const double mn[3], const double mx[3], double res,
                         double sensor_dev, float* dist_out, float* prob_out, int nthreads, int exp_mode) {
  uint32_t size[3];
  orc_grid_sizes(mn, mx, res, size);
  const uint64_t sx = size[0], sy = size[1], sz = size[2];
  const float c1 = (float)(1. / (sensor_dev * std::sqrt(2 * kPi)));
  const float c2 = (float)(1. / (2. * sensor_dev * sensor_dev));
  KdTree idx;
  idx.build(xyz, n);
  const bool empty = idx.count() == 0;
  if (nthreads < 1) nthreads = 1;
  auto work = [&](int t) {
    for (uint64_t iz = t; iz < sz; iz += nthreads) {
      const float qz = (float)(mn[2] + (iz * res));
      for (uint64_t iy = 0; iy < sy; ++iy) {
        const float qy = (float)(mn[1] + (iy * res));
        for (uint64_t ix = 0; ix < sx; ++ix) {
          const float qx = (float)(mn[0] + (ix * res));
          const uint64_t o = ix + iy * sx + iz * sx * sy;
          if (empty) {
            dist_out[o] = -1.0f;
            prob_out[o] = 0.0f;
            continue;
          }
          const float d = idx.nearest_d2(qx, qy, qz);
          dist_out[o] = d;
          const float e = -d * d * c2;
          prob_out[o] = c1 * (exp_mode == 0 ? expf(e) : (float)std::exp((double)e));
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  return (int64_t)idx.count();
}

// Nearest float squared distance for arbitrary query points (used to check
// sampled cells of large maps).
void orc_nearest_d2(const float* xyz, size_t n, const double mn[3], const double mx[3], double res,
                    const float* q, size_t nq, float* d2_out) {
  (void)mn;
  (void)mx;
  (void)res;
  KdTree idx;
  idx.build(xyz, n);
  for (size_t i = 0; i < nq; ++i)
    d2_out[i] = idx.count() ? idx.nearest_d2(q[3 * i], q[3 * i + 1], q[3 * i + 2]) : -1.0f;
}

// Grid3d::computeCloudWeight (Grid3d.cpp:133-199); prob == NULL means "map
// not open" and returns 0 (Grid3d.cpp:136-137).
float orc_cloud_weight(const float* prob, const uint32_t size[3], const double mn[3], const double mx[3],
                       double res, const float* px, const float* py, const float* pz, size_t m, float tx, float ty,
                       float tz, float roll, float pitch, float yaw) {
  GridView g{prob, {size[0], size[1], size[2]}, {mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, res};
  return cloud_weight(g, px, py, pz, m, tx, ty, tz, roll, pitch, yaw);
}

// Many poses, one cloud: out[i] = computeCloudWeight(cloud, t_i, roll, pitch, yaw_i).
void orc_cloud_weight_batch(const float* prob, const uint32_t size[3], const double mn[3], const double mx[3],
                            double res, const float* px, const float* py, const float* pz, size_t m,
                            const float* tx, const float* ty, const float* tz, const float* yaw, size_t n,
                            float roll, float pitch, float* out, int nthreads) {
  GridView g{prob, {size[0], size[1], size[2]}, {mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, res};
  if (nthreads < 1) nthreads = 1;
  auto work = [&](int t) {
    for (size_t i = t; i < n; i += nthreads)
      out[i] = cloud_weight(g, px, py, pz, m, tx[i], ty[i], tz[i], roll, pitch, yaw[i]);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
}

// Cells gathered by computeCloudWeight(cloud, t_i, roll, pitch, yaw_i):
// out[i*m + k] = the cell index of point k under pose i, or -1 when the point
// is outside the map.
void orc_cloud_cells(const uint32_t size[3], const double mn[3], const double mx[3], double res, const float* px,
                     const float* py, const float* pz, size_t m, const float* tx, const float* ty, const float* tz,
                     const float* yaw, size_t n, float roll, float pitch, int64_t* out) {
  static const float dummy = 0.f;  // the map is open
  GridView g{&dummy, {size[0], size[1], size[2]}, {mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, res};
  for (size_t i = 0; i < n; ++i) {
    int64_t* o = out + i * m;
    for (size_t k = 0; k < m; ++k) o[k] = -1;
    cloud_cells(g, px, py, pz, m, tx[i], ty[i], tz[i], roll, pitch, yaw[i],
                [&](size_t k, uint64_t idx) { o[k] = (int64_t)idx; });
  }
}

int orc_is_into_map(const uint32_t size[3], const double mn[3], const double mx[3], int open, float x, float y,
                    float z) {
  static const float one = 1.f;
  GridView g{open ? &one : nullptr, {size[0], size[1], size[2]}, {mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, 0};
  return is_into_map(g, x, y, z) ? 1 : 0;
}

float orc_range_weight(float x, float y, float z, const float* ranges, size_t nb, double sigma) {
  return range_weight(x, y, z, ranges, nb, sigma);
}

// ParticleFilter::init (ParticleFilter.cpp:46-95).  p: n particles of 7
// floats {x,y,z,a,w,wp,wr}; z4: 4*n standard normals, z4[4i..4i+3] are the
// x,y,z,a draws of particle i (the first four are unused: particle 0 is the
// initial pose itself).  mean_out: {x,y,z,a}.
void orc_pf_init(float* pbuf, int n, float x_init, float y_init, float z_init, float a_init, float x_dev,
                 float y_dev, float z_dev, float a_dev, const float* z4, float* mean_out) {
  Particle* p = reinterpret_cast<Particle*>(pbuf);
  const float dev = std::max(std::max(x_dev, y_dev), z_dev);
  const float gauss_const_1 = (float)(1. / (dev * std::sqrt(2 * kPi)));
  const float gauss_const_2 = (float)(1. / (2 * dev * dev));
  p[0].x = x_init;
  p[0].y = y_init;
  p[0].z = z_init;
  p[0].a = a_init;
  p[0].w = gauss_const_1;
  float wt = p[0].w;
  for (int i = 1; i < n; ++i) {
    p[i].x = p[0].x + gauss_draw(z4[4 * i + 0], x_dev);
    p[i].y = p[0].y + gauss_draw(z4[4 * i + 1], y_dev);
    p[i].z = p[0].z + gauss_draw(z4[4 * i + 2], z_dev);
    p[i].a = p[0].a + gauss_draw(z4[4 * i + 3], a_dev);
    const float dist = (float)std::sqrt((double)((p[i].x - p[0].x) * (p[i].x - p[0].x) +
                                                 (p[i].y - p[0].y) * (p[i].y - p[0].y) +
                                                 (p[i].z - p[0].z) * (p[i].z - p[0].z)));
    p[i].w = (float)(gauss_const_1 * std::exp((double)(-dist * dist * gauss_const_2)));
    wt += p[i].w;
  }
  float mx = 0, my = 0, mz = 0, ma = 0;
  for (int i = 0; i < n; ++i) {
    p[i].w /= wt;
    mx += p[i].w * p[i].x;
    my += p[i].w * p[i].y;
    mz += p[i].w * p[i].z;
    ma += p[i].w * p[i].a;
  }
  mean_out[0] = mx;
  mean_out[1] = my;
  mean_out[2] = mz;
  mean_out[3] = ma;
}

// ParticleFilter::predict (ParticleFilter.cpp:97-119).  mods/deltas: x,y,z,a.
void orc_pf_predict(float* pbuf, int n, const double mods[4], const double deltas[4], const float* z4) {
  Particle* p = reinterpret_cast<Particle*>(pbuf);
  const double x_dev = std::fabs(deltas[0] * mods[0]);
  const double y_dev = std::fabs(deltas[1] * mods[1]);
  const double z_dev = std::fabs(deltas[2] * mods[2]);
  const double a_dev = std::fabs(deltas[3] * mods[3]);
  for (int i = 0; i < n; ++i) {
    const float sa = (float)std::sin((double)p[i].a);
    const float ca = (float)std::cos((double)p[i].a);
    const float rand_x = (float)(deltas[0] + (double)gauss_draw(z4[4 * i + 0], x_dev));
    const float rand_y = (float)(deltas[1] + (double)gauss_draw(z4[4 * i + 1], y_dev));
    p[i].x += ca * rand_x - sa * rand_y;
    p[i].y += sa * rand_x + ca * rand_y;
    p[i].z = (float)((double)p[i].z + (deltas[2] + (double)gauss_draw(z4[4 * i + 2], z_dev)));
    p[i].a = (float)((double)p[i].a + (deltas[3] + (double)gauss_draw(z4[4 * i + 3], a_dev)));
  }
}

// ParticleFilter::update (ParticleFilter.cpp:121-196).  ranges: nb x
// {r, ax, ay, az}.  prob == NULL means the map is not open.
void orc_pf_update(float* pbuf, int n, const float* prob, const uint32_t size[3], const double mn[3],
                   const double mx[3], double res, const float* px, const float* py, const float* pz, size_t m,
                   const float* ranges, size_t nb, double alpha, double sigma, double roll, double pitch,
                   float* mean_out, int nthreads) {
  Particle* p = reinterpret_cast<Particle*>(pbuf);
  GridView g{prob, {size[0], size[1], size[2]}, {mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, res};
  // The per-particle terms are independent; threads only split that loop.
  std::vector<uint8_t> in(n);
  if (nthreads < 1) nthreads = 1;
  auto work = [&](int t) {
    for (int i = t; i < n; i += nthreads) {
      const float tx = p[i].x, ty = p[i].y, tz = p[i].z;
      in[i] = is_into_map(g, tx, ty, tz);
      if (!in[i]) continue;
      p[i].wp = cloud_weight(g, px, py, pz, m, tx, ty, tz, (float)roll, (float)pitch, p[i].a);
      p[i].wr = range_weight(tx, ty, tz, ranges, nb, sigma);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  float wtp = 0, wtr = 0;
  for (int i = 0; i < n; ++i) {
    if (!in[i]) {
      p[i].w = 0;
      continue;
    }
    wtp += p[i].wp;
    wtr += p[i].wr;
  }
  float wt = 0;
  for (int i = 0; i < n; ++i) {
    if (wtp > 0)
      p[i].wp /= wtp;
    else
      p[i].wp = 0;
    if (wtr > 0)
      p[i].wr /= wtr;
    else
      p[i].wr = 0;
    if (!is_into_map(g, p[i].x, p[i].y, p[i].z))
      p[i].w = 0;
    else
      p[i].w = (float)(p[i].wp * alpha + p[i].wr * (1 - alpha));
    wt += p[i].w;
  }
  float mxx = 0, myy = 0, mzz = 0, maa = 0;
  for (int i = 0; i < n; ++i) {
    if (wt > 0)
      p[i].w /= wt;
    else
      p[i].w = 0;
    mxx += p[i].w * p[i].x;
    myy += p[i].w * p[i].y;
    mzz += p[i].w * p[i].z;
    maa += p[i].w * p[i].a;
  }
  mean_out[0] = mxx;
  mean_out[1] = myy;
  mean_out[2] = mzz;
  mean_out[3] = maa;
}

// Resampling index selection.
// mode 0: ParticleFilter.cpp:198-222 (float running sum), index clamped to
//         n-1 where the reference would read past the end.
// mode 1: FP64 prefix sum over the weights.  Blocked association, identical on
//         the device: within each block of 1024 weights a sequential double
//         sum; block totals summed sequentially; u_m = (u01 + m) / n in double.
void orc_resample_indices(const float* w, int n, float u01, int mode, uint32_t* idx_out) {
  if (mode == 0) {
    const float factor = 1.f / (size_t)n;
    const float r = factor * u01;
    float c = w[0];
    uint32_t i = 0;
    for (uint32_t m = 0; m < (uint32_t)n; ++m) {
      const float u = r + factor * m;
      while (u > c) {
        if (++i >= (uint32_t)n) break;
        c += w[i];
      }
      idx_out[m] = std::min<uint32_t>(i, (uint32_t)n - 1);
    }
    return;
  }
  const int B = 1024;
  std::vector<double> c(n);
  double carry = 0.0;
  for (int b0 = 0; b0 < n; b0 += B) {
    double s = 0.0;
    const int b1 = std::min(n, b0 + B);
    for (int i = b0; i < b1; ++i) {
      s += (double)w[i];
      c[i] = s;
    }
    for (int i = b0; i < b1; ++i) c[i] = carry + c[i];
    carry = carry + s;
  }
  const double inv = 1.0 / (double)n;
  for (int m = 0; m < n; ++m) {
    const double u = ((double)u01 + (double)m) * inv;
    // first i with u <= c[i]
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (u <= c[mid])
        hi = mid;
      else
        lo = mid + 1;
    }
    idx_out[m] = (uint32_t)std::min(lo, n - 1);
  }
}

// ParticleFilter::resample (ParticleFilter.cpp:198-222): new_p[m] = p[idx[m]],
// new_p[m].w = 1/n.
void orc_pf_resample(float* pbuf, int n, float u01, int mode, uint32_t* idx_out) {
  Particle* p = reinterpret_cast<Particle*>(pbuf);
  std::vector<float> w(n);
  for (int i = 0; i < n; ++i) w[i] = p[i].w;
  std::vector<uint32_t> idx(n);
  orc_resample_indices(w.data(), n, u01, mode, idx.data());
  std::vector<Particle> np(n);
  const float factor = 1.f / (size_t)n;
  for (int m = 0; m < n; ++m) {
    np[m] = p[idx[m]];
    np[m].w = factor;
  }
  std::memcpy(p, np.data(), sizeof(Particle) * n);
  if (idx_out) std::memcpy(idx_out, idx.data(), sizeof(uint32_t) * n);
}

// Slice of the grid as Grid3d::buildGridSliceMsg scales it (Grid3d.cpp:100-118):
// int8(prob * (100 / max_prob)) over one z-plane.  Used only as a test pin.
void orc_slice_int8(const float* prob_plane, size_t count, int8_t* out) {
  float max_prob = -1.0;
  for (size_t i = 0; i < count; ++i)
    if (prob_plane[i] > max_prob) max_prob = prob_plane[i];
  if (max_prob < 0.000001f) max_prob = 0.000001f;
  max_prob = 100.f / max_prob;
  for (size_t i = 0; i < count; ++i) out[i] = static_cast<int8_t>(prob_plane[i] * max_prob);
}

// pcl::VoxelGrid<pcl::PointXYZ>::filter as the node calls it (Node.cpp:130-141:
// setLeafSize(voxel_size) x3, defaults otherwise).  PCL is not vendored in
// the reference; this restates VoxelGrid::applyFilter of PCL 1.7.2 (the ROS
// Kinetic / Ubuntu 16.04 package) with its defaults: no filter field,
// min_points_per_voxel 0, downsample_all_data true (CentroidPoint: float
// sums of x, y, z over the voxel's points in the order of the sorted index
// vector, divided by the count as float).  Non-finite points are skipped (the
// is_dense == false path).  inverse leaf = 1.0f / leaf; bounds and voxel
// coordinates use float floor:
//   min_b = (int)floor(min_p * inv); ijk = (int)(floor(p * inv) - (float)min_b)
//   idx = ijk0 + ijk1 * div0 + ijk2 * div0 * div1   (int)
// If the voxel count (int64)((max-min)*inv)+1 per axis overflows int32 the
// input is returned unchanged, as PCL does.  order_mode 0 sorts (idx, point)
// pairs by idx with std::sort (PCL's unstable sort; the order of points
// inside a voxel is whatever libstdc++'s introsort leaves), 1 by (idx, point
// index).  Returns the number of output points written to out_xyz (3 floats
// each; capacity n).
size_t orc_voxel_grid(const float* xyz, size_t n, float leaf, int order_mode, float* out_xyz) {
  const float inv = 1.0f / leaf;
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  size_t nfin = 0;
  for (size_t i = 0; i < n; ++i) {
    const float* p = xyz + 3 * i;
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
    ++nfin;
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], p[k]);
      mx[k] = std::max(mx[k], p[k]);
    }
  }
  if (nfin == 0) return 0;
  int64_t d[3];
  for (int k = 0; k < 3; ++k) d[k] = static_cast<int64_t>((mx[k] - mn[k]) * inv) + 1;
  if (d[0] * d[1] * d[2] > static_cast<int64_t>(INT32_MAX)) {  // indices would overflow: input unchanged
    std::memcpy(out_xyz, xyz, sizeof(float) * 3 * n);
    return n;
  }
  int min_b[3], max_b[3], div_b[3];
  for (int k = 0; k < 3; ++k) {
    min_b[k] = static_cast<int>(std::floor(mn[k] * inv));
    max_b[k] = static_cast<int>(std::floor(mx[k] * inv));
    div_b[k] = max_b[k] - min_b[k] + 1;
  }
  const int mul[3] = {1, div_b[0], div_b[0] * div_b[1]};
  struct Entry {
    unsigned int idx;
    unsigned int point;
  };
  std::vector<Entry> v;
  v.reserve(nfin);
  for (size_t i = 0; i < n; ++i) {
    const float* p = xyz + 3 * i;
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
    int ijk[3];
    for (int k = 0; k < 3; ++k) ijk[k] = static_cast<int>(std::floor(p[k] * inv) - static_cast<float>(min_b[k]));
    const int idx = ijk[0] * mul[0] + ijk[1] * mul[1] + ijk[2] * mul[2];
    v.push_back(Entry{static_cast<unsigned int>(idx), static_cast<unsigned int>(i)});
  }
  if (order_mode == 0)
    std::sort(v.begin(), v.end(), [](const Entry& a, const Entry& b) { return a.idx < b.idx; });
  else
    std::sort(v.begin(), v.end(),
              [](const Entry& a, const Entry& b) { return a.idx < b.idx || (a.idx == b.idx && a.point < b.point); });
  size_t out = 0;
  for (size_t first = 0; first < v.size();) {
    size_t last = first + 1;
    while (last < v.size() && v[last].idx == v[first].idx) ++last;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (size_t li = first; li < last; ++li) {
      const float* p = xyz + 3 * (size_t)v[li].point;
      sx += p[0];
      sy += p[1];
      sz += p[2];
    }
    const float cnt = static_cast<float>(last - first);
    out_xyz[3 * out + 0] = sx / cnt;
    out_xyz[3 * out + 1] = sy / cnt;
    out_xyz[3 * out + 2] = sz / cnt;
    ++out;
    first = last;
  }
  return out;
}

}  // extern "C"
