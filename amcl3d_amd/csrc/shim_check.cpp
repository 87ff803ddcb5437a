// Drives the header-only C++ layer (include/amcl3d_gpu.hpp) the way the
// reference's node drives Grid3d and ParticleFilter (Node.cpp:41,154,162,181),
// on inputs read from a file, and writes every result to an output file.
//
//   shim_check unopened               reference soft failures without a map
//   shim_check map <file> <sd> <out>  Grid3d::open(map_path, sensor_dev)
//   shim_check <input> <output> [neg] one init -> predict -> update -> resample;
//                                     with neg, init is given -n particles
//                                     (the reference resizes to abs(n))
//
// Input (little-endian): u32 npts, m, nb, n; f64 min[3], max[3], res,
// sensor_dev, mods[4], deltas[4], alpha, sigma; f32 u01, pose[4], devs[4];
// f32 pts[3*npts], cloud[3*m], ranges[4*nb], z_init[4n], z_pred[4n].
// Output: f32 init particles[7n], mean after init[4], updated particles[7n],
// mean after update[4], resampled particles[7n], weight of the mean pose[1],
// isIntoMap(mean)[1].
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/amcl3d_gpu.hpp"

namespace {

struct Point {
  float x, y, z, pad;
};
struct Cloud {
  std::vector<Point> points;
};

// Replays given draws in the order the filter asks for them.
class Replay : public amcl3d::DrawSource {
 public:
  std::vector<float> z;
  size_t pos = 0;
  float u = 0.f;
  void normals(float* out, size_t count) override {
    for (size_t i = 0; i < count; ++i) out[i] = pos < z.size() ? z[pos++] : 0.f;
  }
  float uniform01() override { return u; }
};

template <class T>
bool rd(FILE* f, T* p, size_t count) {
  return std::fread(p, sizeof(T), count, f) == count;
}

int unopened() {
  amcl3d::Grid3d g;
  Cloud* c = new Cloud;
  c->points.resize(20, Point{0.5f, 0.5f, 0.5f, 0.f});
  const float w = g.computeCloudWeight(c, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
  const bool in = g.isIntoMap(0.5f, 0.5f, 0.5f);
  delete c;
  std::printf("{\"weight\": %.9g, \"into_map\": %d, \"open\": %d}\n", w, in ? 1 : 0, g.isOpen() ? 1 : 0);
  return (w == 0.f && !in && !g.isOpen()) ? 0 : 1;
}

// Grid3d::open(map_path, sensor_dev) as the node calls it (Node.cpp:41):
// writes the grid's sizes, bounds and prob cells; exit 3 when open returns
// false, as it does for a missing or empty map.
int open_map(const char* path, double sd, const char* outp) {
  amcl3d::Grid3d g;
  if (!g.open(std::string(path), sd)) {
    std::printf("{\"open\": 0, \"error\": \"%s\"}\n", amcl3d::lastError().c_str());
    return 3;
  }
  uint32_t size[3];
  double mn[3], mx[3], res, s;
  if (a3d_grid_get_info(g.handle(), size, mn, mx, &res, &s) != A3D_OK) return 1;
  const uint64_t cells = (uint64_t)size[0] * size[1] * size[2];
  std::vector<float> prob(cells);
  if (a3d_grid_read_cells(g.handle(), 0, cells, prob.data(), nullptr) != A3D_OK) return 1;
  FILE* o = std::fopen(outp, "wb");
  if (!o) return 1;
  const bool ok = std::fwrite(size, 4, 3, o) == 3 && std::fwrite(mn, 8, 3, o) == 3 && std::fwrite(mx, 8, 3, o) == 3 &&
                  std::fwrite(prob.data(), 4, cells, o) == cells;
  std::fclose(o);
  std::printf("{\"open\": 1, \"cells\": %llu}\n", (unsigned long long)cells);
  return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 2 && std::strcmp(argv[1], "unopened") == 0) return unopened();
  if (argc == 5 && std::strcmp(argv[1], "map") == 0) return open_map(argv[2], std::atof(argv[3]), argv[4]);
  const bool neg = argc == 4 && std::strcmp(argv[3], "neg") == 0;
  if (argc != 3 && !neg) {
    std::fprintf(stderr, "usage: shim_check unopened | shim_check map <file.bt|.ot> <sensor_dev> <output> | "
                 "shim_check <input> <output> [neg]\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t hdr[4];
  double mn[3], mx[3], res, sd, mods[4], deltas[4], alpha, sigma;
  float u01, pose[4], devs[4];
  bool ok = rd(f, hdr, 4) && rd(f, mn, 3) && rd(f, mx, 3) && rd(f, &res, 1) && rd(f, &sd, 1) && rd(f, mods, 4) &&
            rd(f, deltas, 4) && rd(f, &alpha, 1) && rd(f, &sigma, 1) && rd(f, &u01, 1) && rd(f, pose, 4) &&
            rd(f, devs, 4);
  const uint32_t npts = hdr[0], m = hdr[1], nb = hdr[2], n = hdr[3];
  std::vector<float> pts(3 * (size_t)npts), cl(3 * (size_t)m), rg(4 * (size_t)nb);
  Replay draws;
  draws.z.resize(8 * (size_t)n);
  draws.u = u01;
  ok = ok && rd(f, pts.data(), pts.size()) && rd(f, cl.data(), cl.size()) && rd(f, rg.data(), rg.size()) &&
       rd(f, draws.z.data(), draws.z.size());
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "shim_check: short input\n");
    return 2;
  }
  Cloud* cloud = new Cloud;
  for (uint32_t j = 0; j < m; ++j) cloud->points.push_back(Point{cl[3 * j], cl[3 * j + 1], cl[3 * j + 2], 1.f});
  std::vector<amcl3d::Range> ranges;
  for (uint32_t b = 0; b < nb; ++b) ranges.emplace_back(rg[4 * b], rg[4 * b + 1], rg[4 * b + 2], rg[4 * b + 3]);

  amcl3d::Grid3d grid;
  if (!grid.open(pts.data(), npts, mn, mx, res, sd)) {
    std::fprintf(stderr, "shim_check: open failed: %s\n", amcl3d::lastError().c_str());
    return 1;
  }
  amcl3d::ParticleFilter pf(&draws);
  std::vector<float> out;
  auto put_particles = [&]() {
    for (const auto& p : pf.particles()) out.insert(out.end(), {p.x, p.y, p.z, p.a, p.w, p.wp, p.wr});
  };
  auto put_mean = [&]() {
    const amcl3d::Particle q = pf.getMean();
    out.insert(out.end(), {q.x, q.y, q.z, q.a});
  };
  pf.init(neg ? -(int)n : (int)n, pose[0], pose[1], pose[2], pose[3], devs[0], devs[1], devs[2], devs[3]);
  if (!pf.isInitialized()) {
    std::fprintf(stderr, "shim_check: init failed: %s\n", amcl3d::lastError().c_str());
    return 1;
  }
  put_particles();
  put_mean();
  pf.predict(mods[0], mods[1], mods[2], mods[3], deltas[0], deltas[1], deltas[2], deltas[3]);
  pf.update(grid, cloud, ranges, alpha, sigma);
  put_particles();
  put_mean();
  pf.resample();
  put_particles();
  const amcl3d::Particle q = pf.getMean();
  out.push_back(grid.computeCloudWeight(cloud, q.x, q.y, q.z, 0.f, 0.f, q.a));
  out.push_back(grid.isIntoMap(q.x, q.y, q.z) ? 1.f : 0.f);
  delete cloud;
  FILE* o = std::fopen(argv[2], "wb");
  if (!o || std::fwrite(out.data(), sizeof(float), out.size(), o) != out.size()) return 1;
  std::fclose(o);
  std::printf("shim_check: %u particles, %u points, %u beacons\n", n, m, nb);
  return 0;
}
