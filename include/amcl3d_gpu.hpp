/*
 * amcl3d_gpu.hpp -- header-only C++ layer over the C ABI (amcl3d_gpu.h) with
 * the class names and method shapes of the reference's library target
 * `amcl3d`, so that the node's call sites compile against it unchanged apart
 * from the include and the draw source:
 *
 *   amcl3d::Grid3d          amcl3d/src/Grid3d.h:34-159
 *   amcl3d::ParticleFilter  amcl3d/src/ParticleFilter.h:67-213
 *   amcl3d::Particle/Range  amcl3d/src/ParticleFilter.h:35-63
 *
 * Behaviour kept from the reference:
 *  - Grid3d::open returns bool and never throws (Grid3d.cpp:25-78);
 *    computeCloudWeight returns 0 and isIntoMap false while no map is open
 *    (Grid3d.cpp:136-137, 203-204).
 *  - ParticleFilter methods are void; getMean() is the mean of the last
 *    init/update (ParticleFilter.h:94-97).
 * Differences, each forced by the boundary:
 *  - Clouds: any point cloud whose points have float members x, y, z
 *    (pcl::PointCloud<pcl::PointXYZ>::Ptr included), taken as a template
 *    parameter so this header needs no PCL.
 *  - Random draws come from a DrawSource (default: std::mt19937 seeded from
 *    std::random_device with std::normal_distribution<float>(0, 1) and
 *    std::uniform_real_distribution<float>(0, 1), the reference's generator
 *    family, ParticleFilter.cpp:22,246-256).  The device applies each
 *    particle's deviation to the standard normal.
 *  - Grid3d::open(map_path, sensor_dev) reads the OcTree file as the
 *    reference does; two more overloads take the map's occupied points and
 *    bounds, or a .grid file plus bounds.
 *  - Failures of the device path (no GPU, out of memory) are reported by
 *    lastError(); the reference has no such failure mode.
 */
#ifndef AMCL3D_GPU_HPP
#define AMCL3D_GPU_HPP

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "amcl3d_gpu.h"

namespace amcl3d {

struct Particle {  // ParticleFilter.h:35-49
  float x = 0, y = 0, z = 0, a = 0;
  float w = 0, wp = 0, wr = 0;
};

struct Range {  // ParticleFilter.h:53-63
  float r, ax, ay, az;
  Range(float r_, float ax_, float ay_, float az_) : r(r_), ax(ax_), ay(ay_), az(az_) {}
};

inline std::string lastError() { return a3d_last_error(); }

// Standard normals and uniforms for predict / init / resample.
class DrawSource {
 public:
  DrawSource() : gen_(std::random_device()()) {}
  explicit DrawSource(uint32_t seed) : gen_(seed) {}
  virtual ~DrawSource() = default;
  virtual void normals(float* z, size_t count) {
    std::normal_distribution<float> d(0.f, 1.f);
    for (size_t i = 0; i < count; ++i) z[i] = d(gen_);
  }
  virtual float uniform01() {
    std::uniform_real_distribution<float> d(0.f, 1.f);
    return d(gen_);
  }

 private:
  std::mt19937 gen_;
};

namespace detail {
// Splits a cloud into x, y, z arrays (pcl::PointXYZ is 16-byte AoS).
template <class CloudPtr>
inline size_t split_cloud(const CloudPtr& cloud, std::vector<float>& x, std::vector<float>& y,
                          std::vector<float>& z) {
  const size_t m = cloud ? cloud->points.size() : 0;
  x.resize(m);
  y.resize(m);
  z.resize(m);
  for (size_t j = 0; j < m; ++j) {
    x[j] = cloud->points[j].x;
    y[j] = cloud->points[j].y;
    z[j] = cloud->points[j].z;
  }
  return m;
}
}  // namespace detail

class Grid3d {
 public:
  Grid3d() = default;
  Grid3d(const Grid3d&) = delete;
  Grid3d& operator=(const Grid3d&) = delete;
  ~Grid3d() { close(); }

  // Grid3d::open(map_path, sensor_dev) (Grid3d.cpp:25-78): the OcTree file
  // (.bt / .ot), its .grid cache, else computeGrid on the device.
  bool open(const std::string& map_path, double sensor_dev) {
    close();
    return a3d_grid_open(map_path.c_str(), sensor_dev, 0, &g_) == A3D_OK;
  }
  // Builds the grid on the device from the occupied points (computeGrid).
  bool open(const float* xyz, size_t n, const double min[3], const double max[3], double res, double sensor_dev) {
    close();
    return a3d_grid_create_from_points(xyz, n, min, max, res, sensor_dev, 0, &g_) == A3D_OK;
  }
  // Loads a .grid cache (Grid3d::loadGrid layout); bounds as for open().
  bool open(const std::string& grid_path, const double min[3], const double max[3], double res, double sensor_dev) {
    close();
    return a3d_grid_load(grid_path.c_str(), min, max, res, sensor_dev, &g_) == A3D_OK;
  }
  bool save(const std::string& grid_path) const { return g_ && a3d_grid_save(g_, grid_path.c_str()) == A3D_OK; }
  void close() {
    if (g_) a3d_grid_destroy(g_);
    g_ = nullptr;
  }
  bool isOpen() const { return g_ != nullptr; }

  // Grid3d::computeCloudWeight (Grid3d.cpp:133-199).
  template <class CloudPtr>
  float computeCloudWeight(const CloudPtr& cloud, float tx, float ty, float tz, float roll, float pitch,
                           float yaw) const {
    if (!g_) return 0.f;
    std::vector<float> x, y, z;
    const size_t m = detail::split_cloud(cloud, x, y, z);
    float out = 0.f;
    if (a3d_grid_cloud_weight(g_, x.data(), y.data(), z.data(), m, &tx, &ty, &tz, &yaw, 1, roll, pitch, &out) !=
        A3D_OK)
      return 0.f;
    return out;
  }

  // Grid3d::isIntoMap (Grid3d.cpp:201-208).
  bool isIntoMap(float x, float y, float z) const {
    int in = 0;
    return g_ && a3d_grid_is_into_map(g_, x, y, z, &in) == A3D_OK && in != 0;
  }

  const a3d_grid* handle() const { return g_; }

 private:
  a3d_grid* g_ = nullptr;
};

class ParticleFilter {
 public:
  explicit ParticleFilter(DrawSource* draws = nullptr) : draws_(draws) {
    if (!draws_) {
      own_.reset(new DrawSource());
      draws_ = own_.get();
    }
    if (a3d_pf_create(&pf_) != A3D_OK) pf_ = nullptr;
  }
  ParticleFilter(const ParticleFilter&) = delete;
  ParticleFilter& operator=(const ParticleFilter&) = delete;
  ~ParticleFilter() {
    if (pf_) a3d_pf_destroy(pf_);
  }

  bool isInitialized() const {
    int v = 0;
    return pf_ && a3d_pf_is_initialized(pf_, &v) == A3D_OK && v != 0;
  }

  Particle getMean() const {
    Particle p;
    float m[4] = {0, 0, 0, 0};
    if (pf_ && a3d_pf_get_mean(pf_, m) == A3D_OK) {
      p.x = m[0];
      p.y = m[1];
      p.z = m[2];
      p.a = m[3];
    }
    return p;
  }

  // ParticleFilter::init (ParticleFilter.cpp:46-95).
  void init(int num_particles, float x_init, float y_init, float z_init, float a_init, float x_dev, float y_dev,
            float z_dev, float a_dev) {
    // p_.resize(abs(num_particles)) (ParticleFilter.cpp:51): a negative
    // count is taken by its magnitude
    const size_t n = (size_t)std::llabs((long long)num_particles);
    if (!pf_ || n == 0) return;
    z4_.resize(4 * n);
    draws_->normals(z4_.data(), z4_.size());
    a3d_pf_init(pf_, num_particles, x_init, y_init, z_init, a_init, x_dev, y_dev, z_dev, a_dev, z4_.data());
  }

  // ParticleFilter::predict (ParticleFilter.cpp:97-119).
  void predict(double odom_x_mod, double odom_y_mod, double odom_z_mod, double odom_a_mod, double delta_x,
               double delta_y, double delta_z, double delta_a) {
    size_t n = 0;
    if (!pf_ || a3d_pf_size(pf_, &n) != A3D_OK || n == 0) return;
    z4_.resize(4 * n);
    draws_->normals(z4_.data(), z4_.size());
    a3d_pf_predict(pf_, odom_x_mod, odom_y_mod, odom_z_mod, odom_a_mod, delta_x, delta_y, delta_z, delta_a,
                   z4_.data());
  }

  // ParticleFilter::update (ParticleFilter.cpp:121-196).
  template <class CloudPtr>
  void update(const Grid3d& grid3d, const CloudPtr& cloud, const std::vector<Range>& range_data, double alpha,
              double sigma, double roll = 0.0, double pitch = 0.0) {
    if (!pf_) return;
    std::vector<float> x, y, z;
    const size_t m = detail::split_cloud(cloud, x, y, z);
    std::vector<float> r(4 * range_data.size());
    for (size_t b = 0; b < range_data.size(); ++b) {
      r[4 * b] = range_data[b].r;
      r[4 * b + 1] = range_data[b].ax;
      r[4 * b + 2] = range_data[b].ay;
      r[4 * b + 3] = range_data[b].az;
    }
    a3d_pf_update(pf_, grid3d.handle(), x.data(), y.data(), z.data(), m, r.data(), range_data.size(), alpha, sigma,
                  roll, pitch);
  }

  // ParticleFilter::resample (ParticleFilter.cpp:198-222).
  void resample() {
    if (pf_) a3d_pf_resample(pf_, draws_->uniform01(), A3D_RESAMPLE_REFERENCE, nullptr);
  }

  // Particle state (replaces buildParticlesPoseMsg, ParticleFilter.cpp:30-44).
  std::vector<Particle> particles() const {
    size_t n = 0;
    std::vector<Particle> out;
    if (!pf_ || a3d_pf_size(pf_, &n) != A3D_OK) return out;
    out.resize(n);
    static_assert(sizeof(Particle) == 28, "Particle is 7 floats");
    if (n && a3d_pf_get_particles(pf_, reinterpret_cast<float*>(out.data()), n) != A3D_OK) out.clear();
    return out;
  }

  a3d_pf* handle() const { return pf_; }

 private:
  a3d_pf* pf_ = nullptr;
  DrawSource* draws_ = nullptr;
  std::unique_ptr<DrawSource> own_;
  std::vector<float> z4_;
};

}  // namespace amcl3d

#endif  // AMCL3D_GPU_HPP
