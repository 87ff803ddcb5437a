// The ROS node's glue around the filter, without ROS (amcl3d/src/Node.cpp):
// odometry increments -> predict deltas, update gating, the VoxelGrid step,
// resample interval, NaN/Inf/jump detection and the base/odom -> world
// transforms the node publishes.  Host code over the library's own entry
// points; the transforms are restated from the tf (Bullet LinearMath) double
// arithmetic the node uses.
#include <cmath>
#include <cstring>
#include <vector>

#include "a3d_handles.h"

using namespace a3d;

namespace {

// tf::Transform: a 3x3 double basis (rows) and an origin.
struct Tf {
  double b[3][3];
  double o[3];
};

Tf tf_identity() {
  Tf t{};
  for (int i = 0; i < 3; ++i) t.b[i][i] = 1.0;
  return t;
}

// Matrix3x3::setRotation(q), q = (x, y, z, w).
void basis_from_quat(const double q[4], double b[3][3]) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double d = x * x + y * y + z * z + w * w;
  const double s = 2.0 / d;
  const double xs = x * s, ys = y * s, zs = z * s;
  const double wx = w * xs, wy = w * ys, wz = w * zs;
  const double xx = x * xs, xy = x * ys, xz = x * zs;
  const double yy = y * ys, yz = y * zs, zz = z * zs;
  b[0][0] = 1.0 - (yy + zz);
  b[0][1] = xy - wz;
  b[0][2] = xz + wy;
  b[1][0] = xy + wz;
  b[1][1] = 1.0 - (xx + zz);
  b[1][2] = yz - wx;
  b[2][0] = xz - wy;
  b[2][1] = yz + wx;
  b[2][2] = 1.0 - (xx + yy);
}

Tf tf_from(const double t[3], const double q[4]) {
  Tf r;
  basis_from_quat(q, r.b);
  for (int i = 0; i < 3; ++i) r.o[i] = t[i];
  return r;
}

double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Transform::inverse: basis transposed, origin -(B^T o).
Tf tf_inv(const Tf& t) {
  Tf r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.b[i][j] = t.b[j][i];
  const double no[3] = {-t.o[0], -t.o[1], -t.o[2]};
  for (int i = 0; i < 3; ++i) r.o[i] = dot3(r.b[i], no);
  return r;
}

// Transform::operator*: (B1 B2, B1 o2 + o1); the product's entries as
// Matrix3x3's operator* forms them (row of B1 against a column of B2).
Tf tf_mul(const Tf& a, const Tf& c) {
  Tf r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.b[i][j] = c.b[0][j] * a.b[i][0] + c.b[1][j] * a.b[i][1] + c.b[2][j] * a.b[i][2];
  for (int i = 0; i < 3; ++i) r.o[i] = dot3(a.b[i], c.o) + a.o[i];
  return r;
}

// Matrix3x3::getRPY (getEulerYPR, first solution).
void tf_rpy(const Tf& t, double& roll, double& pitch, double& yaw) {
  const double(&m)[3][3] = t.b;
  if (std::fabs(m[2][0]) >= 1) {
    yaw = 0;
    const double delta = std::atan2(m[2][1], m[2][2]);
    pitch = m[2][0] < 0 ? M_PI / 2.0 : -M_PI / 2.0;
    roll = delta;
  } else {
    pitch = -std::asin(m[2][0]);
    const double c = std::cos(pitch);
    roll = std::atan2(m[2][1] / c, m[2][2] / c);
    yaw = std::atan2(m[1][0] / c, m[0][0] / c);
  }
}

// Matrix3x3::getRotation: the basis as a quaternion (x, y, z, w).
void tf_quat(const Tf& t, double q[4]) {
  const double(&m)[3][3] = t.b;
  const double trace = m[0][0] + m[1][1] + m[2][2];
  double tmp[4];
  if (trace > 0.0) {
    double s = std::sqrt(trace + 1.0);
    tmp[3] = s * 0.5;
    s = 0.5 / s;
    tmp[0] = (m[2][1] - m[1][2]) * s;
    tmp[1] = (m[0][2] - m[2][0]) * s;
    tmp[2] = (m[1][0] - m[0][1]) * s;
  } else {
    const int i = m[0][0] < m[1][1] ? (m[1][1] < m[2][2] ? 2 : 1) : (m[0][0] < m[2][2] ? 2 : 0);
    const int j = (i + 1) % 3;
    const int k = (i + 2) % 3;
    double s = std::sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0);
    tmp[i] = s * 0.5;
    s = 0.5 / s;
    tmp[3] = (m[k][j] - m[j][k]) * s;
    tmp[j] = (m[j][i] + m[i][j]) * s;
    tmp[k] = (m[k][i] + m[i][k]) * s;
  }
  for (int a = 0; a < 4; ++a) q[a] = tmp[a];
}

// Quaternion::setRPY (setEulerZYX(yaw, pitch, roll)).
void quat_from_rpy(double roll, double pitch, double yaw, double q[4]) {
  const double hy = yaw * 0.5, hp = pitch * 0.5, hr = roll * 0.5;
  const double cy = std::cos(hy), sy = std::sin(hy);
  const double cp = std::cos(hp), sp = std::sin(hp);
  const double cr = std::cos(hr), sr = std::sin(hr);
  q[0] = sr * cp * cy - cr * sp * sy;
  q[1] = cr * sp * cy + sr * cp * sy;
  q[2] = cr * cp * sy - sr * sp * cy;
  q[3] = cr * cp * cy + sr * sp * sy;
}

double yaw_of(const Tf& t) {
  double r, p, y;
  tf_rpy(t, r, p, y);
  return y;
}

// ros::Duration(d) in nanoseconds: whole seconds by floor, the rest rounded
// half away from zero.
int64_t duration_ns(double d) {
  const double sec = std::floor(d);
  return (int64_t)sec * 1000000000LL + (int64_t)std::llround((d - sec) * 1e9);
}

void tf_out(const Tf& t, double tr[3], double q[4]) {
  for (int i = 0; i < 3; ++i) tr[i] = t.o[i];
  tf_quat(t, q);
}

}  // namespace

struct a3d_node {
  a3d_node_params p{};
  const a3d_grid* grid = nullptr;
  a3d_pf* pf = nullptr;
  uint64_t seed = 0;
  uint64_t step = 0;  // draw step: 0 for init, then one per predict
  bool is_odom = false;
  bool amcl_out = false;
  bool has_takenoff = false;
  int n_updates = 0;
  double roll = 0.0;
  double pitch = 0.0;
  float mean[4] = {0.f, 0.f, 0.f, 0.f};
  float lastmean[4] = {0.f, 0.f, 0.f, 0.f};
  double deltas[4] = {0.0, 0.0, 0.0, 0.0};
  Tf lastbase_2_world = tf_identity();
  Tf initodom_2_world = tf_identity();
  Tf lastodom_2_world = tf_identity();
  Tf amcl_out_lastbase_2_odom = tf_identity();
  Tf lastupdatebase_2_odom = tf_identity();
  Tf base_2_odom = tf_identity();
  int64_t nextupdate_ns = 0;
  std::vector<float> ranges;  // {r, ax, ay, az} per measurement since the last update
  std::vector<float> z4;
  // device draws of a predict step: 4 normals per particle, the step's
  // uniform, and the (seed, step) key (made on the filter's stream)
  uint64_t key[2] = {0, 0};
  float* d_z4 = nullptr;
  float* d_u01 = nullptr;
  uint64_t* d_key = nullptr;
  // the scan on the device: raw points, voxel centroids (x, y, z rows of
  // v_cap floats), ranges; grown when a larger input arrives
  void* d_raw = nullptr;
  size_t raw_cap = 0;  // bytes
  float* d_v = nullptr;
  size_t v_cap = 0;  // points
  float* d_rng = nullptr;
  size_t rng_cap = 0;  // floats
  VoxelWorkspace* vws = nullptr;  // VoxelGrid's sort buffers
};

namespace {

int node_init_pose(a3d_node* nd, const Tf& init_pose, float xdev, float ydev, float zdev, float adev) {
  // Node::setInitialPose (Node.cpp:375-398)
  nd->initodom_2_world = init_pose;
  const float x = (float)init_pose.o[0], y = (float)init_pose.o[1], z = (float)init_pose.o[2];
  const float a = (float)yaw_of(init_pose);
  const int n = nd->p.num_particles;
  nd->z4.resize(4 * (size_t)n);
  int rc = a3d_session_draws(nd->seed, 0, n, nd->z4.data(), nullptr);
  if (rc) return rc;
  nd->step = 1;
  rc = a3d_pf_init(nd->pf, n, x, y, z, a, xdev, ydev, zdev, adev, nd->z4.data());
  if (rc) return rc;
  rc = a3d_pf_get_mean(nd->pf, nd->mean);
  if (rc) return rc;
  std::memcpy(nd->lastmean, nd->mean, sizeof(nd->mean));
  nd->lastupdatebase_2_odom = nd->base_2_odom;
  return A3D_OK;
}

bool update_due(a3d_node* nd, int64_t now_ns) {
  // Node::checkUpdateThresholds (Node.cpp:347-373)
  if (now_ns < nd->nextupdate_ns) return false;
  const Tf inc = tf_mul(tf_inv(nd->lastupdatebase_2_odom), nd->base_2_odom);
  if (std::sqrt(dot3(inc.o, inc.o)) > nd->p.d_th) return true;
  double r, p, y;
  tf_rpy(inc, r, p, y);
  return std::fabs(y) > nd->p.a_th;
}

}  // namespace

extern "C" {

int a3d_node_create(const a3d_node_params* params, const a3d_grid* grid, uint64_t seed, a3d_node** out) {
  if (!params || !grid || !out) return fail(A3D_ERR_ARG, "a3d_node_create: null argument");
  *out = nullptr;
  if (params->num_particles <= 0 || !(params->update_rate > 0) || !(params->voxel_size > 0))
    return fail(A3D_ERR_ARG, "a3d_node_create: num_particles, update_rate and voxel_size must be positive");
  // VoxelGrid takes the leaf as a float (setLeafSize(float, float, float),
  // Node.cpp:134): it must stay positive and finite after that rounding
  const float leaf = (float)params->voxel_size;
  if (!(leaf > 0.0f) || !std::isfinite(leaf))
    return fail(A3D_ERR_ARG, "a3d_node_create: voxel_size must be a positive, finite float");
  a3d_node* nd = new a3d_node();
  nd->p = *params;
  nd->grid = grid;
  nd->seed = seed;
  const int rc = a3d_pf_create(&nd->pf);
  if (rc) {
    delete nd;
    return rc;
  }
  if (hipMalloc(&nd->d_z4, 16 * (size_t)params->num_particles) != hipSuccess ||
      hipMalloc(&nd->d_u01, 4) != hipSuccess || hipMalloc(&nd->d_key, 16) != hipSuccess ||
      !(nd->vws = voxel_workspace_new())) {
    a3d_node_destroy(nd);
    return fail(A3D_ERR_NOMEM, "a3d_node_create: cannot allocate the draw buffers");
  }
  *out = nd;
  return A3D_OK;
}

int a3d_node_destroy(a3d_node* nd) {
  if (!nd) return A3D_OK;
  if (nd->pf) a3d_pf_destroy(nd->pf);  // waits for the filter's stream
  (void)hipFree(nd->d_z4);
  (void)hipFree(nd->d_u01);
  (void)hipFree(nd->d_key);
  (void)hipFree(nd->d_raw);
  (void)hipFree(nd->d_v);
  (void)hipFree(nd->d_rng);
  voxel_workspace_free(nd->vws);
  delete nd;
  return A3D_OK;
}

int a3d_node_set_initial_pose(a3d_node* nd, const double t[3], const double q[4], float xdev, float ydev, float zdev,
                              float adev) {
  if (!nd || !t || !q) return fail(A3D_ERR_ARG, "a3d_node_set_initial_pose: null argument");
  return node_init_pose(nd, tf_from(t, q), xdev, ydev, zdev, adev);
}

int a3d_node_odometry(a3d_node* nd, const double t[3], const double q[4], a3d_node_output* out) {
  // Node::odomCallback (Node.cpp:193-324)
  if (!nd || !t || !q) return fail(A3D_ERR_ARG, "a3d_node_odometry: null argument");
  if (out) std::memset(out, 0, sizeof(*out));
  nd->base_2_odom = tf_from(t, q);
  int init = 0;
  int rc = a3d_pf_is_initialized(nd->pf, &init);
  if (rc) return rc;
  if (!init) {
    if (nd->p.set_initial_pose) {
      const double it[3] = {nd->p.init_x, nd->p.init_y, nd->p.init_z};
      const double iq[4] = {0.0, 0.0, std::sin(nd->p.init_a * 0.5), std::cos(nd->p.init_a * 0.5)};
      rc = node_init_pose(nd, tf_from(it, iq), (float)nd->p.init_x_dev, (float)nd->p.init_y_dev,
                          (float)nd->p.init_z_dev, (float)nd->p.init_a_dev);
      if (rc) return rc;
    }
    return A3D_OK;
  }
  double yaw;
  tf_rpy(nd->base_2_odom, nd->roll, nd->pitch, yaw);
  if (!nd->is_odom) {
    nd->is_odom = true;
    nd->lastbase_2_world = tf_mul(nd->initodom_2_world, nd->base_2_odom);
    nd->lastodom_2_world = nd->initodom_2_world;
  }
  if (!nd->has_takenoff) {
    nd->has_takenoff = nd->base_2_odom.o[2] > nd->p.take_off_height;
    nd->lastbase_2_world = tf_mul(nd->initodom_2_world, nd->base_2_odom);
    nd->lastodom_2_world = nd->initodom_2_world;
    std::memcpy(nd->lastmean, nd->mean, sizeof(nd->mean));
  } else {
    for (int a = 0; a < 4; ++a)
      if (std::isnan(nd->mean[a]) || std::isinf(nd->mean[a])) nd->amcl_out = true;
    for (int a = 0; a < 4; ++a)
      if (std::fabs(nd->mean[a] - nd->lastmean[a]) > 1.) nd->amcl_out = true;
    if (!nd->amcl_out) {
      double qq[4];
      quat_from_rpy(nd->roll, nd->pitch, nd->mean[3], qq);
      const double mt[3] = {nd->mean[0], nd->mean[1], nd->mean[2]};
      Tf b2w = tf_from(mt, qq);
      b2w = tf_mul(tf_mul(b2w, tf_inv(nd->lastupdatebase_2_odom)), nd->base_2_odom);
      std::memcpy(nd->lastmean, nd->mean, sizeof(nd->mean));
      nd->lastbase_2_world = b2w;
      nd->lastodom_2_world = tf_mul(b2w, tf_inv(nd->base_2_odom));
      nd->amcl_out_lastbase_2_odom = nd->lastupdatebase_2_odom;
    } else {
      nd->lastbase_2_world =
          tf_mul(tf_mul(nd->lastbase_2_world, tf_inv(nd->amcl_out_lastbase_2_odom)), nd->base_2_odom);
      nd->amcl_out_lastbase_2_odom = nd->base_2_odom;
    }
  }
  if (out) {
    out->published = 1;
    tf_out(nd->lastbase_2_world, out->base_t, out->base_q);
    tf_out(nd->lastodom_2_world, out->odom_t, out->odom_q);
    out->amcl_out = nd->amcl_out ? 1 : 0;
    out->has_takenoff = nd->has_takenoff ? 1 : 0;
  }
  return A3D_OK;
}

int a3d_node_range(a3d_node* nd, float range, double ax, double ay, double az) {
  // Node::rangeCallback (Node.cpp:326-345): Range(float(range), x, y, z)
  if (!nd) return fail(A3D_ERR_ARG, "a3d_node_range: null node");
  nd->ranges.push_back(range);
  nd->ranges.push_back((float)ax);
  nd->ranges.push_back((float)ay);
  nd->ranges.push_back((float)az);
  return A3D_OK;
}

int a3d_node_point_cloud(a3d_node* nd, int64_t now_ns, const void* points, size_t n, size_t stride, int* updated) {
  // Node::pointcloudCallback (Node.cpp:112-191)
  if (!nd || (n && !points) || !updated) return fail(A3D_ERR_ARG, "a3d_node_point_cloud: null argument");
  if (n && (stride < 12 || stride % 4 != 0 || n > 0x7fffffffull))
    return fail(A3D_ERR_ARG, "a3d_node_point_cloud: stride must be >= 12 and a multiple of 4, n < 2^31");
  *updated = 0;
  if (!nd->is_odom) return A3D_OK;
  if (!update_due(nd, now_ns)) return A3D_OK;
  nd->nextupdate_ns = now_ns + duration_ns(1.0 / nd->p.update_rate);
  // the scan and the ranges to the device, then VoxelGrid there
  int rc = A3D_OK;
  size_t m = 0;
  const size_t nr = nd->ranges.size();
  if (n * stride > nd->raw_cap || n > nd->v_cap || nr > nd->rng_cap) {
    if (n * stride > nd->raw_cap) {
      (void)hipFree(nd->d_raw);
      nd->d_raw = nullptr;
      nd->raw_cap = 0;
      if (hipMalloc(&nd->d_raw, n * stride) != hipSuccess) return fail(A3D_ERR_NOMEM, "a3d_node_point_cloud: scan");
      nd->raw_cap = n * stride;
    }
    if (n > nd->v_cap) {
      (void)hipFree(nd->d_v);
      nd->d_v = nullptr;
      nd->v_cap = 0;
      if (hipMalloc(&nd->d_v, 3 * n * sizeof(float)) != hipSuccess)
        return fail(A3D_ERR_NOMEM, "a3d_node_point_cloud: voxels");
      nd->v_cap = n;
    }
    if (nr > nd->rng_cap) {
      (void)hipFree(nd->d_rng);
      nd->d_rng = nullptr;
      nd->rng_cap = 0;
      if (hipMalloc(&nd->d_rng, nr * sizeof(float)) != hipSuccess)
        return fail(A3D_ERR_NOMEM, "a3d_node_point_cloud: ranges");
      nd->rng_cap = nr;
    }
  }
  float* const d_vx = nd->d_v;
  float* const d_vy = nd->d_v ? nd->d_v + nd->v_cap : nullptr;
  float* const d_vz = nd->d_v ? nd->d_v + 2 * nd->v_cap : nullptr;
  if (n) {
    A3D_HIP(hipMemcpy(nd->d_raw, points, n * stride, hipMemcpyHostToDevice));
    rc = voxel_grid_device(nd->d_raw, n, stride, (float)nd->p.voxel_size, d_vx, d_vy, d_vz, &m, nullptr, nd->vws);
    if (rc) return rc;
  }
  if (nr) A3D_HIP(hipMemcpy(nd->d_rng, nd->ranges.data(), nr * sizeof(float), hipMemcpyHostToDevice));
  const Tf inc = tf_mul(tf_inv(nd->lastupdatebase_2_odom), nd->base_2_odom);
  nd->deltas[0] = inc.o[0];
  nd->deltas[1] = inc.o[1];
  nd->deltas[2] = inc.o[2];
  nd->deltas[3] = yaw_of(inc);
  size_t cur = 0;
  rc = a3d_pf_size(nd->pf, &cur);
  if (rc) return rc;
  if (cur > (size_t)nd->p.num_particles) return fail(A3D_ERR_STATE, "a3d_node_point_cloud: filter outgrew the node");
  // this step's draws, as a3d_session_draws gives them, made on the device
  // on the filter's stream, so predict and resample read them in order and
  // the host never waits for them
  const hipStream_t ps = pf_stream(nd->pf);
  nd->key[0] = nd->seed;
  nd->key[1] = nd->step;
  FilterDev f{};
  f.n = (uint32_t)cur;
  f.z4 = nd->d_z4;
  A3D_HIP(hipMemcpyAsync(nd->d_key, nd->key, 16, hipMemcpyHostToDevice, ps));
  launch_draws(single(f), nd->d_key, nd->d_key + 1, nd->d_u01, ps);
  A3D_HIP(hipGetLastError());
  ++nd->step;
  const double mods[4] = {nd->p.odom_x_mod, nd->p.odom_y_mod, nd->p.odom_z_mod, nd->p.odom_a_mod};
  rc = a3d_pf_predict_dev(nd->pf, mods, nd->deltas, nd->d_z4);
  if (rc) return rc;
  rc = a3d_pf_update_dev(nd->pf, nd->grid, d_vx, d_vy, d_vz, m, nd->d_rng, nr / 4, nd->p.alpha, nd->p.sensor_range,
                         nd->roll, nd->pitch);
  if (rc) return rc;
  rc = a3d_pf_get_mean(nd->pf, nd->mean);
  if (rc) return rc;
  nd->ranges.clear();
  nd->lastupdatebase_2_odom = nd->base_2_odom;
  if (++nd->n_updates > nd->p.resample_interval) {
    nd->n_updates = 0;
    // enqueued behind the update; the next callback's work waits for it on
    // the same stream
    rc = pf_resample_device_u01(nd->pf, nd->d_u01);
    if (rc) return rc;
  }
  *updated = 1;
  return A3D_OK;
}

int a3d_node_get_state(const a3d_node* nd, float mean[4], double deltas[4], double* roll, double* pitch,
                       int* is_odom, int* amcl_out, uint64_t* draw_step) {
  if (!nd) return fail(A3D_ERR_ARG, "a3d_node_get_state: null node");
  for (int a = 0; a < 4; ++a) {
    if (mean) mean[a] = nd->mean[a];
    if (deltas) deltas[a] = nd->deltas[a];
  }
  if (roll) *roll = nd->roll;
  if (pitch) *pitch = nd->pitch;
  if (is_odom) *is_odom = nd->is_odom ? 1 : 0;
  if (amcl_out) *amcl_out = nd->amcl_out ? 1 : 0;
  if (draw_step) *draw_step = nd->step;
  return A3D_OK;
}

int a3d_node_get_filter(a3d_node* nd, a3d_pf** pf) {
  if (!nd || !pf) return fail(A3D_ERR_ARG, "a3d_node_get_filter: null argument");
  *pf = nd->pf;
  return A3D_OK;
}

}  // extern "C"
