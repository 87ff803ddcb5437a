// C ABI of libamcl3d_gpu (include/amcl3d_gpu.h): handle management, host
// staging and the Grid3d / ParticleFilter / fleet entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "a3d_handles.h"
#include "a3d_math.h"

namespace a3d {

static thread_local std::string g_err;
static int g_ppw = 16;
static int g_chunk = 64;  // points per LDS cloud tile (particle-lane kernels)
static int g_deal = 0;    // work-list dealing of persistent launches
static int g_mapping = A3D_MAP_PARTICLE_LANES;
static int g_order = A3D_ORDER_SPATIAL;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int current_ppw() { return g_ppw; }
int current_chunk() { return g_chunk; }
int current_deal() { return g_deal; }
int current_mapping() { return g_mapping; }
int current_order() { return g_order; }

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

// Whether the device's fast cell index along one axis, trunc(q * inv_res),
// equals the reference's floor(q / res) (correctly rounded division) for
// every float q >= 0 whose cell is below size.  The two products are within
// 2^-52 relative of q/res, so they can straddle an integer n only when
// |q - n*res| <= q * 2^-51.9, far below the spacing of floats around q: q is
// then one of the floats bracketing n*res, which are among the float nearest
// to the double n*res and its two neighbours.  Checking those three floats for
// n = 1..size covers every q; longer axes are not checked (reported unclean).
static bool axis_clean(double res, double inv_res, uint32_t size) {
  if (size > (1u << 24) || !(res > 1e-30)) return false;
  for (uint32_t n = 1; n <= size; ++n) {
    const float p = (float)((double)n * res);
    const float c[3] = {std::nextafter(p, -INFINITY), p, std::nextafter(p, INFINITY)};
    for (float q : c) {
      if (!(q >= 0.0f)) continue;
      const double qd = (double)q;
      if (std::trunc(qd * inv_res) != std::floor(qd / res)) return false;
    }
  }
  return true;
}

// Grid metadata exactly as PointCloudTools.cpp:93-101 and Grid3d.cpp:151-153.
void grid_meta(const double mn[3], const double mx[3], double res, GridDev& g) {
  for (int a = 0; a < 3; ++a) {
    g.min[a] = mn[a];
    g.max[a] = mx[a];
    g.ext[a] = mx[a] - mn[a];
    g.size[a] = (uint32_t)std::ceil(g.ext[a] / res);
    float f = (float)g.ext[a];
    if ((double)f < g.ext[a]) f = std::nextafter(f, INFINITY);
    g.ext_hi[a] = f;
  }
  g.res = res;
  g.inv_res = 1.0 / res;
  for (int a = 0; a < 3; ++a) {
    // t = the smallest non-negative float with floor((double)t / res) >= size
    // (bit patterns of non-negative floats are ordered like their values)
    auto past = [&](uint32_t b) {
      float f;
      std::memcpy(&f, &b, 4);
      return std::floor((double)f / res) >= (double)g.size[a];
    };
    uint32_t lo = 0, hi = 0x7f800000u;  // +inf is past
    if (past(0)) hi = 0;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      (past(mid) ? hi : lo) = mid;
    }
    uint32_t eb;
    std::memcpy(&eb, &g.ext_hi[a], 4);
    g.lim[a] = std::min(eb, hi);
    g.near_w[a] = std::max(1e-9, std::ldexp((double)g.size[a], -48));
    g.near_w2[a] = 2.0 * g.near_w[a];
  }
  g.clean = (axis_clean(res, g.inv_res, g.size[0]) && axis_clean(res, g.inv_res, g.size[1]) &&
             axis_clean(res, g.inv_res, g.size[2])) ? 1u : 0u;
  // Test knob: A3D_FORCE_NEAR_TEST=1 keeps the window test on any map.
  const char* near_test = std::getenv("A3D_FORCE_NEAR_TEST");
  if (near_test && near_test[0] == '1') g.clean = 0u;
  g.step_z = (uint64_t)g.size[0] * g.size[1];
  g.cells = g.step_z * g.size[2];
  // 32-bit cell indices use 24-bit multiply-adds: (iy + sy * iz) * sx + ix
  // needs sx < 2^24 and sy * sz < 2^24, and the zero cell after the last
  // (index cells) must fit a 32-bit record count; other grids take the
  // 64-bit path.
  const bool mad24 = g.size[0] < (1u << 24) && (uint64_t)g.size[1] * g.size[2] < (1ull << 24);
  g.idx64 = (g.cells >= 0xffffffffull || !mad24) ? 1u : 0u;
  // A structured buffer forms record * 4 in 32 bits (modulo 2^32): a 2^30-cell
  // segment of any grid whose indices the multiply-adds form is one buffer.
  g.seg_bits = mad24 ? 30u : 0u;
  // Test knobs: A3D_FORCE_IDX64=1 takes the 64-bit index path on any map
  // (=2: the same kernels, with their per-wavefront segment loads).
  const char* force = std::getenv("A3D_FORCE_IDX64");
  if (force && (force[0] == '1' || force[0] == '2')) {
    g.idx64 = 1u;
    if (force[0] == '1') g.seg_bits = 0u;
  }
  // A structured buffer based at prob: only a prob array whose bytes, the
  // zero cell included, fit 2^32 may be read through one.
  g.buf32 = (!g.idx64 && (g.cells + 1) * 4 <= 0x100000000ull) ? 1u : 0u;
  // Test knob: A3D_FORCE_IDX32=1 reads 32-bit indices through 64-bit
  // addresses on any map that has 32-bit indices (=2: the same kernels, with
  // their per-wavefront segment loads).
  const char* force32 = std::getenv("A3D_FORCE_IDX32");
  if (force32 && (force32[0] == '1' || force32[0] == '2')) {
    g.buf32 = 0u;
    if (force32[0] == '1') g.seg_bits = 0u;
  }
  // Test knob: A3D_SEG_TEST_BITS=b (8..29) lets a wavefront take the segment
  // loads only when its cells lie in one 2^b-cell block, so that small maps
  // mix wavefronts of both forms; b = 0: no segment loads.
  const char* segb = std::getenv("A3D_SEG_TEST_BITS");
  if (segb && g.seg_bits) {
    const int b = std::atoi(segb);
    if (b == 0 || (b >= 8 && b < 30)) g.seg_bits = (uint32_t)b;
  }
}

// Bytes of a grid's prob allocation: the cells and the zero cell, rounded up
// to whole 2^30-cell segments when segment loads may read it (a segment's
// buffer then never reaches past the allocation).
inline size_t prob_alloc_bytes(const GridDev& g) {
  const uint64_t b = (g.cells + 1) * sizeof(float);
  const uint64_t seg = 1ull << 32;
  if (g.seg_bits && g.cells + 1 > (1ull << 30)) return (size_t)((b + seg - 1) / seg * seg);
  return (size_t)b;
}

}  // namespace a3d

using namespace a3d;

// ------------------------------------------------------------------ handles
struct a3d_pf {
  hipStream_t stream = nullptr;
  uint32_t n = 0;
  float* x = nullptr;  // one allocation: x y z a w wp wr
  float* scratch = nullptr;
  uint8_t* inmap = nullptr;
  float* mean = nullptr;
  uint32_t* idx = nullptr;
  void* ord = nullptr;     // particle visiting order buffers (order_view)
  float* stage = nullptr;  // host-input staging
  size_t stage_cap = 0;
  bool initialized = false;
  bool capturing = false;
  bool prefix_valid = false;  // the last update left resample's running sum c
  TimingPool timing;

  FilterDev dev() const {
    FilterDev f{};
    f.x = x;
    f.y = x + n;
    f.z = x + 2 * (size_t)n;
    f.a = x + 3 * (size_t)n;
    f.w = x + 4 * (size_t)n;
    f.wp = x + 5 * (size_t)n;
    f.wr = x + 6 * (size_t)n;
    f.scratch = scratch;
    f.zr = scratch ? scratch + sc_sums(n) + kSumZr : nullptr;
    f.inmap = inmap;
    f.mean = mean;
    f.n = n;
    return f;
  }
};


// -------------------------------------------------------------- utilities
static int ensure_stage(float*& buf, size_t& cap, size_t floats) {
  if (floats <= cap) return A3D_OK;
  (void)hipFree(buf);
  buf = nullptr;
  cap = 0;
  const size_t want = std::max(floats, (size_t)1 << 16);
  A3D_HIP(hipMalloc(&buf, want * sizeof(float)));
  cap = want;
  return A3D_OK;
}

static bool finite_pos(double v) { return std::isfinite(v) && v > 0; }

extern "C" {

const char* a3d_last_error(void) { return g_err.c_str(); }

const char* a3d_version(void) { return "amcl3d_gpu 0.1 (gfx950)"; }

int a3d_set_device(int device) {
  A3D_HIP(hipSetDevice(device));
  return A3D_OK;
}

int a3d_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes) {
  int dev = 0;
  A3D_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  A3D_HIP(hipGetDeviceProperties(&p, dev));
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
  return A3D_OK;
}

int a3d_device_synchronize(void) {
  A3D_HIP(hipDeviceSynchronize());
  return A3D_OK;
}

int a3d_buffer_alloc(size_t bytes, void** dptr) {
  if (!dptr || bytes == 0) return fail(A3D_ERR_ARG, "a3d_buffer_alloc: null output or zero size");
  *dptr = nullptr;
  if (hipMalloc(dptr, bytes) != hipSuccess) return fail(A3D_ERR_NOMEM, "a3d_buffer_alloc: hipMalloc failed");
  return A3D_OK;
}

int a3d_buffer_free(void* dptr) {
  A3D_HIP(hipFree(dptr));
  return A3D_OK;
}

int a3d_buffer_upload(void* dptr, const void* host, size_t bytes) {
  if (!dptr || (!host && bytes)) return fail(A3D_ERR_ARG, "a3d_buffer_upload: null pointer");
  A3D_HIP(hipMemcpy(dptr, host, bytes, hipMemcpyHostToDevice));
  return A3D_OK;
}

int a3d_buffer_download(void* host, const void* dptr, size_t bytes) {
  if (!dptr || (!host && bytes)) return fail(A3D_ERR_ARG, "a3d_buffer_download: null pointer");
  A3D_HIP(hipMemcpy(host, dptr, bytes, hipMemcpyDeviceToHost));
  return A3D_OK;
}

int a3d_buffer_copy(void* dst, const void* src, size_t bytes) {
  if (!dst || !src) return fail(A3D_ERR_ARG, "a3d_buffer_copy: null pointer");
  A3D_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return A3D_OK;
}

static int voxel_args(size_t n, size_t stride, float leaf, const void* pts, const float* x, const float* y,
                      const float* z, size_t* out_n) {
  if (!out_n || (n && (!pts || !x || !y || !z))) return fail(A3D_ERR_ARG, "a3d_voxel_grid: null argument");
  if (stride < 12 || stride % 4 != 0) return fail(A3D_ERR_ARG, "a3d_voxel_grid: stride must be >= 12, a multiple of 4");
  if (!(leaf > 0.0f) || !std::isfinite(leaf)) return fail(A3D_ERR_ARG, "a3d_voxel_grid: leaf must be positive");
  if (n > 0x7fffffffull) return fail(A3D_ERR_ARG, "a3d_voxel_grid: too many points");
  return A3D_OK;
}

int a3d_voxel_grid_dev(const void* d_points, size_t n, size_t stride, float leaf, float* d_x, float* d_y, float* d_z,
                       size_t* out_n) {
  int rc = voxel_args(n, stride, leaf, d_points, d_x, d_y, d_z, out_n);
  if (rc) return rc;
  return voxel_grid_device(d_points, n, stride, leaf, d_x, d_y, d_z, out_n, nullptr);
}

int a3d_voxel_grid(const void* points, size_t n, size_t stride, float leaf, float* out_x, float* out_y, float* out_z,
                   size_t* out_n) {
  int rc = voxel_args(n, stride, leaf, points, out_x, out_y, out_z, out_n);
  if (rc) return rc;
  *out_n = 0;
  if (n == 0) return A3D_OK;
  void* d_in = nullptr;
  float* d_out = nullptr;
  if (hipMalloc(&d_in, n * stride) != hipSuccess || hipMalloc(&d_out, 3 * n * 4) != hipSuccess) {
    (void)hipFree(d_in);
    return fail(A3D_ERR_NOMEM, "a3d_voxel_grid: device allocation failed");
  }
  size_t k = 0;
  rc = hip_ok(hipMemcpy(d_in, points, n * stride, hipMemcpyHostToDevice), "a3d_voxel_grid: upload") ? A3D_OK
                                                                                                    : A3D_ERR_HIP;
  if (!rc) rc = voxel_grid_device(d_in, n, stride, leaf, d_out, d_out + n, d_out + 2 * n, &k, nullptr);
  if (!rc && k) {
    const bool ok = hipMemcpy(out_x, d_out, k * 4, hipMemcpyDeviceToHost) == hipSuccess &&
                    hipMemcpy(out_y, d_out + n, k * 4, hipMemcpyDeviceToHost) == hipSuccess &&
                    hipMemcpy(out_z, d_out + 2 * n, k * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) rc = fail(A3D_ERR_HIP, "a3d_voxel_grid: download failed");
  }
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  if (!rc) *out_n = k;
  return rc;
}

int a3d_set_tile(int particles_per_wave) {
  if (particles_per_wave != 16 && particles_per_wave != 32 && particles_per_wave != 64)
    return fail(A3D_ERR_ARG, "a3d_set_tile: particles per wavefront must be 16, 32 or 64");
  g_ppw = particles_per_wave;
  return A3D_OK;
}

int a3d_get_tile(int* particles_per_wave) {
  if (!particles_per_wave) return fail(A3D_ERR_ARG, "a3d_get_tile: null output");
  *particles_per_wave = g_ppw;
  return A3D_OK;
}

int a3d_set_cloud_tile(int points) {
  if (points != 32 && points != 64 && points != 128 && points != 256)
    return fail(A3D_ERR_ARG, "a3d_set_cloud_tile: points per tile must be 32, 64, 128 or 256");
  g_chunk = points;
  return A3D_OK;
}

int a3d_get_cloud_tile(int* points) {
  if (!points) return fail(A3D_ERR_ARG, "a3d_get_cloud_tile: null output");
  *points = g_chunk;
  return A3D_OK;
}

int a3d_set_deal(int run) {
  if (run < 0 || run > 4096) return fail(A3D_ERR_ARG, "a3d_set_deal: run length must be 0 (contiguous) to 4096");
  g_deal = run;
  return A3D_OK;
}

int a3d_get_deal(int* run) {
  if (!run) return fail(A3D_ERR_ARG, "a3d_get_deal: null output");
  *run = g_deal;
  return A3D_OK;
}

int a3d_set_mapping(int mapping) {
  if (mapping != A3D_MAP_PARTICLE_LANES && mapping != A3D_MAP_POINT_LANES)
    return fail(A3D_ERR_ARG, "a3d_set_mapping: unknown mapping");
  g_mapping = mapping;
  return A3D_OK;
}

int a3d_set_order(int order) {
  if (order != A3D_ORDER_INDEX && order != A3D_ORDER_SPATIAL) return fail(A3D_ERR_ARG, "a3d_set_order: unknown order");
  g_order = order;
  return A3D_OK;
}

int a3d_get_order(int* order) {
  if (!order) return fail(A3D_ERR_ARG, "a3d_get_order: null output");
  *order = g_order;
  return A3D_OK;
}

int a3d_get_mapping(int* mapping) {
  if (!mapping) return fail(A3D_ERR_ARG, "a3d_get_mapping: null output");
  *mapping = g_mapping;
  return A3D_OK;
}

// ------------------------------------------------------------------- grids
static int grid_alloc(const double mn[3], const double mx[3], double res, double sensor_dev, bool keep_dist,
                      a3d_grid** out) {
  if (!out || !mn || !mx) return fail(A3D_ERR_ARG, "grid: null argument");
  *out = nullptr;
  if (!finite_pos(res)) return fail(A3D_ERR_ARG, "grid: resolution must be positive");
  for (int a = 0; a < 3; ++a)
    if (!(std::isfinite(mn[a]) && std::isfinite(mx[a]) && mx[a] > mn[a]))
      return fail(A3D_ERR_ARG, "grid: bounds must be finite with max > min");
  if (!(std::isfinite(sensor_dev) && sensor_dev > 0)) return fail(A3D_ERR_ARG, "grid: sensor_dev must be positive");
  for (int a = 0; a < 3; ++a)  // cell coordinates are compared as int32 on the device
    if (!(std::ceil((mx[a] - mn[a]) / res) < 2147483648.0))
      return fail(A3D_ERR_ARG, "grid: 2^31 or more cells along an axis");
  a3d_grid* G = new a3d_grid();
  grid_meta(mn, mx, res, G->g);
  G->sensor_dev = sensor_dev;
  const size_t bytes = G->g.cells * sizeof(float);
  // one more float, 0, after the last cell: the cell that rejected points read
  if (hipMalloc(&G->prob, prob_alloc_bytes(G->g)) != hipSuccess) {
    delete G;
    return fail(A3D_ERR_NOMEM, "grid: cannot allocate " + std::to_string(bytes) + " bytes of prob");
  }
  if (hipMemset(G->prob + G->g.cells, 0, sizeof(float)) != hipSuccess) {
    (void)hipFree(G->prob);
    delete G;
    return fail(A3D_ERR_HIP, "grid: cannot clear the zero cell");
  }
  if (keep_dist && hipMalloc(&G->dist, bytes) != hipSuccess) {
    (void)hipFree(G->prob);
    delete G;
    return fail(A3D_ERR_NOMEM, "grid: cannot allocate the distance field");
  }
  if (hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipFree(G->prob);
    (void)hipFree(G->dist);
    delete G;
    return fail(A3D_ERR_HIP, "grid: cannot create a stream");
  }
  G->g.prob = G->prob;
  G->g.dist = G->dist;
  *out = G;
  return A3D_OK;
}

int a3d_grid_create_from_points(const float* xyz, size_t n, const double min[3], const double max[3], double res,
                                double sensor_dev, int keep_dist, a3d_grid** out) {
  if (n && !xyz) return fail(A3D_ERR_ARG, "a3d_grid_create_from_points: null points");
  a3d_grid* G = nullptr;
  int rc = grid_alloc(min, max, res, sensor_dev, keep_dist != 0, &G);
  if (rc) return rc;
  float* d_xyz = nullptr;
  if (n) {
    if (hipMalloc(&d_xyz, n * 3 * sizeof(float)) != hipSuccess) {
      a3d_grid_destroy(G);
      return fail(A3D_ERR_NOMEM, "a3d_grid_create_from_points: cannot stage points");
    }
    if (!hip_ok(hipMemcpy(d_xyz, xyz, n * 3 * sizeof(float), hipMemcpyHostToDevice), "upload points")) {
      (void)hipFree(d_xyz);
      a3d_grid_destroy(G);
      return A3D_ERR_HIP;
    }
  }
  rc = build_grid_device(d_xyz, n, G->g, G->prob, G->dist, sensor_dev, G->stream);
  if (rc == A3D_OK && !hip_ok(hipStreamSynchronize(G->stream), "computeGrid")) rc = A3D_ERR_HIP;
  (void)hipFree(d_xyz);
  if (rc) {
    a3d_grid_destroy(G);
    return rc;
  }
  *out = G;
  return A3D_OK;
}

int a3d_grid_create_from_cells(const float* prob, const float* dist, const uint32_t size[3], const double min[3],
                               const double max[3], double res, double sensor_dev, a3d_grid** out) {
  if (!prob || !size) return fail(A3D_ERR_ARG, "a3d_grid_create_from_cells: null argument");
  a3d_grid* G = nullptr;
  int rc = grid_alloc(min, max, res, sensor_dev, dist != nullptr, &G);
  if (rc) return rc;
  for (int a = 0; a < 3; ++a)
    if (size[a] != G->g.size[a]) {
      a3d_grid_destroy(G);
      return fail(A3D_ERR_ARG, "a3d_grid_create_from_cells: size does not match ceil((max-min)/res)");
    }
  const size_t bytes = G->g.cells * sizeof(float);
  if (!hip_ok(hipMemcpy(G->prob, prob, bytes, hipMemcpyHostToDevice), "upload prob") ||
      (dist && !hip_ok(hipMemcpy(G->dist, dist, bytes, hipMemcpyHostToDevice), "upload dist"))) {
    a3d_grid_destroy(G);
    return A3D_ERR_HIP;
  }
  *out = G;
  return A3D_OK;
}

int a3d_grid_destroy(a3d_grid* g) {
  if (!g) return A3D_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  (void)hipFree(g->prob);
  (void)hipFree(g->dist);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return A3D_OK;
}

int a3d_grid_get_info(const a3d_grid* g, uint32_t size[3], double min[3], double max[3], double* res,
                      double* sensor_dev) {
  if (!g) return fail(A3D_ERR_ARG, "a3d_grid_get_info: null grid");
  for (int a = 0; a < 3; ++a) {
    if (size) size[a] = g->g.size[a];
    if (min) min[a] = g->g.min[a];
    if (max) max[a] = g->g.max[a];
  }
  if (res) *res = g->g.res;
  if (sensor_dev) *sensor_dev = g->sensor_dev;
  return A3D_OK;
}

int a3d_grid_read_cells(const a3d_grid* g, uint64_t first, uint64_t count, float* prob, float* dist) {
  if (!g) return fail(A3D_ERR_ARG, "a3d_grid_read_cells: null grid");
  if (first > g->g.cells || count > g->g.cells - first) return fail(A3D_ERR_ARG, "a3d_grid_read_cells: range");
  if (dist && !g->dist) return fail(A3D_ERR_STATE, "a3d_grid_read_cells: distance field not kept");
  if (prob) A3D_HIP(hipMemcpy(prob, g->prob + first, count * sizeof(float), hipMemcpyDeviceToHost));
  if (dist) A3D_HIP(hipMemcpy(dist, g->dist + first, count * sizeof(float), hipMemcpyDeviceToHost));
  return A3D_OK;
}

int a3d_grid_is_into_map(const a3d_grid* g, float x, float y, float z, int* inside) {
  if (!inside) return fail(A3D_ERR_ARG, "a3d_grid_is_into_map: null output");
  *inside = (g && into_map(g->g, x, y, z)) ? 1 : 0;  // Grid3d.cpp:203-204: no map -> false
  return A3D_OK;
}

int a3d_grid_cloud_weight(const a3d_grid* g, const float* x, const float* y, const float* z, size_t m,
                          const float* tx, const float* ty, const float* tz, const float* yaw, size_t n, float roll,
                          float pitch, float* out) {
  if (!out || (n && (!tx || !ty || !tz || !yaw)) || (m && (!x || !y || !z)))
    return fail(A3D_ERR_ARG, "a3d_grid_cloud_weight: null argument");
  if (n == 0) return A3D_OK;
  if (!g) {  // Grid3d.cpp:136-137: no map -> 0
    std::fill(out, out + n, 0.f);
    return A3D_OK;
  }
  if (m > 0xffffffffull || n > 0xffffffffull) return fail(A3D_ERR_ARG, "a3d_grid_cloud_weight: too large");
  float* buf = nullptr;
  const size_t floats = 3 * m + 5 * n + 64;
  if (hipMalloc(&buf, floats * sizeof(float)) != hipSuccess)
    return fail(A3D_ERR_NOMEM, "a3d_grid_cloud_weight: staging");
  float* cx = buf;
  float* cy = cx + m;
  float* cz = cy + m;
  float* dtx = cz + m;
  float* dty = dtx + n;
  float* dtz = dty + n;
  float* dya = dtz + n;
  float* dout = dya + n;
  float* dzr = dout + n;  // 2 * kZrParts of the 64 spare floats
  int rc = A3D_OK;
  hipStream_t s = g->stream;
  if (m) {
    if (!hip_ok(hipMemcpyAsync(cx, x, m * 4, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(hipMemcpyAsync(cy, y, m * 4, hipMemcpyHostToDevice, s), "upload") ||
        !hip_ok(hipMemcpyAsync(cz, z, m * 4, hipMemcpyHostToDevice, s), "upload"))
      rc = A3D_ERR_HIP;
  }
  if (!rc && (!hip_ok(hipMemcpyAsync(dtx, tx, n * 4, hipMemcpyHostToDevice, s), "upload") ||
              !hip_ok(hipMemcpyAsync(dty, ty, n * 4, hipMemcpyHostToDevice, s), "upload") ||
              !hip_ok(hipMemcpyAsync(dtz, tz, n * 4, hipMemcpyHostToDevice, s), "upload") ||
              !hip_ok(hipMemcpyAsync(dya, yaw, n * 4, hipMemcpyHostToDevice, s), "upload")))
    rc = A3D_ERR_HIP;
  if (!rc) {
    launch_weight_poses(g->g, cx, cy, cz, (uint32_t)m, dtx, dty, dtz, dya, (uint32_t)n, roll, pitch, dout, dzr, g_ppw,
                        s);
    if (!hip_ok(hipGetLastError(), "cloud weight kernel") ||
        !hip_ok(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, s), "download") ||
        !hip_ok(hipStreamSynchronize(s), "cloud weight"))
      rc = A3D_ERR_HIP;
  }
  (void)hipFree(buf);
  return rc;
}

// .grid file, Grid3d.cpp:210-275
int a3d_grid_save(const a3d_grid* g, const char* path) {
  if (!g || !path) return fail(A3D_ERR_ARG, "a3d_grid_save: null argument");
  if (!g->dist) return fail(A3D_ERR_STATE, "a3d_grid_save: the .grid format stores dist; build with keep_dist");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(A3D_ERR_IO, std::string("a3d_grid_save: cannot open ") + path);
  bool ok = std::fwrite(&g->g.size[0], 4, 1, f) == 1 && std::fwrite(&g->g.size[1], 4, 1, f) == 1 &&
            std::fwrite(&g->g.size[2], 4, 1, f) == 1 && std::fwrite(&g->sensor_dev, 8, 1, f) == 1;
  const uint64_t chunk = 1ull << 22;
  std::vector<float> pb(chunk), db(chunk), cells(2 * chunk);
  for (uint64_t o = 0; ok && o < g->g.cells; o += chunk) {
    const uint64_t c = std::min(chunk, g->g.cells - o);
    if (hipMemcpy(pb.data(), g->prob + o, c * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(db.data(), g->dist + o, c * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      std::fclose(f);
      return fail(A3D_ERR_HIP, "a3d_grid_save: download failed");
    }
    for (uint64_t i = 0; i < c; ++i) {
      cells[2 * i] = db[i];
      cells[2 * i + 1] = pb[i];
    }
    ok = std::fwrite(cells.data(), 8, c, f) == c;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? A3D_OK : fail(A3D_ERR_IO, std::string("a3d_grid_save: write failed: ") + path);
}

int a3d_grid_load(const char* path, const double min[3], const double max[3], double res, double sensor_dev,
                  a3d_grid** out) {
  if (!path || !out) return fail(A3D_ERR_ARG, "a3d_grid_load: null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(A3D_ERR_IO, std::string("a3d_grid_load: cannot open ") + path);
  uint32_t sz[3];
  double sd = 0;
  if (std::fread(sz, 4, 3, f) != 3 || std::fread(&sd, 8, 1, f) != 1) {
    std::fclose(f);
    return fail(A3D_ERR_IO, "a3d_grid_load: truncated header");
  }
  if (std::fabs(sd - sensor_dev) >= DBL_EPSILON) {  // Grid3d.cpp:256-260
    std::fclose(f);
    return fail(A3D_ERR_STATE, "a3d_grid_load: sensor_dev of the file differs");
  }
  a3d_grid* G = nullptr;
  int rc = grid_alloc(min, max, res, sensor_dev, true, &G);
  if (rc) {
    std::fclose(f);
    return rc;
  }
  for (int a = 0; a < 3; ++a)
    if (sz[a] != G->g.size[a]) {
      std::fclose(f);
      a3d_grid_destroy(G);
      return fail(A3D_ERR_ARG, "a3d_grid_load: file sizes do not match the bounds and resolution");
    }
  const uint64_t chunk = 1ull << 22;
  std::vector<float> pb(chunk), db(chunk), cells(2 * chunk);
  for (uint64_t o = 0; o < G->g.cells; o += chunk) {
    const uint64_t c = std::min(chunk, G->g.cells - o);
    if (std::fread(cells.data(), 8, c, f) != c) {
      std::fclose(f);
      a3d_grid_destroy(G);
      return fail(A3D_ERR_IO, "a3d_grid_load: truncated cells");
    }
    for (uint64_t i = 0; i < c; ++i) {
      db[i] = cells[2 * i];
      pb[i] = cells[2 * i + 1];
    }
    if (hipMemcpy(G->prob + o, pb.data(), c * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(G->dist + o, db.data(), c * 4, hipMemcpyHostToDevice) != hipSuccess) {
      std::fclose(f);
      a3d_grid_destroy(G);
      return fail(A3D_ERR_HIP, "a3d_grid_load: upload failed");
    }
  }
  std::fclose(f);
  *out = G;
  return A3D_OK;
}

// ---------------------------------------------------------- ParticleFilter
int a3d_pf_create(a3d_pf** out) {
  if (!out) return fail(A3D_ERR_ARG, "a3d_pf_create: null output");
  a3d_pf* p = new a3d_pf();
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return fail(A3D_ERR_HIP, "a3d_pf_create: cannot create a stream");
  }
  *out = p;
  return A3D_OK;
}

static void pf_free_state(a3d_pf* p) {
  (void)hipFree(p->x);
  (void)hipFree(p->scratch);
  (void)hipFree(p->inmap);
  (void)hipFree(p->mean);
  (void)hipFree(p->idx);
  (void)hipFree(p->ord);
  p->x = p->scratch = p->mean = nullptr;
  p->inmap = nullptr;
  p->idx = nullptr;
  p->ord = nullptr;
  p->n = 0;
}

int a3d_pf_destroy(a3d_pf* p) {
  if (!p) return A3D_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  p->timing.destroy();
  pf_free_state(p);
  (void)hipFree(p->stage);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return A3D_OK;
}

// Resizes the particle set like std::vector::resize: surviving particles keep
// their values, new ones are zero (Particle() constructor).
static int pf_resize(a3d_pf* p, uint32_t n) {
  if (n == p->n && p->x) return A3D_OK;
  float* x = nullptr;
  float* sc = nullptr;
  uint8_t* in = nullptr;
  float* mean = nullptr;
  uint32_t* idx = nullptr;
  void* ord = nullptr;
  if (hipMalloc(&x, 7 * (size_t)n * 4) != hipSuccess || hipMalloc(&sc, scratch_floats(n) * 4) != hipSuccess ||
      hipMalloc(&in, n) != hipSuccess || hipMalloc(&mean, 4 * 4) != hipSuccess ||
      hipMalloc(&idx, (size_t)n * 4) != hipSuccess || hipMalloc(&ord, order_bytes(n)) != hipSuccess) {
    (void)hipFree(x);
    (void)hipFree(sc);
    (void)hipFree(in);
    (void)hipFree(mean);
    (void)hipFree(idx);
    (void)hipFree(ord);
    return fail(A3D_ERR_NOMEM, "particle filter: cannot allocate particle state");
  }
  A3D_HIP(hipMemsetAsync(x, 0, 7 * (size_t)n * 4, p->stream));
  A3D_HIP(hipMemsetAsync(mean, 0, 16, p->stream));
  const uint32_t keep = std::min(n, p->n);
  for (int k = 0; k < 7 && keep; ++k)
    A3D_HIP(hipMemcpyAsync(x + (size_t)k * n, p->x + (size_t)k * p->n, (size_t)keep * 4, hipMemcpyDeviceToDevice,
                           p->stream));
  A3D_HIP(hipStreamSynchronize(p->stream));
  pf_free_state(p);
  p->x = x;
  p->scratch = sc;
  p->inmap = in;
  p->mean = mean;
  p->idx = idx;
  p->ord = ord;
  p->n = n;
  return A3D_OK;
}

static int pf_stage(a3d_pf* p, size_t floats) { return ensure_stage(p->stage, p->stage_cap, floats); }

int a3d_pf_init(a3d_pf* p, int num_particles, float x_init, float y_init, float z_init, float a_init, float x_dev,
                float y_dev, float z_dev, float a_dev, const float* z4) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_init: null filter");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_init: not allowed during capture");
  const uint32_t n = (uint32_t)std::abs(num_particles);  // p_.resize(abs(num_particles))
  if (n == 0) return fail(A3D_ERR_ARG, "a3d_pf_init: num_particles must be non-zero (the reference indexes p_[0])");
  if (n > 1 && !z4) return fail(A3D_ERR_ARG, "a3d_pf_init: null draws");
  int rc = pf_resize(p, n);
  if (rc) return rc;
  rc = pf_stage(p, 4 * (size_t)n);
  if (rc) return rc;
  if (n > 1) A3D_HIP(hipMemcpyAsync(p->stage, z4, 4 * (size_t)n * 4, hipMemcpyHostToDevice, p->stream));
  FilterDev f = p->dev();
  // ParticleFilter.cpp:54-56
  const float dev = std::max(std::max(x_dev, y_dev), z_dev);
  const float gc1 = (float)(1. / (dev * std::sqrt(2 * M_PI)));
  const float gc2 = (float)(1. / (2 * dev * dev));
  f.dev[0] = gc1;
  f.dev[1] = gc2;
  f.delta[0] = x_dev;
  f.delta[1] = y_dev;
  f.delta[2] = z_dev;
  f.delta[3] = a_dev;
  f.pose[0] = x_init;
  f.pose[1] = y_init;
  f.pose[2] = z_init;
  f.pose[3] = a_init;
  f.z4 = p->stage;
  launch_init(single(f), p->stream);
  p->prefix_valid = false;
  A3D_HIP(hipGetLastError());
  A3D_HIP(hipStreamSynchronize(p->stream));
  p->initialized = true;
  return A3D_OK;
}

int a3d_pf_predict_dev(a3d_pf* p, const double mods[4], const double deltas[4], const float* d_z4) {
  if (!p || !mods || !deltas || !d_z4) return fail(A3D_ERR_ARG, "a3d_pf_predict: null argument");
  if (!p->n) return fail(A3D_ERR_STATE, "a3d_pf_predict: filter has no particles (call init)");
  FilterDev f = p->dev();
  for (int k = 0; k < 4; ++k) {
    f.delta[k] = deltas[k];
    f.dev[k] = std::fabs(deltas[k] * mods[k]);  // ParticleFilter.cpp:101-104
  }
  f.z4 = d_z4;
  launch_predict(single(f), p->stream);
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_pf_predict(a3d_pf* p, double odom_x_mod, double odom_y_mod, double odom_z_mod, double odom_a_mod,
                   double delta_x, double delta_y, double delta_z, double delta_a, const float* z4) {
  if (!p || !z4) return fail(A3D_ERR_ARG, "a3d_pf_predict: null argument");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_predict: use the _dev form during capture");
  if (!p->n) return fail(A3D_ERR_STATE, "a3d_pf_predict: filter has no particles (call init)");
  int rc = pf_stage(p, 4 * (size_t)p->n);
  if (rc) return rc;
  A3D_HIP(hipMemcpyAsync(p->stage, z4, 4 * (size_t)p->n * 4, hipMemcpyHostToDevice, p->stream));
  const double mods[4] = {odom_x_mod, odom_y_mod, odom_z_mod, odom_a_mod};
  const double deltas[4] = {delta_x, delta_y, delta_z, delta_a};
  rc = a3d_pf_predict_dev(p, mods, deltas, p->stage);
  if (rc) return rc;
  A3D_HIP(hipStreamSynchronize(p->stream));
  return A3D_OK;
}

int a3d_pf_update_dev(a3d_pf* p, const a3d_grid* g, const float* d_x, const float* d_y, const float* d_z,
                      size_t m, const float* d_ranges, size_t nb, double alpha, double sigma, double roll,
                      double pitch) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_update: null filter");
  if (!p->n) return fail(A3D_ERR_STATE, "a3d_pf_update: filter has no particles (call init)");
  if (!g) return fail(A3D_ERR_ARG, "a3d_pf_update: null grid");
  if (m && (!d_x || !d_y || !d_z)) return fail(A3D_ERR_ARG, "a3d_pf_update: null cloud");
  if (nb && !d_ranges) return fail(A3D_ERR_ARG, "a3d_pf_update: null ranges");
  if (m > 0xffffffffull || nb > 0xffffffffull) return fail(A3D_ERR_ARG, "a3d_pf_update: too large");
  FilterDev f = p->dev();
  f.cx = d_x;
  f.cy = d_y;
  f.cz = d_z;
  f.m = (uint32_t)m;
  f.ranges = d_ranges;
  f.nb = (uint32_t)nb;
  f.alpha = alpha;
  f.sigma = sigma;
  f.roll = (float)roll;  // double -> float parameter of computeCloudWeight
  f.pitch = (float)pitch;
  f.rp0 = (f.roll == 0.0f && f.pitch == 0.0f) ? 1u : 0u;
  if (g_order == A3D_ORDER_SPATIAL) {
    const OrderDev o = order_view(p->ord, p->n);
    launch_order(f, o, p->stream);
    f.order = o.order;
  }
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (!p->capturing) ev = p->timing.take();
  if (ev.first) (void)hipEventRecord(ev.first, p->stream);
  launch_weight(g->g, single(f), nullptr, 0, nullptr, g_ppw, 0, p->stream);
  if (ev.first) {
    (void)hipEventRecord(ev.second, p->stream);
    p->timing.pending.push_back(ev);
    p->timing.launches++;
    p->timing.evals += (uint64_t)p->n * m;
  }
  launch_update_tail(g->g, single(f), p->stream);
  p->prefix_valid = true;
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_pf_update(a3d_pf* p, const a3d_grid* g, const float* x, const float* y, const float* z, size_t m,
                  const float* ranges, size_t nb, double alpha, double sigma, double roll, double pitch) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_update: null filter");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_update: use the _dev form during capture");
  if (!p->n) return fail(A3D_ERR_STATE, "a3d_pf_update: filter has no particles (call init)");
  if (m && (!x || !y || !z)) return fail(A3D_ERR_ARG, "a3d_pf_update: null cloud");
  if (nb && !ranges) return fail(A3D_ERR_ARG, "a3d_pf_update: null ranges");
  if (!g) {
    // No map: every particle is outside (Grid3d::isIntoMap -> false), so all
    // weights become 0 and the mean 0 (ParticleFilter.cpp:137-141,158-195).
    std::vector<float> pp(7 * (size_t)p->n);
    int rc = a3d_pf_get_particles(p, pp.data(), p->n);
    if (rc) return rc;
    for (uint32_t i = 0; i < p->n; ++i) {
      float* q = &pp[7 * (size_t)i];
      q[5] = 0.f;  // wtp == 0 -> wp = 0
      q[6] = 0.f;
      q[4] = 0.f;
    }
    rc = a3d_pf_set_particles(p, pp.data(), p->n);
    if (rc) return rc;
    A3D_HIP(hipMemset(p->mean, 0, 16));
    return A3D_OK;
  }
  int rc = pf_stage(p, 3 * m + 4 * nb + 16);
  if (rc) return rc;
  float* cx = p->stage;
  float* cy = cx + m;
  float* cz = cy + m;
  float* rg = cz + m;
  if (m) {
    A3D_HIP(hipMemcpyAsync(cx, x, m * 4, hipMemcpyHostToDevice, p->stream));
    A3D_HIP(hipMemcpyAsync(cy, y, m * 4, hipMemcpyHostToDevice, p->stream));
    A3D_HIP(hipMemcpyAsync(cz, z, m * 4, hipMemcpyHostToDevice, p->stream));
  }
  if (nb) A3D_HIP(hipMemcpyAsync(rg, ranges, nb * 16, hipMemcpyHostToDevice, p->stream));
  rc = a3d_pf_update_dev(p, g, cx, cy, cz, m, nb ? rg : nullptr, nb, alpha, sigma, roll, pitch);
  if (rc) return rc;
  A3D_HIP(hipStreamSynchronize(p->stream));
  return A3D_OK;
}

}  // extern "C"

namespace a3d {
hipStream_t pf_stream(const a3d_pf* p) { return p->stream; }

// a3d_pf_resample_dev with the uniform read from device memory when the
// resample runs (the node's draws leave it there), reference mode.
int pf_resample_device_u01(a3d_pf* p, const float* d_u01) {
  if (!p || !d_u01) return fail(A3D_ERR_ARG, "resample: null argument");
  if (!p->n) return fail(A3D_ERR_STATE, "resample: filter has no particles (call init)");
  Filters fs = single(p->dev());
  fs.one.idx_out = p->idx;
  fs.u01 = d_u01;
  launch_resample(fs, A3D_RESAMPLE_REFERENCE, p->prefix_valid, p->stream);
  p->prefix_valid = false;
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}
}  // namespace a3d

extern "C" {

int a3d_pf_resample_dev(a3d_pf* p, float u01, int mode) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_resample: null filter");
  if (!p->n) return fail(A3D_ERR_STATE, "a3d_pf_resample: filter has no particles (call init)");
  if (mode != A3D_RESAMPLE_REFERENCE && mode != A3D_RESAMPLE_FP64)
    return fail(A3D_ERR_ARG, "a3d_pf_resample: unknown mode");
  FilterDev f = p->dev();
  f.u01 = u01;
  f.idx_out = p->idx;
  launch_resample(single(f), mode, p->prefix_valid, p->stream);
  p->prefix_valid = false;
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_pf_resample(a3d_pf* p, float u01, int mode, uint32_t* idx_out) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_resample: null filter");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_resample: use the _dev form during capture");
  int rc = a3d_pf_resample_dev(p, u01, mode);
  if (rc) return rc;
  if (idx_out) A3D_HIP(hipMemcpyAsync(idx_out, p->idx, (size_t)p->n * 4, hipMemcpyDeviceToHost, p->stream));
  A3D_HIP(hipStreamSynchronize(p->stream));
  return A3D_OK;
}

int a3d_pf_synchronize(a3d_pf* p) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_synchronize: null filter");
  A3D_HIP(hipStreamSynchronize(p->stream));
  return A3D_OK;
}

int a3d_pf_get_mean(const a3d_pf* p, float mean[4]) {
  if (!p || !mean) return fail(A3D_ERR_ARG, "a3d_pf_get_mean: null argument");
  if (!p->mean) {
    std::fill(mean, mean + 4, 0.f);
    return A3D_OK;
  }
  A3D_HIP(hipMemcpyAsync(mean, p->mean, 16, hipMemcpyDeviceToHost, p->stream));
  A3D_HIP(hipStreamSynchronize(p->stream));
  return A3D_OK;
}

int a3d_pf_is_initialized(const a3d_pf* p, int* initialized) {
  if (!p || !initialized) return fail(A3D_ERR_ARG, "a3d_pf_is_initialized: null argument");
  *initialized = p->initialized ? 1 : 0;
  return A3D_OK;
}

int a3d_pf_size(const a3d_pf* p, size_t* n) {
  if (!p || !n) return fail(A3D_ERR_ARG, "a3d_pf_size: null argument");
  *n = p->n;
  return A3D_OK;
}

int a3d_pf_get_particles(const a3d_pf* p, float* p7, size_t n) {
  if (!p || (!p7 && n)) return fail(A3D_ERR_ARG, "a3d_pf_get_particles: null argument");
  if (n != p->n) return fail(A3D_ERR_ARG, "a3d_pf_get_particles: n must equal the particle count");
  if (!n) return A3D_OK;
  std::vector<float> soa(7 * n);
  A3D_HIP(hipMemcpyAsync(soa.data(), p->x, 7 * n * 4, hipMemcpyDeviceToHost, p->stream));
  A3D_HIP(hipStreamSynchronize(p->stream));
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 7; ++k) p7[7 * i + k] = soa[k * n + i];
  return A3D_OK;
}

int a3d_pf_set_particles(a3d_pf* p, const float* p7, size_t n) {
  if (!p || !p7 || n == 0) return fail(A3D_ERR_ARG, "a3d_pf_set_particles: null argument or n == 0");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_set_particles: not allowed during capture");
  if (n > 0xffffffffull) return fail(A3D_ERR_ARG, "a3d_pf_set_particles: too many particles");
  int rc = pf_resize(p, (uint32_t)n);
  if (rc) return rc;
  std::vector<float> soa(7 * n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 7; ++k) soa[k * n + i] = p7[7 * i + k];
  A3D_HIP(hipMemcpyAsync(p->x, soa.data(), 7 * n * 4, hipMemcpyHostToDevice, p->stream));
  A3D_HIP(hipStreamSynchronize(p->stream));
  p->initialized = true;
  p->prefix_valid = false;
  return A3D_OK;
}

int a3d_pf_capture_begin(a3d_pf* p) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_capture_begin: null filter");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_capture_begin: already capturing");
  A3D_HIP(hipStreamSynchronize(p->stream));
  A3D_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  p->capturing = true;
  return A3D_OK;
}

int a3d_pf_capture_end(a3d_pf* p, a3d_graph** out) {
  if (!p || !out) return fail(A3D_ERR_ARG, "a3d_pf_capture_end: null argument");
  if (!p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_capture_end: not capturing");
  p->capturing = false;
  a3d_graph* G = new a3d_graph();
  G->stream = p->stream;
  if (!hip_ok(hipStreamEndCapture(p->stream, &G->graph), "hipStreamEndCapture") ||
      !hip_ok(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
    if (G->graph) (void)hipGraphDestroy(G->graph);
    delete G;
    return A3D_ERR_HIP;
  }
  *out = G;
  return A3D_OK;
}

int a3d_graph_launch(a3d_graph* G) {
  if (!G) return fail(A3D_ERR_ARG, "a3d_graph_launch: null graph");
  A3D_HIP(hipGraphLaunch(G->exec, G->stream));
  return A3D_OK;
}

int a3d_graph_destroy(a3d_graph* G) {
  if (!G) return A3D_OK;
  (void)hipStreamSynchronize(G->stream);
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  delete G;
  return A3D_OK;
}

int a3d_pf_timing(a3d_pf* p, int reset, double* weight_ms_avg, uint64_t* launches, uint64_t* evals) {
  if (!p) return fail(A3D_ERR_ARG, "a3d_pf_timing: null filter");
  if (p->capturing) return fail(A3D_ERR_STATE, "a3d_pf_timing: not during capture");
  A3D_HIP(hipStreamSynchronize(p->stream));
  p->timing.resolve();
  if (weight_ms_avg) *weight_ms_avg = p->timing.launches ? p->timing.acc_ms / p->timing.launches : 0.0;
  if (launches) *launches = p->timing.launches;
  if (evals) *evals = p->timing.evals;
  if (reset) {
    p->timing.acc_ms = 0.0;
    p->timing.launches = 0;
    p->timing.evals = 0;
  }
  return A3D_OK;
}

}  // extern "C"
