// Internal declarations shared by the HIP sources of libamcl3d_gpu.
//
// Device layout (HBM):
//  * grid: prob as one float per cell, x fastest, index ix + iy*sx + iz*sx*sy
//    (the reference's cell order, Grid3d.cpp:187), 64-bit when the map has
//    2^32 - 1 cells or more, followed by one float 0 that rejected points
//    read; the squared distance field is a second, optional, array of the
//    cells' shape.
//  * filter: structure of arrays x, y, z, a, w, wp, wr (one float each per
//    particle, the fields of struct Particle, ParticleFilter.h:35-49).
//  * cloud: three float arrays x, y, z in cloud order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/amcl3d_gpu.h"

namespace a3d {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
bool hip_ok(hipError_t e, const char* what);

#define A3D_HIP(call)                                      \
  do {                                                     \
    if (!::a3d::hip_ok((call), #call)) return A3D_ERR_HIP; \
  } while (0)

// ------------------------------------------------------------ device views
struct GridDev {
  const float* prob;  // cells + 1 floats: prob[cells] = 0 is the cell of a rejected point
  float* dist;        // optional (map build output)
  uint32_t size[3];
  uint32_t idx64;
  uint64_t step_z;  // sx*sy
  uint64_t cells;
  double min[3];
  double max[3];
  double ext[3];    // max - min, as Grid3d.cpp:151-153
  float ext_hi[3];  // for a float q: (double)q < ext  <=>  q < ext_hi
  float pad_;
  double res;
  double inv_res;
  // For a float q other than -0.0: q >= 0 && q < ext_hi && floor(q/res) < size
  // <=> bits(q) < lim (both bounds are monotone in q, so they fold into one
  // float bound, and a negative q or a NaN has a larger bit pattern).
  uint32_t lim[3];
  // 1 when, on every axis, trunc(q * inv_res) equals floor(q / res) for every
  // float q in [0, lim): the cell index needs no window test at all.
  uint32_t clean;
  // 1 when the prob array with its zero cell, (cells + 1) * 4 bytes, fits in
  // 2^32 bytes: the cloud weight may load it through a structured buffer
  // (32-bit byte offsets).  idx64 == 0 and buf32 == 0: 32-bit cell indices
  // with 64-bit addresses.
  uint32_t buf32;
  // Half-width of the window around an integer inside which q*inv_res may
  // floor differently from the correctly rounded q/res (max(1e-9, size*2^-48)),
  // and twice that.
  double near_w[3];
  double near_w2[3];
  // 30 when cell indices may be formed by the 24-bit multiply-adds (sx <
  // 2^24 and sy * sz < 2^24; the index then comes out modulo 2^32): a
  // wavefront whose cells all lie in one 2^30-cell segment k may load them
  // through a structured buffer based at prob + k * 2^30, whose 32-bit byte
  // offset (record * 4 modulo 2^32) is then the offset inside the segment.
  // 0: never.  The allocation of such a grid of more than 2^30 cells is
  // rounded up to whole segments, so every such load stays inside it.
  uint32_t seg_bits;
  uint32_t pad2_;
};

// One filter and the parameters of the call being made on it.  Arrays are
// device pointers.
struct FilterDev {
  float* x;
  float* y;
  float* z;
  float* a;
  float* w;
  float* wp;
  float* wr;
  float* scratch;  // scratch_floats(n): tail sums, resample prefix, gather staging
  // 2 * kZrParts floats: the range of the finite rotated z of each of
  // kZrParts parts of the cloud, written before the cloud weight
  // (launch_weight) for its per-wavefront segment test; null: no segment
  // loads
  float* zr;
  uint8_t* inmap;  // n
  float* mean;     // 4
  const GridDev* gdev;  // device copy of this filter's grid; null: the launch's grid
  const uint32_t* order;  // cloud-weight visiting order of the particles; null: index order
  uint32_t n;
  // update
  uint32_t m;
  uint32_t nb;
  uint32_t rp0;  // roll == pitch == 0
  const float* cx;
  const float* cy;
  const float* cz;
  const float* ranges;  // nb x {r, ax, ay, az}
  float roll;
  float pitch;
  double alpha;
  double sigma;
  // predict / init draws: 4 standard normals per particle
  const float* z4;
  double delta[4];  // predict: delta x, y, z, a
  double dev[4];    // predict: fabs(delta*mod); init: x, y, z, a deviations
  float pose[4];    // init
  // resample
  float u01;
  uint32_t no_resample;  // a batch member that did not ask to resample
  uint32_t* idx_out;     // optional: ancestor indices
};

// A unit of cloud-weight work: particles [p0, p0+count) of one filter.
struct WorkItem {
  uint32_t filter;
  uint32_t p0;
  uint32_t count;
  uint32_t pad_;
};

// Filters a launch works on: one descriptor by value, or a device array.
struct Filters {
  const FilterDev* arr;  // device array, or null
  uint32_t count;        // number of filters
  uint32_t max_n;        // largest particle count
  const float* u01;      // optional device array: resample uniform of filter i
  FilterDev one;         // used when arr == null
};

inline Filters single(const FilterDev& f) {
  Filters s{};
  s.arr = nullptr;
  s.count = 1;
  s.max_n = f.n;
  s.one = f;
  return s;
}

// Buffers of the particle visiting order (launch_order).
constexpr uint32_t kOrderKeys = 4096;
struct OrderDev {
  double* part;     // 64 x 8 partial sums (moments of x, y, z, a)
  uint32_t* key;    // n
  uint32_t* order;  // n
  uint32_t* hist;   // kOrderKeys + 1
};
constexpr size_t order_bytes(uint32_t n) { return 64 * 8 * 8 + (2 * (size_t)n + kOrderKeys + 1) * 4; }
// order_bytes rounded up to keep consecutive views 256-byte aligned
constexpr size_t order_stride(uint32_t n) { return (order_bytes(n) + 255) & ~(size_t)255; }
inline OrderDev order_view(void* base, uint32_t n) {
  OrderDev o;
  o.part = static_cast<double*>(base);
  o.key = reinterpret_cast<uint32_t*>(o.part + 64 * 8);
  o.order = o.key + n;
  o.hist = o.order + n;
  return o;
}

// Host-side grid metadata, computed exactly as the reference does.
void grid_meta(const double mn[3], const double mx[3], double res, GridDev& g);

// ------------------------------------------------------------ launchers
// Cloud weight (Grid3d::computeCloudWeight for every in-map particle) and the
// range weight; writes wp, wr of in-map particles (out-of-map ones keep their
// previous wp, wr as in ParticleFilter.cpp:137-141).  A persistent launch
// over a work list uses d_counter[0..7] (one queue counter per XCD).
void launch_weight(const GridDev& g, const Filters& fs, const WorkItem* d_items, uint32_t nitems,
                   uint32_t* d_counter, int ppw, int persistent, hipStream_t s);
// Poses given explicitly (a3d_grid_cloud_weight): out[i]; zr: 2 * kZrParts
// floats of device memory for the cloud's z ranges.
void launch_weight_poses(const GridDev& g, const float* cx, const float* cy, const float* cz, uint32_t m,
                         const float* tx, const float* ty, const float* tz, const float* yaw, uint32_t n, float roll,
                         float pitch, float* out, float* zr, int ppw, hipStream_t s);
// Normalisation, blend and mean of ParticleFilter::update (ParticleFilter.cpp:150-195).
void launch_update_tail(const GridDev& g, const Filters& fs, hipStream_t s);
void launch_predict(const Filters& fs, hipStream_t s);
void launch_init(const Filters& fs, hipStream_t s);
// have_prefix: the update tail already left the running sum c of these weights.
void launch_resample(const Filters& fs, int mode, bool have_prefix, hipStream_t s);
// out[4i .. 4i+3] = mean of filter i
void launch_gather_means(const Filters& fs, float* out, hipStream_t s);
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);
// Visiting order of f's particles for the cloud weight, into o.order.
void launch_order(const FilterDev& f, const OrderDev& o, hipStream_t s);
// The same for count filters at once (device arrays fs[i], os[i]; max_n the
// largest particle count).
void launch_order_many(const FilterDev* fs, const OrderDev* os, uint32_t count, uint32_t max_n, hipStream_t s);

// Device map build (computeGrid, PointCloudTools.cpp:84-149).
int build_grid_device(const float* d_xyz, size_t n, const GridDev& g, float* d_prob, float* d_dist,
                      double sensor_dev, hipStream_t s);

// pcl::VoxelGrid as Node.cpp:130-141 applies it; synchronous on stream s.
// ws (optional): device buffers kept between calls (voxel_workspace_new).
struct VoxelWorkspace;
VoxelWorkspace* voxel_workspace_new();
void voxel_workspace_free(VoxelWorkspace* w);
int voxel_grid_device(const void* d_pts, size_t n, size_t stride, float leaf, float* d_x, float* d_y, float* d_z,
                      size_t* out_n, hipStream_t s, VoxelWorkspace* ws = nullptr);

// Counter-based draws for served sessions (Philox4x32-10 + Box-Muller).
// Writes 4 normals per particle to each filter's z4 and one uniform per
// filter to d_u01 (may be null).
void launch_draws(const Filters& fs, const uint64_t* d_seed, const uint64_t* d_step, float* d_u01, hipStream_t s);

// A filter's stream, and its resample with the uniform in device memory
// (capi.cpp; for the node glue).
hipStream_t pf_stream(const a3d_pf* p);
int pf_resample_device_u01(a3d_pf* p, const float* d_u01);

// Cloud-weight kernel tile sizes.
int clamp_ppw(int ppw);
// Particles per work item of a work list for the given mapping and ppw (the
// particle-lane kernels take an item per workgroup, the point-lane ones an
// item per wavefront).
int item_particles(int mapping, int ppw);
int current_ppw();
// Points per LDS cloud tile of the particle-lane kernels (32, 64, 128, 256).
int clamp_chunk(int points);
int current_chunk();
// Work-list dealing of persistent (fleet, served) launches: 0 contiguous
// ranges per XCD, k > 0 runs of k items dealt round-robin.
int current_deal();
int current_mapping();  // A3D_MAP_PARTICLE_LANES or A3D_MAP_POINT_LANES
int current_order();    // A3D_ORDER_INDEX or A3D_ORDER_SPATIAL

// Per-filter scratch layout, in floats (n = particles):
//   [0, 4n)    update tail: in-map wp, wr, then the weighted x, y, z, a
//   [4n, 10n)  resample: gathered x, y, z, a, wp, wr
//   [10n, 12n) resample FP64 prefix (n doubles, 8-byte aligned)
//   [12n, 13n) resample float running sum c: its value at the end of every
//              group of 64 weights (ceil(n / 64) floats)
//   then FP64 block totals, then 32 floats of per-filter sums (wtp, wtr, wt,
//   c, the irregular-weights flag; from kSumZr the cloud's z ranges, zr), then
//   the tile records of the tiled exact sums (kTiledSums sums of
//   tiled_tiles(n) records of 16 floats, then their sub-tile records, 16 per
//   tile of 8 floats each)
__host__ __device__ constexpr size_t sc_c64(uint32_t n) { return 10 * (size_t)n; }
__host__ __device__ constexpr size_t sc_c32(uint32_t n) { return 12 * (size_t)n; }
__host__ __device__ constexpr size_t sc_tot(uint32_t n) { return (13 * (size_t)n + 1) & ~(size_t)1; }
__host__ __device__ constexpr size_t sc_sums(uint32_t n) { return sc_tot(n) + 2 * ((size_t)n / 1024 + 2); }
constexpr uint32_t kTiledTile = 1024;  // terms per tile of a tiled exact sum
constexpr uint32_t kTiledSums = 5;     // sums one tiled launch computes at most
__host__ __device__ constexpr size_t tiled_tiles(uint32_t n) { return ((size_t)n + kTiledTile - 1) / kTiledTile; }
__host__ __device__ constexpr size_t sc_tiles(uint32_t n) { return (sc_sums(n) + 32 + 7) & ~(size_t)7; }
constexpr uint32_t kSumZr = 8;    // per-filter sums slots kSumZr .. kSumZr + 2 * kZrParts - 1: FilterDev::zr
constexpr uint32_t kZrParts = 12;  // parts of a cloud whose z ranges are taken in parallel
static_assert(kSumZr + 2 * kZrParts <= 32, "the z ranges must fit the 32 per-filter sums");
__host__ __device__ constexpr size_t scratch_floats(uint32_t n) {
  return sc_tiles(n) + (16 + 16 * 8) * kTiledSums * tiled_tiles(n);
}

// Event pairs around the cloud-weight launches.  Completed pairs are folded
// into acc_ms and reused as later launches take pairs, so a caller that never
// reads the timing (a node's callback loop) holds at most kMaxPairs pairs;
// while that many are still in flight, a launch goes untimed.
struct TimingPool {
  static constexpr size_t kMaxPairs = 64;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
  double acc_ms = 0.0;
  uint64_t launches = 0;
  uint64_t evals = 0;

  // Moves the pairs whose end event has completed from pending to spare.  A
  // query of an event still in flight answers hipErrorNotReady, which the
  // runtime may keep as the thread's last error; that is cleared again unless
  // an error was already pending before the queries.
  void reap() {
    const bool clean = hipPeekAtLastError() == hipSuccess;
    size_t k = 0;
    for (size_t i = 0; i < pending.size(); ++i) {
      const auto e = pending[i];
      if (hipEventQuery(e.second) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) acc_ms += ms;
        spare.push_back(e);
      } else {
        pending[k++] = e;
      }
    }
    pending.resize(k);
    if (clean) (void)hipGetLastError();
  }
  size_t held() const { return pending.size() + spare.size(); }

  std::pair<hipEvent_t, hipEvent_t> take() {
    if (spare.empty()) reap();
    if (!spare.empty()) {
      auto e = spare.back();
      spare.pop_back();
      return e;
    }
    if (held() >= kMaxPairs) return {nullptr, nullptr};
    std::pair<hipEvent_t, hipEvent_t> e{nullptr, nullptr};
    if (hipEventCreate(&e.first) != hipSuccess) return {nullptr, nullptr};
    if (hipEventCreate(&e.second) != hipSuccess) {
      (void)hipEventDestroy(e.first);
      return {nullptr, nullptr};
    }
    return e;
  }
  void resolve() {
    for (auto& e : pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) acc_ms += ms;
      spare.push_back(e);
    }
    pending.clear();
  }
  void destroy() {
    for (auto& e : pending) spare.push_back(e);
    pending.clear();
    for (auto& e : spare) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    spare.clear();
  }
};

}  // namespace a3d
