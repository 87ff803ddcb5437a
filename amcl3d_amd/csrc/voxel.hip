// pcl::VoxelGrid<pcl::PointXYZ> on the device, as the node applies it to
// every scan before ParticleFilter::update (Node.cpp:130-141: setLeafSize
// with voxel_size on all three axes, PCL defaults otherwise).
//
// PCL 1.7.2's VoxelGrid::applyFilter, with its defaults (no filter field,
// min_points_per_voxel 0, downsample_all_data true), computes:
//   inv = 1.0f / leaf; min_p, max_p over the finite points;
//   (int64)((max_p - min_p) * inv) + 1 voxels per axis, and if their product
//   exceeds INT32_MAX the input is returned unchanged;
//   min_b = (int)floor(min_p * inv), div_b = floor(max_p * inv) - min_b + 1;
//   per finite point idx = ijk0 + ijk1 * div0 + ijk2 * div0 * div1 with
//   ijk = (int)(floor(p * inv) - (float)min_b), all in float;
//   one output point per distinct idx, in increasing idx, at the float
//   centroid (sum of x, y, z in order, divided by the count as float).
// Here the (idx, point) pairs are radix sorted, which is stable: the points of
// one voxel are summed in point-index order.  PCL's std::sort leaves them in
// an order of its own, so a centroid of three or more points can differ from
// PCL's in the last bits; the voxel set and its order are the same.
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <vector>

#include "a3d_math.h"

namespace a3d {
namespace {

constexpr int kVT = 1024;
constexpr uint32_t kNoVoxel = 0xffffffffu;

struct VoxelMeta {
  float inv;
  uint32_t nfin;
  int overflow;
  int min_b[3];
  int mul[3];
};

__device__ inline bool finite3(const float* p) { return isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]); }

__device__ inline const float* point_at(const uint8_t* base, size_t stride, uint32_t i) {
  return reinterpret_cast<const float*>(base + (size_t)i * stride);
}

// One workgroup: bounds of the finite points and the voxel lattice.
__global__ __launch_bounds__(kVT) void voxel_meta_kernel(const uint8_t* pts, uint32_t n, size_t stride, float leaf,
                                                         VoxelMeta* meta) {
  __shared__ float red[6][kVT / 64];
  __shared__ uint32_t cnt[kVT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  uint32_t c = 0;
  for (uint32_t i = tid; i < n; i += kVT) {
    const float* p = point_at(pts, stride, i);
    if (!finite3(p)) continue;
    ++c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], p[k]);
      mx[k] = fmaxf(mx[k], p[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], d));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d));
    }
  }
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      red[k][wave] = mn[k];
      red[3 + k][wave] = mx[k];
    }
    cnt[wave] = c;
  }
  __syncthreads();
  if (tid != 0) return;
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  uint32_t total = 0;
  for (int w = 0; w < kVT / 64; ++w) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(lo[k], red[k][w]);
      hi[k] = fmaxf(hi[k], red[3 + k][w]);
    }
    total += cnt[w];
  }
  VoxelMeta m{};
  m.inv = 1.0f / leaf;
  m.nfin = total;
  if (total > 0) {
    int64_t d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = (int64_t)((hi[k] - lo[k]) * m.inv) + 1;
    m.overflow = (d[0] * d[1] * d[2] > (int64_t)INT32_MAX) ? 1 : 0;
    int div[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      m.min_b[k] = (int)floorf(lo[k] * m.inv);
      div[k] = (int)floorf(hi[k] * m.inv) - m.min_b[k] + 1;
    }
    m.mul[0] = 1;
    m.mul[1] = div[0];
    m.mul[2] = div[0] * div[1];
  }
  *meta = m;
}

__global__ void voxel_key_kernel(const uint8_t* pts, uint32_t n, size_t stride, const VoxelMeta* meta,
                                 uint32_t* key, uint32_t* val) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const VoxelMeta m = *meta;
  const float* p = point_at(pts, stride, i);
  uint32_t k = kNoVoxel;
  if (finite3(p)) {
    int ijk[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ijk[a] = (int)(floorf(p[a] * m.inv) - (float)m.min_b[a]);
    k = (uint32_t)(ijk[0] * m.mul[0] + ijk[1] * m.mul[1] + ijk[2] * m.mul[2]);
  }
  key[i] = k;
  val[i] = i;
}

__global__ void voxel_head_kernel(const uint32_t* key, uint32_t n, uint32_t* head) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (key[i] != kNoVoxel && (i == 0 || key[i] != key[i - 1])) ? 1u : 0u;
}

__global__ void voxel_centroid_kernel(const uint8_t* pts, size_t stride, const uint32_t* key, const uint32_t* val,
                                      const uint32_t* head, const uint32_t* pos, uint32_t n, float* ox, float* oy,
                                      float* oz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  const uint32_t k = key[i];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  uint32_t j = i;
  for (; j < n && key[j] == k; ++j) {
    const float* p = point_at(pts, stride, val[j]);
    sx += p[0];
    sy += p[1];
    sz += p[2];
  }
  const float c = (float)(j - i);
  const uint32_t o = pos[i];
  ox[o] = sx / c;
  oy[o] = sy / c;
  oz[o] = sz / c;
}

__global__ void voxel_copy_kernel(const uint8_t* pts, size_t stride, uint32_t n, float* ox, float* oy, float* oz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* p = point_at(pts, stride, i);
  ox[i] = p[0];
  oy[i] = p[1];
  oz[i] = p[2];
}

}  // namespace

// Device buffers of voxel_grid_device, kept by a caller that filters many
// clouds: the meta record, six arrays of `cap` words (keys, values, their
// sorted copies, voxel heads and positions) and the sort / scan temporary.
struct VoxelWorkspace {
  VoxelMeta* meta = nullptr;
  uint32_t* buf = nullptr;
  size_t cap = 0;
  void* tmp = nullptr;
  size_t tmp_cap = 0;
};

VoxelWorkspace* voxel_workspace_new() { return new VoxelWorkspace(); }

void voxel_workspace_free(VoxelWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->meta);
  (void)hipFree(w->buf);
  (void)hipFree(w->tmp);
  delete w;
}

// d_pts: n points, `stride` bytes apart (x, y, z the first three floats).
// Writes the filtered cloud as d_x, d_y, d_z (capacity n) and its size.
// ws: buffers to reuse (grown as needed), or null for buffers of this call.
int voxel_grid_device(const void* d_pts, size_t n, size_t stride, float leaf, float* d_x, float* d_y, float* d_z,
                      size_t* out_n, hipStream_t s, VoxelWorkspace* ws) {
  *out_n = 0;
  if (n == 0) return A3D_OK;
  const uint8_t* pts = static_cast<const uint8_t*>(d_pts);
  const uint32_t nn = (uint32_t)n;
  const uint32_t blocks = (nn + 255) / 256;
  VoxelWorkspace local;
  VoxelWorkspace& w = ws ? *ws : local;
  uint32_t *key = nullptr, *val = nullptr, *key2 = nullptr, *val2 = nullptr, *head = nullptr, *pos = nullptr;
  size_t tmp_bytes = 0, tmp2 = 0;
  auto cleanup = [&]() {
    if (ws) return;
    (void)hipFree(local.meta);
    (void)hipFree(local.buf);
    (void)hipFree(local.tmp);
  };
#define A3D_TRY(call)           \
  if (!hip_ok((call), #call)) { \
    cleanup();                  \
    return A3D_ERR_HIP;         \
  }
  if (!w.meta) A3D_TRY(hipMalloc(&w.meta, sizeof(VoxelMeta)));
  VoxelMeta* const meta = w.meta;
  voxel_meta_kernel<<<1, kVT, 0, s>>>(pts, nn, stride, leaf, meta);
  A3D_TRY(hipGetLastError());
  VoxelMeta hm{};
  A3D_TRY(hipMemcpyAsync(&hm, meta, sizeof(VoxelMeta), hipMemcpyDeviceToHost, s));
  A3D_TRY(hipStreamSynchronize(s));
  if (hm.nfin == 0) {
    cleanup();
    return A3D_OK;
  }
  if (hm.overflow) {  // PCL: integer indices would overflow, output = input
    voxel_copy_kernel<<<blocks, 256, 0, s>>>(pts, stride, nn, d_x, d_y, d_z);
    A3D_TRY(hipGetLastError());
    A3D_TRY(hipStreamSynchronize(s));
    cleanup();
    *out_n = n;
    return A3D_OK;
  }
  if (w.cap < n) {
    (void)hipFree(w.buf);
    w.buf = nullptr;
    w.cap = 0;
    A3D_TRY(hipMalloc(&w.buf, 6 * n * 4));
    w.cap = n;
  }
  key = w.buf;
  val = key + w.cap;
  key2 = val + w.cap;
  val2 = key2 + w.cap;
  head = val2 + w.cap;
  pos = head + w.cap;
  voxel_key_kernel<<<blocks, 256, 0, s>>>(pts, nn, stride, meta, key, val);
  A3D_TRY(hipGetLastError());
  A3D_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key2, val, val2, (int)nn, 0, 32, s));
  A3D_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, head, pos, (int)nn, s));
  tmp_bytes = std::max(tmp_bytes, tmp2);
  if (w.tmp_cap < tmp_bytes) {
    (void)hipFree(w.tmp);
    w.tmp = nullptr;
    w.tmp_cap = 0;
    A3D_TRY(hipMalloc(&w.tmp, tmp_bytes));
    w.tmp_cap = tmp_bytes;
  }
  void* const tmp = w.tmp;
  A3D_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key2, val, val2, (int)nn, 0, 32, s));
  voxel_head_kernel<<<blocks, 256, 0, s>>>(key2, nn, head);
  A3D_TRY(hipGetLastError());
  A3D_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, head, pos, (int)nn, s));
  voxel_centroid_kernel<<<blocks, 256, 0, s>>>(pts, stride, key2, val2, head, pos, nn, d_x, d_y, d_z);
  A3D_TRY(hipGetLastError());
  uint32_t last[2] = {0, 0};
  A3D_TRY(hipMemcpyAsync(&last[0], pos + (nn - 1), 4, hipMemcpyDeviceToHost, s));
  A3D_TRY(hipMemcpyAsync(&last[1], head + (nn - 1), 4, hipMemcpyDeviceToHost, s));
  A3D_TRY(hipStreamSynchronize(s));
#undef A3D_TRY
  cleanup();
  *out_n = (size_t)last[0] + last[1];
  return A3D_OK;
}

}  // namespace a3d
