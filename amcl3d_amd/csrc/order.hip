// Processing order of a filter's particles for the cloud-weight kernel.
//
// A wavefront gathers prob for the same cloud point under P particles' poses;
// the closer those poses, the fewer cache lines one gather instruction
// touches and the smaller the set of cells resident caches must hold.  The
// particle arrays themselves stay in index order (the reference's sums run
// over that order, ParticleFilter.cpp:151-193), so only the order in which
// the kernel visits particles changes, and with it nothing but speed: each
// particle's weight is computed by one lane exactly as before.
//
// The order is a counting sort of the particles by a 16-bit key: a Morton
// code of (x, y, z, yaw) quantised to 16 levels each over the particles' own
// bounding box.  Five kernels, no library calls and no memset, so the
// sequence can be captured into a HIP graph:
//   bbox     one workgroup: min/max of x, y, z, a; clears the histogram
//   keys     key per particle, histogram by atomics
//   scan     one workgroup: exclusive scan of the 65,536 counts
//   scatter  order[offset[key]++] = i (order inside a key is arbitrary)
#include "a3d_math.h"

namespace a3d {
namespace {

constexpr int kBT = 1024;  // threads of the single-workgroup kernels

__device__ inline uint32_t spread4(uint32_t v) {  // 4 bits -> bits 0, 4, 8, 12
  return (v & 1u) | ((v & 2u) << 3) | ((v & 4u) << 6) | ((v & 8u) << 9);
}

__device__ inline uint32_t level16(float v, float lo, float scale) {
  float t = (v - lo) * scale;  // scale = 16 / extent (0 for an empty extent)
  t = fminf(fmaxf(t, 0.0f), 15.0f);  // NaN -> 0
  return (uint32_t)t;
}

__global__ __launch_bounds__(kBT) void order_bbox_kernel(FilterDev f, OrderDev o) {
  __shared__ float red[8][kBT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t k = tid; k < kOrderKeys + 1; k += kBT) o.hist[k] = 0u;
  float mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  float mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = tid; i < f.n; i += kBT) {
    const float v[4] = {f.x[i], f.y[i], f.z[i], f.a[i]};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      mn[k] = fminf(mn[k], v[k]);  // fminf ignores a NaN operand
      mx[k] = fmaxf(mx[k], v[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], d));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d));
    }
    if (lane == 0) {
      red[k][wave] = mn[k];
      red[4 + k][wave] = mx[k];
    }
  }
  __syncthreads();
  if (tid < 4) {
    float a = INFINITY, b = -INFINITY;
    for (int w = 0; w < kBT / 64; ++w) {
      a = fminf(a, red[tid][w]);
      b = fmaxf(b, red[4 + tid][w]);
    }
    const float ext = b - a;
    o.box[tid] = a;
    o.box[4 + tid] = (ext > 0.0f && ext < INFINITY) ? 16.0f / ext : 0.0f;
  }
}

__global__ void order_keys_kernel(FilterDev f, OrderDev o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const uint32_t key = spread4(level16(f.x[i], o.box[0], o.box[4])) |
                       (spread4(level16(f.y[i], o.box[1], o.box[5])) << 1) |
                       (spread4(level16(f.z[i], o.box[2], o.box[6])) << 2) |
                       (spread4(level16(f.a[i], o.box[3], o.box[7])) << 3);
  o.key[i] = key;
  atomicAdd(&o.hist[key], 1u);
}

__global__ __launch_bounds__(kBT) void order_scan_kernel(OrderDev o) {
  // kOrderKeys / kBT = 64 consecutive counts per thread
  __shared__ uint32_t part[kBT];
  constexpr int kPer = kOrderKeys / kBT;
  const int tid = threadIdx.x;
  uint32_t* h = o.hist + (size_t)tid * kPer;
  uint32_t s = 0;
  for (int k = 0; k < kPer; ++k) s += h[k];
  part[tid] = s;
  __syncthreads();
  for (int d = 1; d < kBT; d <<= 1) {  // inclusive scan of the per-thread totals
    const uint32_t v = (tid >= d) ? part[tid - d] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - s;  // exclusive
  for (int k = 0; k < kPer; ++k) {
    const uint32_t c = h[k];
    h[k] = run;
    run += c;
  }
}

__global__ void order_scatter_kernel(FilterDev f, OrderDev o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const uint32_t pos = atomicAdd(&o.hist[o.key[i]], 1u);
  o.order[pos] = i;  // pos < n: the offsets of key k cover exactly its count
}

}  // namespace

void launch_order(const FilterDev& f, const OrderDev& o, hipStream_t s) {
  if (f.n == 0) return;
  const uint32_t blocks = (f.n + 255) / 256;
  order_bbox_kernel<<<1, kBT, 0, s>>>(f, o);
  order_keys_kernel<<<blocks, 256, 0, s>>>(f, o);
  order_scan_kernel<<<1, kBT, 0, s>>>(o);
  order_scatter_kernel<<<blocks, 256, 0, s>>>(f, o);
}

}  // namespace a3d
