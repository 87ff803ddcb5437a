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
// The order is a counting sort of the particles by a 12-bit key: a Morton
// code of (x, y, z, yaw), each quantised to 8 levels of half a standard
// deviation around the particles' mean (so the levels stay populated however
// concentrated the set is).  Four kernels, no library calls and no memset,
// so the sequence can be captured into a HIP graph:
//   moments  one workgroup: mean and deviation of x, y, z, a; clears the histogram
//   keys     key per particle, histogram by atomics
//   scan     one workgroup: exclusive scan of the 4,096 counts
//   scatter  order[offset[key]++] = i (order inside a key is arbitrary)
#include "a3d_math.h"

namespace a3d {
namespace {

constexpr int kBT = 1024;  // threads of the single-workgroup kernels

__device__ inline uint32_t spread3(uint32_t v) {  // 3 bits -> bits 0, 4, 8
  return (v & 1u) | ((v & 2u) << 3) | ((v & 4u) << 6);
}

__device__ inline uint32_t level8(float v, float lo, float scale) {
  float t = (v - lo) * scale;  // lo = mean - 2 sd, scale = 2 / sd (0 when sd is 0)
  t = fminf(fmaxf(t, 0.0f), 7.0f);  // NaN -> 0
  return (uint32_t)t;
}

__global__ __launch_bounds__(kBT) void order_moments_kernel(FilterDev f, OrderDev o) {
  __shared__ double red[8][kBT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t k = tid; k < kOrderKeys + 1; k += kBT) o.hist[k] = 0u;
  double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
  const float* arr[4] = {f.x, f.y, f.z, f.a};
  constexpr int kU = 8;  // loads in flight per thread
  for (uint32_t i0 = tid; i0 < f.n; i0 += kBT * kU) {
    float v[kU][4];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const uint32_t i = i0 + (uint32_t)u * kBT;
#pragma unroll
      for (int k = 0; k < 4; ++k) v[u][k] = (i < f.n) ? arr[k][i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double d = isfinite(v[u][k]) ? (double)v[u][k] : 0.0;
        s1[k] += d;
        s2[k] += d * d;
      }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      s1[k] += __shfl_xor(s1[k], d);
      s2[k] += __shfl_xor(s2[k], d);
    }
    if (lane == 0) {
      red[k][wave] = s1[k];
      red[4 + k][wave] = s2[k];
    }
  }
  __syncthreads();
  if (tid < 4) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kBT / 64; ++w) {
      a += red[tid][w];
      b += red[4 + tid][w];
    }
    const double n = (double)(f.n ? f.n : 1u);
    const double mean = a / n;
    const double var = b / n - mean * mean;
    const double sd = var > 0.0 ? sqrt(var) : 0.0;
    o.box[tid] = (float)(mean - 2.0 * sd);
    o.box[4 + tid] = (sd > 0.0 && sd < 1e30) ? (float)(2.0 / sd) : 0.0f;
  }
}

__global__ void order_keys_kernel(FilterDev f, OrderDev o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const uint32_t key = spread3(level8(f.x[i], o.box[0], o.box[4])) |
                       (spread3(level8(f.y[i], o.box[1], o.box[5])) << 1) |
                       (spread3(level8(f.z[i], o.box[2], o.box[6])) << 2) |
                       (spread3(level8(f.a[i], o.box[3], o.box[7])) << 3);
  o.key[i] = key;
  atomicAdd(&o.hist[key], 1u);
}

__global__ __launch_bounds__(kBT) void order_scan_kernel(OrderDev o) {
  // kOrderKeys / kBT consecutive counts per thread
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
  order_moments_kernel<<<1, kBT, 0, s>>>(f, o);
  order_keys_kernel<<<blocks, 256, 0, s>>>(f, o);
  order_scan_kernel<<<1, kBT, 0, s>>>(o);
  order_scatter_kernel<<<blocks, 256, 0, s>>>(f, o);
}

}  // namespace a3d
