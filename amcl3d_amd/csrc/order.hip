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
// so the sequence can be captured into a HIP graph; the batched form sorts
// many filters at once, one filter per grid row (blockIdx.y):
//   moments  64 blocks: partial sums of x, y, z, a and their squares; clears the histogram
//   keys     every block derives the levels from the partial sums; key per
//            particle, histogram per block in LDS, added to the global one
//   scan     one workgroup: exclusive scan of the 4,096 counts
//   scatter  order[offset[key]++] = i (order inside a key is arbitrary)
#include <algorithm>

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

constexpr int kMB = 64;   // blocks of the moments kernel
constexpr int kKT = 256;  // threads per block of the keys and scatter kernels

// The filter of this block: the launch's one, or row base + blockIdx.y of
// the arrays.
struct OrderJob {
  FilterDev f;
  OrderDev o;
  const FilterDev* fs;
  const OrderDev* os;
  uint32_t base;
};

template <bool MANY>
__device__ inline void job_filter(const OrderJob& j, FilterDev& f, OrderDev& o) {
  if (MANY) {
    f = j.fs[j.base + blockIdx.y];
    o = j.os[j.base + blockIdx.y];
  } else {
    f = j.f;
    o = j.o;
  }
}

// Partial sums of x, y, z, a and their squares (in double) per block, and the
// histogram cleared for the keys kernel.
template <bool MANY>
__global__ __launch_bounds__(kKT) void order_moments_kernel(OrderJob j) {
  FilterDev f;
  OrderDev o;
  job_filter<MANY>(j, f, o);
  __shared__ double red[8][kKT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t gt = blockIdx.x * kKT + tid, gs = kMB * kKT;
  for (uint32_t k = gt; k < kOrderKeys + 1; k += gs) o.hist[k] = 0u;
  double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
  const float* arr[4] = {f.x, f.y, f.z, f.a};
  for (uint32_t i = gt; i < f.n; i += gs) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float v = arr[k][i];
      const double d = isfinite(v) ? (double)v : 0.0;
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
  if (tid < 8) {
    double a = 0.0;
    for (int w = 0; w < kKT / 64; ++w) a += red[tid][w];
    o.part[blockIdx.x * 8 + tid] = a;
  }
}

// The levels of one block: mean - 2 sd and 2 / sd per coordinate.
__device__ inline void order_levels(const FilterDev& f, const OrderDev& o, float* box) {
  if (threadIdx.x < 4) {
    double a = 0.0, b = 0.0;
    for (int p = 0; p < kMB; ++p) {
      a += o.part[p * 8 + threadIdx.x];
      b += o.part[p * 8 + 4 + threadIdx.x];
    }
    const double n = (double)(f.n ? f.n : 1u);
    const double mean = a / n;
    const double var = b / n - mean * mean;
    const double sd = var > 0.0 ? sqrt(var) : 0.0;
    box[threadIdx.x] = (float)(mean - 2.0 * sd);
    box[4 + threadIdx.x] = (sd > 0.0 && sd < 1e30) ? (float)(2.0 / sd) : 0.0f;
  }
}

__device__ inline uint32_t order_key(const FilterDev& f, uint32_t i, const float* box) {
  return spread3(level8(f.x[i], box[0], box[4])) | (spread3(level8(f.y[i], box[1], box[5])) << 1) |
         (spread3(level8(f.z[i], box[2], box[6])) << 2) | (spread3(level8(f.a[i], box[3], box[7])) << 3);
}

// Keys, counted per block in LDS and added to the global histogram once per
// occupied key (few global atomics however concentrated the particles are).
template <bool MANY>
__global__ __launch_bounds__(kKT) void order_keys_kernel(OrderJob j) {
  FilterDev f;
  OrderDev o;
  job_filter<MANY>(j, f, o);
  if (blockIdx.x * kKT >= f.n) return;  // whole block past this filter's end (batched launches)
  __shared__ uint32_t h[kOrderKeys];
  __shared__ float box[8];
  for (uint32_t k = threadIdx.x; k < kOrderKeys; k += kKT) h[k] = 0u;
  order_levels(f, o, box);
  __syncthreads();
  const uint32_t i = blockIdx.x * kKT + threadIdx.x;
  if (i < f.n) {
    const uint32_t key = order_key(f, i, box);
    o.key[i] = key;
    atomicAdd(&h[key], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kOrderKeys; k += kKT)
    if (h[k]) atomicAdd(&o.hist[k], h[k]);
}

template <bool MANY>
__global__ __launch_bounds__(kBT) void order_scan_kernel(OrderJob j) {
  FilterDev f;
  OrderDev o;
  job_filter<MANY>(j, f, o);
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

// order[offset[key]++] = i: a block reserves one range per occupied key and
// places its particles of that key inside it (in arbitrary order).
template <bool MANY>
__global__ __launch_bounds__(kKT) void order_scatter_kernel(OrderJob j) {
  FilterDev f;
  OrderDev o;
  job_filter<MANY>(j, f, o);
  if (blockIdx.x * kKT >= f.n) return;  // whole block past this filter's end (batched launches)
  __shared__ uint32_t cnt[kOrderKeys];
  for (uint32_t k = threadIdx.x; k < kOrderKeys; k += kKT) cnt[k] = 0u;
  __syncthreads();
  const uint32_t i = blockIdx.x * kKT + threadIdx.x;
  uint32_t key = 0, rank = 0;
  if (i < f.n) {
    key = o.key[i];
    rank = atomicAdd(&cnt[key], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kOrderKeys; k += kKT)
    if (cnt[k]) cnt[k] = atomicAdd(&o.hist[k], cnt[k]);  // now the block's base for key k
  __syncthreads();
  if (i < f.n) o.order[cnt[key] + rank] = i;  // < n: the ranges of key k cover exactly its count
}

}  // namespace

void launch_order(const FilterDev& f, const OrderDev& o, hipStream_t s) {
  if (f.n == 0) return;
  const uint32_t blocks = (f.n + kKT - 1) / kKT;
  const OrderJob j{f, o, nullptr, nullptr, 0};
  order_moments_kernel<false><<<kMB, kKT, 0, s>>>(j);
  order_keys_kernel<false><<<blocks, kKT, 0, s>>>(j);
  order_scan_kernel<false><<<1, kBT, 0, s>>>(j);
  order_scatter_kernel<false><<<blocks, kKT, 0, s>>>(j);
}

void launch_order_many(const FilterDev* fs, const OrderDev* os, uint32_t count, uint32_t max_n, hipStream_t s) {
  if (count == 0 || max_n == 0) return;
  const uint32_t blocks = (max_n + kKT - 1) / kKT;
  for (uint32_t base = 0; base < count; base += 65535u) {
    const uint32_t rows = std::min(count - base, 65535u);
    const OrderJob j{FilterDev{}, OrderDev{}, fs, os, base};
    order_moments_kernel<true><<<dim3(kMB, rows), kKT, 0, s>>>(j);
    order_keys_kernel<true><<<dim3(blocks, rows), kKT, 0, s>>>(j);
    order_scan_kernel<true><<<dim3(1, rows), kBT, 0, s>>>(j);
    order_scatter_kernel<true><<<dim3(blocks, rows), kKT, 0, s>>>(j);
  }
}

}  // namespace a3d
