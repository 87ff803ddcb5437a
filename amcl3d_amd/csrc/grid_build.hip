// Device map build: computeGrid (PointCloudTools.cpp:84-149).
//
// For every cell corner s = (float)(min + i*res) (PointCloudTools.cpp:127-129)
// the squared distance to the nearest map point, as KdTreeFLANN's L2_Simple
// <float> computes it: d2 = ((dx*dx + dy*dy) + dz*dz), dx = s.x - p.x in
// float, minimised over all points; then prob = c1 * expf(-d2*d2*c2)
// (PointCloudTools.cpp:133-137).  The minimum is exact, not approximate:
//
//  1. points are counting-sorted into a bucket grid of h = 4 cells;
//  2. one workgroup per tile of 8x8x8 cells searches buckets in rounds of
//     growing radius r (h, 2h, 4h, ...): a round adds every bucket whose box
//     lies within r of the tile's box of query points, streaming their points
//     through LDS, each thread updating four cells;
//  3. a tile stops when every cell's best d2 <= r^2 (1 - 1e-6): any point not
//     yet seen is farther than r from the whole tile, and a float squared
//     distance is at least (1 - 3e-7) times the exact one, so it cannot be
//     smaller.
//
// expf is evaluated as (float)exp((double)x), the correctly rounded value.
//
// When only prob is built (no dist field), a tile also stops once its radius
// passes d2_cut, the smallest float d2 whose prob is exactly 0: prob is a
// non-increasing function of d2, so every cell whose nearest point is not yet
// found has prob 0 whatever that point is.  prob is unchanged bit for bit;
// in maps with large empty volumes this bounds the search.
#include <hipcub/hipcub.hpp>

#include <vector>

#include "a3d_math.h"

namespace a3d {
namespace {

constexpr int kBC = 4;         // cells per bucket edge
constexpr int kT = 8;          // cells per tile edge
constexpr int kGThreads = 128; // 4 cells per thread
constexpr int kBatch = 1024;   // points per LDS batch

struct BucketDev {
  double org[3];
  double h;
  uint32_t nb[3];
  uint32_t pad_;
  const uint32_t* start;  // nb0*nb1*nb2 + 1
  const float* px;
  const float* py;
  const float* pz;
};

__host__ __device__ inline uint32_t bucket_axis(double v, double org, double h, uint32_t nb) {
  double t = floor((v - org) / h);
  if (!(t >= 0.0)) t = 0.0;
  if (t > (double)(nb - 1)) t = (double)(nb - 1);
  return (uint32_t)t;
}

__global__ void bucket_count_kernel(const float* xyz, uint32_t n, BucketDev b, uint32_t* cnt, uint32_t* key) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) {  // PCL's kd-tree drops invalid points
    key[i] = 0xffffffffu;
    return;
  }
  const uint32_t k = bucket_axis(x, b.org[0], b.h, b.nb[0]) +
                     b.nb[0] * (bucket_axis(y, b.org[1], b.h, b.nb[1]) + b.nb[1] * bucket_axis(z, b.org[2], b.h, b.nb[2]));
  key[i] = k;
  atomicAdd(&cnt[k], 1u);
}

__global__ void bucket_scatter_kernel(const float* xyz, uint32_t n, const uint32_t* key, uint32_t* fill, float* px,
                                      float* py, float* pz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  if (k == 0xffffffffu) return;
  const uint32_t o = atomicAdd(&fill[k], 1u);
  px[o] = xyz[3 * (size_t)i];
  py[o] = xyz[3 * (size_t)i + 1];
  pz[o] = xyz[3 * (size_t)i + 2];
}

// prob of a cell whose nearest point is at float squared distance d
// (PointCloudTools.cpp:137): c1 * expf(-d * d * c2).
__device__ inline float cell_prob(float d, float c1, float c2) {
  const float e = -d * d * c2;
  return c1 * (float)exp((double)e);
}

// d2_cut: bisection over the bit patterns of non-negative floats (ordered as
// integers) for the smallest d with cell_prob(d) == 0; +inf gives 0.
__global__ void prob_cut_kernel(float c1, float c2, float* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t lo = 0u, hi = 0x7f800000u;  // cell_prob(0) = c1 > 0, cell_prob(inf) = 0
  while (hi - lo > 1u) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (cell_prob(__uint_as_float(mid), c1, c2) == 0.0f) hi = mid;
    else lo = mid;
  }
  out[0] = __uint_as_float(hi);
}

__global__ void fill_empty_kernel(float* prob, float* dist, uint64_t cells) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  prob[i] = 0.0f;
  if (dist) dist[i] = -1.0f;
}

// Lower bound of the distance between the box [lo, hi] of query points and
// bucket (bx, by, bz).  Edge buckets also hold the points clamped into them,
// so their outer side is open.  Bucket boxes are widened slightly to cover
// rounding in the bucket assignment.
__device__ inline double box_gap(const BucketDev& b, const float lo[3], const float hi[3], const uint32_t bi[3]) {
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double eps = b.h * 1e-6;
    const double blo = (bi[a] == 0) ? -INFINITY : b.org[a] + (double)bi[a] * b.h - eps;
    const double bhi = (bi[a] == b.nb[a] - 1) ? INFINITY : b.org[a] + (double)(bi[a] + 1) * b.h + eps;
    double d = 0.0;
    if (bhi < (double)lo[a]) d = (double)lo[a] - bhi;
    else if (blo > (double)hi[a]) d = blo - (double)hi[a];
    s += d * d;
  }
  return sqrt(s);
}

__global__ __launch_bounds__(kGThreads) void grid_tile_kernel(GridDev g, BucketDev b, float* prob, float* dist,
                                                              float c1, float c2, uint64_t tile0, uint32_t ntx,
                                                              uint32_t nty, double rmax, const float* d2_cut) {
  __shared__ float4 pts[kBatch];
  __shared__ uint32_t blist[kGThreads];
  __shared__ uint32_t bpre[kGThreads + 1];
  __shared__ uint32_t nlist;
  __shared__ uint32_t notdone;
  const int tid = threadIdx.x;
  const uint64_t tile = tile0 + blockIdx.x;
  const uint32_t tx = (uint32_t)(tile % ntx);
  const uint32_t ty = (uint32_t)((tile / ntx) % nty);
  const uint32_t tz = (uint32_t)(tile / ((uint64_t)ntx * nty));
  const uint32_t ix = tx * kT + (tid & 7);
  const uint32_t iy = ty * kT + ((tid >> 3) & 7);
  const uint32_t iz0 = tz * kT + (tid >> 6);  // cells iz0, iz0+2, iz0+4, iz0+6
  // query points, PointCloudTools.cpp:127-129
  const float qx = (float)(g.min[0] + (double)ix * g.res);
  const float qy = (float)(g.min[1] + (double)iy * g.res);
  float qz[4];
  bool cv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t iz = iz0 + 2 * c;
    qz[c] = (float)(g.min[2] + (double)iz * g.res);
    cv[c] = ix < g.size[0] && iy < g.size[1] && iz < g.size[2];
  }
  float best[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  // box of this tile's query points
  float lo[3], hi[3];
  {
    const uint32_t c0[3] = {tx * kT, ty * kT, tz * kT};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t c1_ = min(c0[a] + kT - 1, g.size[a] - 1);
      lo[a] = (float)(g.min[a] + (double)c0[a] * g.res);
      hi[a] = (float)(g.min[a] + (double)c1_ * g.res);
    }
  }
  // A round processes the buckets inside its cube (the bucket ranges covering
  // the tile box grown by r) whose gap is <= r, minus those the previous
  // round took by the same test.  Cubes and radii grow, so that union is
  // exactly what every earlier round processed, and every point within r of
  // the tile box lies in a processed bucket.
  const double cut = d2_cut ? (double)d2_cut[0] : INFINITY;
  double r_prev = -1.0;
  double r = b.h;
  uint32_t plo[3] = {1u, 1u, 1u}, phi[3] = {0u, 0u, 0u};  // previous cube (empty)
  for (;;) {
    uint32_t blo[3], bhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      blo[a] = bucket_axis((double)lo[a] - r, b.org[a], b.h, b.nb[a]);
      bhi[a] = bucket_axis((double)hi[a] + r, b.org[a], b.h, b.nb[a]);
    }
    const uint32_t wx = bhi[0] - blo[0] + 1, wy = bhi[1] - blo[1] + 1, wz = bhi[2] - blo[2] + 1;
    const uint64_t total = (uint64_t)wx * wy * wz;
    for (uint64_t base = 0; base < total; base += kGThreads) {
      if (tid == 0) nlist = 0;
      __syncthreads();
      const uint64_t e = base + tid;
      if (e < total) {
        uint32_t bi[3];
        bi[0] = blo[0] + (uint32_t)(e % wx);
        bi[1] = blo[1] + (uint32_t)((e / wx) % wy);
        bi[2] = blo[2] + (uint32_t)(e / ((uint64_t)wx * wy));
        const uint32_t id = bi[0] + b.nb[0] * (bi[1] + b.nb[1] * bi[2]);
        const uint32_t cnt = b.start[id + 1] - b.start[id];
        if (cnt) {
          const double gap = box_gap(b, lo, hi, bi);
          const bool in_prev = bi[0] >= plo[0] && bi[0] <= phi[0] && bi[1] >= plo[1] && bi[1] <= phi[1] &&
                               bi[2] >= plo[2] && bi[2] <= phi[2] && gap <= r_prev;
          if (gap <= r && !in_prev) {
            const uint32_t slot = atomicAdd(&nlist, 1u);
            blist[slot] = id;
          }
        }
      }
      __syncthreads();
      const uint32_t nl = nlist;
      if (nl == 0) {
        __syncthreads();  // nlist is reset at the top of the next chunk
        continue;
      }
      if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < nl; ++k) {
          bpre[k] = acc;
          acc += b.start[blist[k] + 1] - b.start[blist[k]];
        }
        bpre[nl] = acc;
      }
      __syncthreads();
      const uint32_t tp = bpre[nl];
      for (uint32_t pb = 0; pb < tp; pb += kBatch) {
        const uint32_t len = min((uint32_t)kBatch, tp - pb);
        for (uint32_t k = tid; k < len; k += kGThreads) {
          const uint32_t q = pb + k;
          uint32_t l = 0, h = nl;  // last list entry with bpre <= q
          while (h - l > 1) {
            const uint32_t mid = (l + h) >> 1;
            if (bpre[mid] <= q) l = mid;
            else h = mid;
          }
          const uint32_t gi = b.start[blist[l]] + (q - bpre[l]);
          pts[k] = make_float4(b.px[gi], b.py[gi], b.pz[gi], 0.f);
        }
        __syncthreads();
        for (uint32_t k = 0; k < len; ++k) {
          const float4 p = pts[k];
          const float dx = qx - p.x;
          const float dy = qy - p.y;
          float dxy = 0.f;
          dxy += dx * dx;
          dxy += dy * dy;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float dz = qz[c] - p.z;
            const float d = dxy + dz * dz;
            best[c] = fminf(best[c], d);
          }
        }
        __syncthreads();
      }
    }
    bool ok = true;
    const double lim = r * r * (1.0 - 1e-6);
    const bool past_cut = lim >= cut;  // unseen points give prob 0, as does best
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (cv[c] && !((double)best[c] <= lim) && !past_cut) ok = false;
    // workgroup AND of ok through one LDS word (every thread passes each barrier)
    if (tid == 0) notdone = 0u;
    __syncthreads();
    if (!ok) notdone = 1u;
    __syncthreads();
    const bool finished = (notdone == 0u) || r >= rmax;
    __syncthreads();
    if (finished) break;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      plo[a] = blo[a];
      phi[a] = bhi[a];
    }
    r_prev = r;
    // the radius at which the prob cut applies, if that comes before 2r
    const double rc = sqrt(cut / (1.0 - 1e-6)) * (1.0 + 1e-9);
    r = (rc > r && rc < 2.0 * r) ? rc : 2.0 * r;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (!cv[c]) continue;
    const uint64_t idx = (uint64_t)ix + (uint64_t)iy * g.size[0] + (uint64_t)(iz0 + 2 * c) * g.step_z;
    const float d = best[c];
    prob[idx] = cell_prob(d, c1, c2);
    if (dist) dist[idx] = d;
  }
}

}  // namespace

int build_grid_device(const float* d_xyz, size_t n, const GridDev& g, float* d_prob, float* d_dist,
                      double sensor_dev, hipStream_t s) {
  const float c1 = (float)(1. / (sensor_dev * std::sqrt(2 * kPi)));
  const float c2 = (float)(1. / (2. * sensor_dev * sensor_dev));
  if (n == 0) {
    const uint64_t blocks = (g.cells + 255) / 256;
    fill_empty_kernel<<<(uint32_t)blocks, 256, 0, s>>>(d_prob, d_dist, g.cells);
    A3D_HIP(hipGetLastError());
    return A3D_OK;
  }
  BucketDev b{};
  b.h = kBC * g.res;
  uint64_t nbk = 1;
  for (int a = 0; a < 3; ++a) {
    b.org[a] = g.min[a];
    b.nb[a] = (g.size[a] + kBC - 1) / kBC + 1;
    nbk *= b.nb[a];
  }
  if (nbk >= 0xfffffff0ull || n >= 0xfffffff0ull)
    return fail(A3D_ERR_ARG, "computeGrid: too many buckets or points for one build");
  uint32_t *cnt = nullptr, *start = nullptr, *key = nullptr;
  float *px = nullptr, *py = nullptr, *pz = nullptr, *cut = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = A3D_OK;
  auto cleanup = [&]() {
    (void)hipFree(cnt);
    (void)hipFree(start);
    (void)hipFree(key);
    (void)hipFree(px);
    (void)hipFree(py);
    (void)hipFree(pz);
    (void)hipFree(cut);
    (void)hipFree(tmp);
  };
#define A3D_TRY(call)              \
  if (!hip_ok((call), #call)) {    \
    cleanup();                     \
    return A3D_ERR_HIP;            \
  }
  A3D_TRY(hipMalloc(&cnt, (nbk + 1) * sizeof(uint32_t)));
  A3D_TRY(hipMalloc(&start, (nbk + 1) * sizeof(uint32_t)));
  A3D_TRY(hipMalloc(&key, n * sizeof(uint32_t)));
  A3D_TRY(hipMalloc(&px, n * sizeof(float)));
  A3D_TRY(hipMalloc(&py, n * sizeof(float)));
  A3D_TRY(hipMalloc(&pz, n * sizeof(float)));
  launch_zero_u32(cnt, (uint32_t)(nbk + 1), s);
  bucket_count_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(d_xyz, (uint32_t)n, b, cnt, key);
  A3D_TRY(hipGetLastError());
  A3D_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, start, (int)(nbk + 1), s));
  A3D_TRY(hipMalloc(&tmp, tmp_bytes));
  A3D_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, start, (int)(nbk + 1), s));
  // reuse cnt as the fill cursor
  A3D_TRY(hipMemcpyAsync(cnt, start, nbk * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  bucket_scatter_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(d_xyz, (uint32_t)n, key, cnt, px, py, pz);
  A3D_TRY(hipGetLastError());
  b.start = start;
  b.px = px;
  b.py = py;
  b.pz = pz;
  const uint32_t ntx = (g.size[0] + kT - 1) / kT, nty = (g.size[1] + kT - 1) / kT, ntz = (g.size[2] + kT - 1) / kT;
  const uint64_t ntiles = (uint64_t)ntx * nty * ntz;
  double rmax = 0.0;
  for (int a = 0; a < 3; ++a) rmax += (b.nb[a] * b.h + 1.0) * (b.nb[a] * b.h + 1.0);
  rmax = std::sqrt(rmax) * 2.0;
  if (!d_dist) {  // prob only: the search may stop where prob reaches 0
    A3D_TRY(hipMalloc(&cut, sizeof(float)));
    prob_cut_kernel<<<1, 64, 0, s>>>(c1, c2, cut);
    A3D_TRY(hipGetLastError());
  }
  const uint64_t chunk = 1ull << 22;
  for (uint64_t t0 = 0; t0 < ntiles; t0 += chunk) {
    const uint32_t nblk = (uint32_t)std::min<uint64_t>(chunk, ntiles - t0);
    grid_tile_kernel<<<nblk, kGThreads, 0, s>>>(g, b, d_prob, d_dist, c1, c2, t0, ntx, nty, rmax, cut);
    A3D_TRY(hipGetLastError());
  }
  A3D_TRY(hipStreamSynchronize(s));
#undef A3D_TRY
  cleanup();
  return rc;
}

}  // namespace a3d
