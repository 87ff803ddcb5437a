// Particle-filter kernels around the cloud weight:
//   predict   ParticleFilter.cpp:97-119   one thread per particle
//   init      ParticleFilter.cpp:46-95    one workgroup per filter
//   update    ParticleFilter.cpp:150-195  (totals, blend, normalise, mean)
//   resample  ParticleFilter.cpp:198-222  prefix (one workgroup per filter),
//                                         then select + gather per particle
//
// The reference's sums over particles are sequential float sums in index
// order (wtp, wtr, wt, the mean, the resample running sum c); the result of
// resampling depends on every bit of them.  They are computed here as
// sequential chains too: lane k of wavefront 0 runs chain k while the other
// wavefronts of the workgroup stage the next block of inputs through LDS.
// Everything that does not depend on a chain runs in parallel.
#include "a3d_math.h"

namespace a3d {
namespace {

constexpr int kTailThreads = 1024;
constexpr int kCH = 1024;  // elements per staged block

// scratch layout (floats), n = particle count
__device__ __host__ inline size_t sc_off(int k, uint32_t n) { return (size_t)k * n; }

// Sequential float sums.  Lane k < K of wavefront 0 computes
//   res[k] = (((res[k] + src[k][0]) + src[k][1]) + ...) + src[k][n-1]
// res[] lives in LDS and is read back by every thread after the call.
// With PREFIX (K == 1) the running value after each element is stored to
// prefix[i].
template <int K, bool PREFIX>
__device__ void chain_sums(const float* const* src, uint32_t n, float* res, float* stage, float* pstage,
                           float* prefix) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int nthr = blockDim.x;
  const uint32_t nblk = (n + kCH - 1) / kCH;
  float acc = (wave == 0 && lane < K) ? res[lane] : 0.f;
  if (nblk > 0) {
    for (int e = tid; e < K * kCH; e += nthr) {
      const int k = e / kCH, j = e % kCH;
      stage[k * kCH + j] = ((uint32_t)j < n) ? src[k][j] : 0.f;
    }
  }
  __syncthreads();
  for (uint32_t b = 0; b < nblk; ++b) {
    const int buf = b & 1;
    const uint32_t base = b * kCH;
    const uint32_t len = min((uint32_t)kCH, n - base);
    if (wave != 0) {
      // stage block b+1; write out the prefix of block b-1
      if (b + 1 < nblk) {
        const uint32_t nb = base + kCH;
        for (int e = tid - 64; e < K * kCH; e += nthr - 64) {
          const int k = e / kCH, j = e % kCH;
          stage[((buf ^ 1) * K + k) * kCH + j] = (nb + j < n) ? src[k][nb + j] : 0.f;
        }
      }
      if (PREFIX && b > 0) {
        const uint32_t pb = base - kCH;
        for (int j = tid - 64; j < kCH; j += nthr - 64) prefix[pb + j] = pstage[(buf ^ 1) * kCH + j];
      }
    } else if (lane < K) {
      const float* row = stage + (buf * K + lane) * kCH;
      float* prow = pstage + buf * kCH;
      uint32_t j = 0;
      for (; j + 4 <= len; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(row + j);
        acc += v.x;
        if (PREFIX) prow[j] = acc;
        acc += v.y;
        if (PREFIX) prow[j + 1] = acc;
        acc += v.z;
        if (PREFIX) prow[j + 2] = acc;
        acc += v.w;
        if (PREFIX) prow[j + 3] = acc;
      }
      for (; j < len; ++j) {
        acc += row[j];
        if (PREFIX) prow[j] = acc;
      }
    }
    __syncthreads();
  }
  if (PREFIX && nblk > 0) {
    const uint32_t lb = (nblk - 1) * kCH;
    const int buf = (nblk - 1) & 1;
    for (uint32_t j = tid; lb + j < n && j < (uint32_t)kCH; j += nthr) prefix[lb + j] = pstage[buf * kCH + j];
  }
  if (wave == 0 && lane < K) res[lane] = acc;
  __syncthreads();
}

__device__ inline const FilterDev& fsel(const Filters& fs, uint32_t i) { return fs.arr ? fs.arr[i] : fs.one; }

// ----------------------------------------------------------------- predict
__global__ void predict_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const float* z = f.z4 + 4 * (size_t)i;
  const float a = f.a[i];
  const float sa = (float)sin((double)a);
  const float ca = (float)cos((double)a);
  const float rand_x = (float)(f.delta[0] + (double)gauss_draw(z[0], f.dev[0]));
  const float rand_y = (float)(f.delta[1] + (double)gauss_draw(z[1], f.dev[1]));
  f.x[i] = f.x[i] + (ca * rand_x - sa * rand_y);
  f.y[i] = f.y[i] + (sa * rand_x + ca * rand_y);
  f.z[i] = (float)((double)f.z[i] + (f.delta[2] + (double)gauss_draw(z[2], f.dev[2])));
  f.a[i] = (float)((double)a + (f.delta[3] + (double)gauss_draw(z[3], f.dev[3])));
}

// -------------------------------------------------------------------- init
// gauss constants of ParticleFilter.cpp:54-56, computed on the host:
// f.dev[0] = gauss_const_1, f.dev[1] = gauss_const_2 (as floats in doubles),
// f.pose = initial pose, f.alpha/sigma unused; x/y/z/a deviations in
// f.delta[0..3].
__global__ __launch_bounds__(kTailThreads) void init_kernel(Filters fs) {
  __shared__ float4 stage4[2 * 4 * kCH / 4];
  __shared__ float res[8];
  float* stage = reinterpret_cast<float*>(stage4);
  const FilterDev& f = fsel(fs, blockIdx.x);
  const uint32_t n = f.n;
  const float gc1 = (float)f.dev[0];
  const float gc2 = (float)f.dev[1];
  const float x0 = f.pose[0], y0 = f.pose[1], z0 = f.pose[2], a0 = f.pose[3];
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    if (i == 0) {
      f.x[0] = x0;
      f.y[0] = y0;
      f.z[0] = z0;
      f.a[0] = a0;
      f.w[0] = gc1;
      continue;
    }
    const float* z = f.z4 + 4 * (size_t)i;
    const float xi = x0 + gauss_draw(z[0], (double)(float)f.delta[0]);
    const float yi = y0 + gauss_draw(z[1], (double)(float)f.delta[1]);
    const float zi = z0 + gauss_draw(z[2], (double)(float)f.delta[2]);
    const float ai = a0 + gauss_draw(z[3], (double)(float)f.delta[3]);
    f.x[i] = xi;
    f.y[i] = yi;
    f.z[i] = zi;
    f.a[i] = ai;
    const float dist = sqrtf((xi - x0) * (xi - x0) + (yi - y0) * (yi - y0) + (zi - z0) * (zi - z0));
    f.w[i] = (float)((double)gc1 * exp((double)(-dist * dist * gc2)));
  }
  if (threadIdx.x == 0) res[0] = 0.f;
  __syncthreads();
  const float* s1[1] = {f.w};
  chain_sums<1, false>(s1, n, res, stage, nullptr, nullptr);
  const float wt = res[0];
  float* sc = f.scratch;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const float w = f.w[i] / wt;
    f.w[i] = w;
    sc[sc_off(0, n) + i] = w * f.x[i];
    sc[sc_off(1, n) + i] = w * f.y[i];
    sc[sc_off(2, n) + i] = w * f.z[i];
    sc[sc_off(3, n) + i] = w * f.a[i];
  }
  if (threadIdx.x < 4) res[threadIdx.x] = 0.f;
  __syncthreads();
  const float* s4[4] = {sc + sc_off(0, n), sc + sc_off(1, n), sc + sc_off(2, n), sc + sc_off(3, n)};
  chain_sums<4, false>(s4, n, res, stage, nullptr, nullptr);
  if (threadIdx.x < 4) f.mean[threadIdx.x] = res[threadIdx.x];
}

// ------------------------------------------------------------- update tail
__global__ __launch_bounds__(kTailThreads) void tail_kernel(GridDev g, Filters fs) {
  __shared__ float4 stage4[2 * 4 * kCH / 4];
  __shared__ float res[8];
  float* stage = reinterpret_cast<float*>(stage4);
  const FilterDev& f = fsel(fs, blockIdx.x);
  const uint32_t n = f.n;
  float* sc = f.scratch;
  // ParticleFilter.cpp:129-153: out-of-map particles get w = 0 and add
  // nothing to wtp, wtr (adding +0.0f to a non-negative sum is exact).
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const bool in = into_map(g, f.x[i], f.y[i], f.z[i]);
    f.inmap[i] = in ? 1 : 0;
    sc[sc_off(0, n) + i] = in ? f.wp[i] : 0.f;
    sc[sc_off(1, n) + i] = in ? f.wr[i] : 0.f;
  }
  if (threadIdx.x < 2) res[threadIdx.x] = 0.f;
  __syncthreads();
  {
    const float* s2[2] = {sc + sc_off(0, n), sc + sc_off(1, n)};
    chain_sums<2, false>(s2, n, res, stage, nullptr, nullptr);
  }
  const float wtp = res[0], wtr = res[1];
  // ParticleFilter.cpp:158-180
  const double alpha = f.alpha;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const float wp = (wtp > 0) ? f.wp[i] / wtp : 0.f;
    const float wr = (wtr > 0) ? f.wr[i] / wtr : 0.f;
    f.wp[i] = wp;
    f.wr[i] = wr;
    f.w[i] = f.inmap[i] ? (float)(wp * alpha + wr * (1 - alpha)) : 0.f;
  }
  if (threadIdx.x == 0) res[0] = 0.f;
  __syncthreads();
  {
    const float* s1[1] = {f.w};
    chain_sums<1, false>(s1, n, res, stage, nullptr, nullptr);
  }
  const float wt = res[0];
  // ParticleFilter.cpp:182-195
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const float w = (wt > 0) ? f.w[i] / wt : 0.f;
    f.w[i] = w;
    sc[sc_off(0, n) + i] = w * f.x[i];
    sc[sc_off(1, n) + i] = w * f.y[i];
    sc[sc_off(2, n) + i] = w * f.z[i];
    sc[sc_off(3, n) + i] = w * f.a[i];
  }
  if (threadIdx.x < 4) res[threadIdx.x] = 0.f;
  __syncthreads();
  {
    const float* s4[4] = {sc + sc_off(0, n), sc + sc_off(1, n), sc + sc_off(2, n), sc + sc_off(3, n)};
    chain_sums<4, false>(s4, n, res, stage, nullptr, nullptr);
  }
  if (threadIdx.x < 4) f.mean[threadIdx.x] = res[threadIdx.x];
}

// ---------------------------------------------------------------- resample
// Scratch use: [0, n) float running sum c (mode 0); [4n, 10n) gathered
// x, y, z, a, wp, wr; [10n, 12n) double prefix (mode 1); then block totals.
__global__ __launch_bounds__(kTailThreads) void resample_prefix_kernel(Filters fs, int mode) {
  __shared__ float4 stage4[2 * kCH / 4];
  __shared__ float4 pstage4[2 * kCH / 4];
  __shared__ float res[8];
  const FilterDev& f = fsel(fs, blockIdx.x);
  if (f.no_resample) return;
  const uint32_t n = f.n;
  float* sc = f.scratch;
  if (mode == A3D_RESAMPLE_REFERENCE) {
    // c = p_[0].w; c += p_[i].w ...: the running values c_i (ParticleFilter.cpp:203,214)
    if (threadIdx.x == 0) res[0] = 0.f;
    __syncthreads();
    const float* s1[1] = {f.w};
    chain_sums<1, true>(s1, n, res, reinterpret_cast<float*>(stage4), reinterpret_cast<float*>(pstage4), sc);
    return;
  }
  double* c64 = reinterpret_cast<double*>(sc + sc_off(10, n));
  double* tot = c64 + n;
  const uint32_t nb = (n + 1023) / 1024;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    double s = 0.0;
    const uint32_t e = min(n, (b + 1) * 1024);
    for (uint32_t i = b * 1024; i < e; ++i) {
      s += (double)f.w[i];
      c64[i] = s;
    }
    tot[b] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double carry = 0.0;
    for (uint32_t b = 0; b < nb; ++b) {
      const double t = tot[b];
      tot[b] = carry;
      carry = carry + t;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) c64[i] = tot[i / 1024] + c64[i];
}

// First i with !(u > c[i]) (the reference's while loop stops there); the
// running sum is non-decreasing, and NaN (if any) stays NaN, so the predicate
// is monotone.  Clamped to n-1 (the reference would read p_[n]).
template <typename T, typename U>
__device__ inline uint32_t first_not_below(const T* c, uint32_t n, U u) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (!(u > c[mid]))
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo < n ? lo : n - 1;
}

__global__ void resample_select_kernel(Filters fs, int mode) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = f.n;
  if (m >= n || f.no_resample) return;
  float* sc = f.scratch;
  const float u01 = fs.u01 ? fs.u01[blockIdx.y] : f.u01;
  uint32_t i;
  if (mode == A3D_RESAMPLE_REFERENCE) {
    const float factor = 1.f / (float)n;  // 1.f / p_.size() (size_t -> float)
    const float r = factor * u01;
    const float u = r + factor * (float)m;
    i = first_not_below(sc, n, u);
  } else {
    const double* c64 = reinterpret_cast<const double*>(sc + sc_off(10, n));
    const double u = ((double)u01 + (double)m) * (1.0 / (double)n);
    i = first_not_below(c64, n, u);
  }
  if (f.idx_out) f.idx_out[m] = i;
  sc[sc_off(4, n) + m] = f.x[i];
  sc[sc_off(5, n) + m] = f.y[i];
  sc[sc_off(6, n) + m] = f.z[i];
  sc[sc_off(7, n) + m] = f.a[i];
  sc[sc_off(8, n) + m] = f.wp[i];
  sc[sc_off(9, n) + m] = f.wr[i];
}

__global__ void resample_commit_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = f.n;
  if (m >= n || f.no_resample) return;
  const float* sc = f.scratch;
  f.x[m] = sc[sc_off(4, n) + m];
  f.y[m] = sc[sc_off(5, n) + m];
  f.z[m] = sc[sc_off(6, n) + m];
  f.a[m] = sc[sc_off(7, n) + m];
  f.wp[m] = sc[sc_off(8, n) + m];
  f.wr[m] = sc[sc_off(9, n) + m];
  f.w[m] = 1.f / (float)n;
}

__global__ void gather_means_kernel(Filters fs, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fs.count) return;
  const FilterDev& f = fsel(fs, i);
  for (int k = 0; k < 4; ++k) out[4 * (size_t)i + k] = f.mean[k];
}

}  // namespace

void launch_gather_means(const Filters& fs, float* out, hipStream_t s) {
  if (fs.count == 0) return;
  gather_means_kernel<<<(fs.count + 63) / 64, 64, 0, s>>>(fs, out);
}

void launch_predict(const Filters& fs, hipStream_t s) {
  if (fs.max_n == 0) return;
  const dim3 grid((fs.max_n + 255) / 256, fs.count);
  predict_kernel<<<grid, 256, 0, s>>>(fs);
}

void launch_init(const Filters& fs, hipStream_t s) {
  if (fs.count == 0) return;
  init_kernel<<<fs.count, kTailThreads, 0, s>>>(fs);
}

void launch_update_tail(const GridDev& g, const Filters& fs, hipStream_t s) {
  if (fs.count == 0) return;
  tail_kernel<<<fs.count, kTailThreads, 0, s>>>(g, fs);
}

void launch_resample(const Filters& fs, int mode, hipStream_t s) {
  if (fs.count == 0 || fs.max_n == 0) return;
  resample_prefix_kernel<<<fs.count, kTailThreads, 0, s>>>(fs, mode);
  const dim3 grid((fs.max_n + 255) / 256, fs.count);
  resample_select_kernel<<<grid, 256, 0, s>>>(fs, mode);
  resample_commit_kernel<<<grid, 256, 0, s>>>(fs);
}

}  // namespace a3d
