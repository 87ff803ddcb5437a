// Particle-filter kernels around the cloud weight:
//   predict   ParticleFilter.cpp:97-119   one thread per particle
//   init      ParticleFilter.cpp:46-95
//   update    ParticleFilter.cpp:150-195  (totals, blend, normalise, mean)
//   resample  ParticleFilter.cpp:198-222  (running sum, select, gather)
//
// The reference's sums over particles are sequential float sums in index
// order (wtp, wtr, wt, the mean, resample's running sum c); resampling
// depends on every bit of them, so they are computed as sequential chains:
// chain_kernel runs one workgroup per filter in which lane k of wavefront 0
// adds chain k while the other fifteen wavefronts stage the next block of
// inputs through LDS.  Everything that does not depend on a chain runs in
// grid-wide elementwise kernels (grid y = filter).
//
// update:   tail_a  -> chain(wtp, wtr) -> tail_c -> chain(wt) -> tail_e
//           -> chain(mean x, y, z, a) beside chain(resample's c)
// resample: [chain(c) unless update just produced it] -> select -> commit
// init:     init_a -> chain(wt) -> init_c -> chain(mean)
#include "a3d_math.h"

namespace a3d {
namespace {

constexpr int kChainThreads = 1024;
constexpr int kCH = 2048;  // elements per staged block
constexpr int kRow = kCH + 12;  // LDS row stride: the rows of up to 5 chains start 12 banks apart
constexpr int kEw = 256;   // elementwise block size
constexpr uint32_t kGroup = 64;  // weights per running-sum checkpoint

// which chains a chain_kernel launch runs
enum ChainWhich {
  kTailTotals = 0,   // wtp, wtr over the in-map wp, wr     -> sums[0], sums[1]
  kWeightTotal = 1,  // wt over w                           -> sums[2]
  kPrefixOnly = 3,   // c over w                            -> c
  kInitMean = 4      // mean x, y, z, a                     -> mean[0..3]
};

__device__ inline const FilterDev& fsel(const Filters& fs, uint32_t i) { return fs.arr ? fs.arr[i] : fs.one; }
__device__ inline float* sums_of(const FilterDev& f) { return f.scratch + sc_sums(f.n); }

// LDS of one chain workgroup: two staging rows per chain.
template <int K, int PFX>
struct ChainLds {
  static constexpr int kFloats = 2 * K * kRow;
};

// Sequential float sums src[k][0..n) -> *dst[k] for k < K (lane k of wave 0
// runs chain k); with PFX >= 0, chain PFX's running sum at the end of every
// group of kGroup elements goes to prefix[0 .. ceil(n / kGroup)).
template <int K, int PFX>
__device__ void chain_body(uint32_t n, const float* const (&src)[K], float* const (&dst)[K], float* prefix,
                           float* lds) {
  float* stage = lds;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, nthr = blockDim.x;
  const uint32_t nblk = (n + kCH - 1) / kCH;
  float acc = 0.f;  // every reference sum starts from 0 (0 + x is x exactly)
  for (int e = tid; e < K * kCH; e += nthr) {
    const int k = e / kCH, j = e % kCH;
    stage[k * kRow + j] = ((uint32_t)j < n) ? src[k][j] : 0.f;
  }
  __syncthreads();
  for (uint32_t b = 0; b < nblk; ++b) {
    const int buf = b & 1;
    const uint32_t base = b * kCH;
    const uint32_t len = min((uint32_t)kCH, n - base);
    // Waves are dealt to the CU's four SIMDs round-robin: waves 4, 8, 12
    // share wave 0's SIMD and stay idle, so the chain's dependent adds never
    // wait for another wave's instructions.  The other 12 waves stage.
    if ((wave & 3) != 0) {
      const int st = ((wave >> 2) * 3 + (wave & 3) - 1) * 64 + lane;  // 0 .. kStagers-1
      constexpr int kStagers = (kChainThreads / 64 / 4) * 3 * 64;
      if (b + 1 < nblk) {
        // every load of the next block is issued before any is stored, so
        // the block costs one memory round trip
        const uint32_t nb = base + kCH;
        constexpr int kR = (K * kCH + kStagers - 1) / kStagers;
        float t[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const int e = st + r * kStagers;
          const int k = e / kCH, j = e % kCH;
          t[r] = 0.f;
          if (e < K * kCH && nb + j < n) t[r] = src[k][nb + j];
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const int e = st + r * kStagers;
          if (e < K * kCH) stage[((buf ^ 1) * K + e / kCH) * kRow + e % kCH] = t[r];
        }
      }
    } else if (wave == 0 && lane < K) {
      const float4* row4 = reinterpret_cast<const float4*>(stage + (buf * K + lane) * kRow);
      // The prefix lane keeps one checkpoint per 64 elements, the running
      // sum at the group's end (a store per element would stall the chain);
      // resample re-adds a group's weights from the previous checkpoint,
      // which reproduces every running sum bit for bit.
      const uint32_t n4 = len >> 2;
      uint32_t q = 0;
      for (; q + 16 <= n4; q += 16) {
        float4 v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = row4[q + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          acc += v[u].x;
          acc += v[u].y;
          acc += v[u].z;
          acc += v[u].w;
        }
        if (PFX >= 0 && lane == PFX) prefix[(base + q * 4) / kGroup] = acc;
      }
      const float* row = reinterpret_cast<const float*>(row4);
      for (uint32_t j = q * 4; j < len; ++j) acc += row[j];
      if (PFX >= 0 && lane == PFX && q * 4 < len) prefix[(base + len - 1) / kGroup] = acc;  // last, partial group
    }
    __syncthreads();
  }
  if (wave == 0 && lane < K) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (lane == k) *dst[k] = acc;
  }
}

template <int K, int PFX>
__global__ __launch_bounds__(kChainThreads) void chain_kernel(Filters fs, int which) {
  __shared__ float4 lds4[ChainLds<K, PFX>::kFloats / 4];
  const FilterDev& f = fsel(fs, blockIdx.x);
  if (which == kPrefixOnly && f.no_resample) return;
  const uint32_t n = f.n;
  float* sc = f.scratch;
  const float* src[K];
  float* dst[K];
  float* prefix = nullptr;
  if (which == kTailTotals) {
    src[0] = sc;
    src[1 % K] = sc + n;
    dst[0] = sums_of(f);
    dst[1 % K] = sums_of(f) + 1;
  } else if (which == kWeightTotal || which == kPrefixOnly) {
    src[0] = f.w;
    dst[0] = (which == kWeightTotal) ? sums_of(f) + 2 : sums_of(f) + 3;
  } else {  // kInitMean
#pragma unroll
    for (int k = 0; k < K; ++k) {
      src[k] = sc + (size_t)k * n;
      dst[k] = f.mean + k;
    }
  }
  if (PFX >= 0) prefix = sc + sc_c32(n);
  chain_body<K, PFX>(n, src, dst, prefix, reinterpret_cast<float*>(lds4));
}

// The last chains of update: the four mean terms (workgroup row 0) and
// resample's running sum c over the new weights (row 1) run side by side.
__global__ __launch_bounds__(kChainThreads) void mean_prefix_kernel(Filters fs) {
  __shared__ float4 lds4[ChainLds<4, -1>::kFloats / 4];
  static_assert(ChainLds<1, 0>::kFloats <= ChainLds<4, -1>::kFloats, "LDS of the prefix chain");
  const FilterDev& f = fsel(fs, blockIdx.x);
  const uint32_t n = f.n;
  float* sc = f.scratch;
  if (blockIdx.y == 0) {
    const float* const src[4] = {sc, sc + n, sc + 2 * (size_t)n, sc + 3 * (size_t)n};
    float* const dst[4] = {f.mean, f.mean + 1, f.mean + 2, f.mean + 3};
    chain_body<4, -1>(n, src, dst, nullptr, reinterpret_cast<float*>(lds4));
  } else {
    const float* const src[1] = {f.w};
    float* const dst[1] = {sums_of(f) + 3};
    chain_body<1, 0>(n, src, dst, sc + sc_c32(n), reinterpret_cast<float*>(lds4));
  }
}

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
// gauss constants of ParticleFilter.cpp:54-56 come from the host in f.dev[0]
// (gauss_const_1) and f.dev[1] (gauss_const_2); deviations in f.delta[0..3].
__global__ void init_a_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const float gc1 = (float)f.dev[0];
  const float gc2 = (float)f.dev[1];
  const float x0 = f.pose[0], y0 = f.pose[1], z0 = f.pose[2], a0 = f.pose[3];
  if (i == 0) {
    f.x[0] = x0;
    f.y[0] = y0;
    f.z[0] = z0;
    f.a[0] = a0;
    f.w[0] = gc1;
    return;
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

// w /= wt (float division, ParticleFilter.cpp:85) and the weighted terms of the mean
__global__ void init_c_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const uint32_t n = f.n;
  const float wt = sums_of(f)[2];
  const float w = f.w[i] / wt;
  f.w[i] = w;
  float* sc = f.scratch;
  sc[i] = w * f.x[i];
  sc[n + i] = w * f.y[i];
  sc[2 * (size_t)n + i] = w * f.z[i];
  sc[3 * (size_t)n + i] = w * f.a[i];
}

// ------------------------------------------------------------- update tail
// ParticleFilter.cpp:129-153: out-of-map particles get w = 0 and add nothing
// to wtp, wtr (adding +0.0f to a non-negative sum is exact).
__global__ void tail_a_kernel(GridDev g, Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const bool in = into_map(f.gdev ? *f.gdev : g, f.x[i], f.y[i], f.z[i]);
  f.inmap[i] = in ? 1 : 0;
  f.scratch[i] = in ? f.wp[i] : 0.f;
  f.scratch[f.n + i] = in ? f.wr[i] : 0.f;
}

// ParticleFilter.cpp:158-180
__global__ void tail_c_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const float* s = sums_of(f);
  const float wtp = s[0], wtr = s[1];
  const float wp = (wtp > 0) ? f.wp[i] / wtp : 0.f;
  const float wr = (wtr > 0) ? f.wr[i] / wtr : 0.f;
  f.wp[i] = wp;
  f.wr[i] = wr;
  f.w[i] = f.inmap[i] ? (float)(wp * f.alpha + wr * (1 - f.alpha)) : 0.f;
}

// ParticleFilter.cpp:182-193: w /= wt and the weighted terms of the mean
__global__ void tail_e_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const uint32_t n = f.n;
  const float wt = sums_of(f)[2];
  const float w = (wt > 0) ? f.w[i] / wt : 0.f;
  f.w[i] = w;
  float* sc = f.scratch;
  sc[i] = w * f.x[i];
  sc[n + i] = w * f.y[i];
  sc[2 * (size_t)n + i] = w * f.z[i];
  sc[3 * (size_t)n + i] = w * f.a[i];
}

// ---------------------------------------------------------------- resample
// FP64 prefix: sequential double sums in blocks of 1024, block totals summed
// in order (one workgroup per filter).
__global__ __launch_bounds__(1024) void prefix64_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.x);
  if (f.no_resample) return;
  const uint32_t n = f.n;
  double* c64 = reinterpret_cast<double*>(f.scratch + sc_c64(n));
  double* tot = reinterpret_cast<double*>(f.scratch + sc_tot(n));
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
    // the group holding the first running sum c_i with !(u > c_i), from the
    // checkpoints; then c_i itself, re-added in the reference's order
    const float* cp = sc + sc_c32(n);
    const uint32_t ng = (n + kGroup - 1) / kGroup;
    uint32_t gi = 0, hi = ng;
    while (gi < hi) {
      const uint32_t mid = gi + ((hi - gi) >> 1);
      if (!(u > cp[mid]))
        hi = mid;
      else
        gi = mid + 1;
    }
    if (gi == ng) {
      i = n - 1;  // the float total ends below u: clamped (the reference reads p_[n])
    } else {
      float c = gi ? cp[gi - 1] : 0.0f;  // 0 + w_0 is w_0 exactly
      const uint32_t e = min(n, (gi + 1) * kGroup);
      i = e - 1;
      for (uint32_t k = gi * kGroup; k < e; ++k) {
        c += f.w[k];
        if (!(u > c)) {
          i = k;
          break;
        }
      }
    }
  } else {
    const double* c64 = reinterpret_cast<const double*>(sc + sc_c64(n));
    const double u = ((double)u01 + (double)m) * (1.0 / (double)n);
    i = first_not_below(c64, n, u);
  }
  if (f.idx_out) f.idx_out[m] = i;
  sc[4 * (size_t)n + m] = f.x[i];
  sc[5 * (size_t)n + m] = f.y[i];
  sc[6 * (size_t)n + m] = f.z[i];
  sc[7 * (size_t)n + m] = f.a[i];
  sc[8 * (size_t)n + m] = f.wp[i];
  sc[9 * (size_t)n + m] = f.wr[i];
}

__global__ void resample_commit_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = f.n;
  if (m >= n || f.no_resample) return;
  const float* sc = f.scratch;
  f.x[m] = sc[4 * (size_t)n + m];
  f.y[m] = sc[5 * (size_t)n + m];
  f.z[m] = sc[6 * (size_t)n + m];
  f.a[m] = sc[7 * (size_t)n + m];
  f.wp[m] = sc[8 * (size_t)n + m];
  f.wr[m] = sc[9 * (size_t)n + m];
  f.w[m] = 1.f / (float)n;
}

__global__ void gather_means_kernel(Filters fs, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fs.count) return;
  const FilterDev& f = fsel(fs, i);
  for (int k = 0; k < 4; ++k) out[4 * (size_t)i + k] = f.mean[k];
}

inline dim3 ew_grid(const Filters& fs) { return dim3((fs.max_n + kEw - 1) / kEw, fs.count); }

}  // namespace

void launch_gather_means(const Filters& fs, float* out, hipStream_t s) {
  if (fs.count == 0) return;
  gather_means_kernel<<<(fs.count + 63) / 64, 64, 0, s>>>(fs, out);
}

void launch_predict(const Filters& fs, hipStream_t s) {
  if (fs.max_n == 0) return;
  predict_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
}

void launch_init(const Filters& fs, hipStream_t s) {
  if (fs.count == 0 || fs.max_n == 0) return;
  init_a_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  chain_kernel<1, -1><<<fs.count, kChainThreads, 0, s>>>(fs, kWeightTotal);
  init_c_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  chain_kernel<4, -1><<<fs.count, kChainThreads, 0, s>>>(fs, kInitMean);
}

void launch_update_tail(const GridDev& g, const Filters& fs, hipStream_t s) {
  if (fs.count == 0 || fs.max_n == 0) return;
  tail_a_kernel<<<ew_grid(fs), kEw, 0, s>>>(g, fs);
  chain_kernel<2, -1><<<fs.count, kChainThreads, 0, s>>>(fs, kTailTotals);
  tail_c_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  chain_kernel<1, -1><<<fs.count, kChainThreads, 0, s>>>(fs, kWeightTotal);
  tail_e_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  // the mean and, in a second workgroup per filter, resample's running sum c
  mean_prefix_kernel<<<dim3(fs.count, 2), kChainThreads, 0, s>>>(fs);
}

void launch_resample(const Filters& fs, int mode, bool have_prefix, hipStream_t s) {
  if (fs.count == 0 || fs.max_n == 0) return;
  if (mode == A3D_RESAMPLE_REFERENCE) {
    if (!have_prefix) chain_kernel<1, 0><<<fs.count, kChainThreads, 0, s>>>(fs, kPrefixOnly);
  } else {
    prefix64_kernel<<<fs.count, 1024, 0, s>>>(fs);
  }
  resample_select_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs, mode);
  resample_commit_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
}

}  // namespace a3d
