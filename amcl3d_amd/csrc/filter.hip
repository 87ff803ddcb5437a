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
  kInitMean = 4,     // mean x, y, z, a                     -> mean[0..3]
  kMeanPrefix = 5    // mean x, y, z, a and c (tiled sums)  -> mean[0..3], c
};

__device__ inline const FilterDev& fsel(const Filters& fs, uint32_t i) { return fs.arr ? fs.arr[i] : fs.one; }
__device__ inline float* sums_of(const FilterDev& f) { return f.scratch + sc_sums(f.n); }

// LDS of one chain workgroup: two staging rows per chain.
template <int K, int PFX>
struct ChainLds {
  static constexpr int kFloats = 2 * K * kRow;
};

// Sequential float sums acc0 + src[k][0] + ... + src[k][n-1] -> *dst[k] for
// k < K (lane k of wave 0 runs chain k); with PFX >= 0, chain PFX's running
// sum at the end of every group of kGroup elements goes to prefix[0 ..
// ceil(n / kGroup)).  Every reference sum starts from 0 (0 + x is x exactly).
template <int K, int PFX>
__device__ __forceinline__ void chain_body(uint32_t n, const float* const (&src)[K], float* const (&dst)[K],
                                           float* prefix, float* lds, float acc0 = 0.f) {
  float* stage = lds;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, nthr = blockDim.x;
  const uint32_t nblk = (n + kCH - 1) / kCH;
  float acc = acc0;
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

// ----------------------------------------------- exact sums in parallel
// The sums over non-negative terms (wtp, wtr, wt, resample's running sum c)
// are computed in parallel with the sequential sum's exact result.  While
// the running sum s stays in one float binade with spacing u, s = k*u and
// fl(s + x) = (k + t)*u, where t is x/u rounded to the nearest integer such
// that ties make k + t even: t depends on x and u only, except at exact ties
// where it depends on the parity of k.  A tile of terms is therefore one
// parallel scan of a two-state map (parity in -> increment, parity out).
// Each step is checked, with a conservative test in double, to stay in the
// binade (the exact sum strictly inside it); the first step that may leave it
// is done as the reference's float add, and the scan restarts after it with
// the new binade.  A sum of non-negative terms crosses each binade once, so
// that happens a few dozen times per sum, mostly within its first terms,
// which are added serially.  Runs of zeros (s = +0: the sums start at +0 and
// never reach -0) are skipped by a search for the first non-zero term.  When
// restarts come densely (terms of mixed sign can cross binades all the time)
// or s is not finite, the workgroup finishes the sum serially from LDS,
// still in the reference's order.
constexpr int kExactPer = 10;                                     // terms per thread per tile
constexpr uint32_t kExactTile = kChainThreads * kExactPer;
// Restarts that each advanced fewer than kExactDenseRun terms: after
// kExactDense of them in a row (terms of mixed sign crossing binades all the
// time) the rest of the sum is added serially.
constexpr int kExactDense = 4;
constexpr uint32_t kExactDenseRun = 1024;
// A growing sum crosses a binade each time it doubles, mostly within its
// first terms: those are added serially, one restart per crossing would cost
// more.
constexpr uint32_t kExactHead = 1024;

struct ExactLds {
  uint32_t d0[kChainThreads / 64], d1[kChainThreads / 64], pp[kChainThreads / 64];
  uint32_t first[kChainThreads / 64];
  uint32_t kfirst[kChainThreads / 64];
  uint32_t klast;
  float s;
  float stage[kExactTile];
};

// parity in {0, 1} -> (increment of k mod 2^32, parity out); fn_then(a, b)
// applies a, then b
struct ParityFn {
  uint32_t d0, d1, p0, p1;
};

__device__ inline ParityFn fn_then(const ParityFn& a, const ParityFn& b) {
  ParityFn r;
  r.d0 = a.d0 + (a.p0 ? b.d1 : b.d0);
  r.p0 = a.p0 ? b.p1 : b.p0;
  r.d1 = a.d1 + (a.p1 ? b.d1 : b.d0);
  r.p1 = a.p1 ? b.p1 : b.p0;
  return r;
}

__device__ inline ParityFn shfl_up_fn(const ParityFn& f, int d) {
  ParityFn r;
  r.d0 = (uint32_t)__shfl_up((int)f.d0, d);
  r.d1 = (uint32_t)__shfl_up((int)f.d1, d);
  r.p0 = (uint32_t)__shfl_up((int)f.p0, d);
  r.p1 = (uint32_t)__shfl_up((int)f.p1, d);
  return r;
}

// DPP lane moves of a map (GFX9 row_shr / row_bcast controls): a lane with
// no source lane, or in a row the row mask leaves out, takes the identity.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ ParityFn dpp_fn(const ParityFn& f) {
  return ParityFn{dpp_u32<CTRL, ROW_MASK>(0u, f.d0), dpp_u32<CTRL, ROW_MASK>(0u, f.d1),
                  dpp_u32<CTRL, ROW_MASK>(0u, f.p0), dpp_u32<CTRL, ROW_MASK>(1u, f.p1)};
}

// Inclusive composition within each row of 16 lanes (shifts by 1, 2, 4, 8).
__device__ __forceinline__ ParityFn row_inclusive_fn(ParityFn f) {
  f = fn_then(dpp_fn<0x111, 0xf>(f), f);
  f = fn_then(dpp_fn<0x112, 0xf>(f), f);
  f = fn_then(dpp_fn<0x114, 0xf>(f), f);
  f = fn_then(dpp_fn<0x118, 0xf>(f), f);
  return f;
}

// The composition of the maps of all threads before this one (thread order):
// rows of 16 lanes by DPP shifts, the four rows of a wavefront by row
// broadcasts (lane 15 into rows 1 and 3, then lane 31 into rows 2 and 3),
// then the 16 wavefront totals, composed by wavefront 0 in one row.
template <typename Lds>
__device__ __forceinline__ ParityFn block_exclusive_fn(ParityFn f, Lds& L) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int kW = kChainThreads / 64;
  static_assert(kW == 16, "the wavefront totals are composed in one row of 16 lanes");
  const ParityFn id{0u, 0u, 0u, 1u};
  ParityFn inc = row_inclusive_fn(f);
  inc = fn_then(dpp_fn<0x142, 0xa>(inc), inc);
  inc = fn_then(dpp_fn<0x143, 0xc>(inc), inc);
  ParityFn ex = shfl_up_fn(inc, 1);
  if (lane == 0) ex = id;
  if (lane == 63) {
    L.d0[wave] = inc.d0;
    L.d1[wave] = inc.d1;
    L.pp[wave] = inc.p0 | (inc.p1 << 1);
  }
  __syncthreads();
  if (wave == 0) {  // exclusive composition of the wavefront totals, in place
    ParityFn t = id;
    if (lane < kW) t = ParityFn{L.d0[lane], L.d1[lane], L.pp[lane] & 1u, L.pp[lane] >> 1};
    const ParityFn te = dpp_fn<0x111, 0xf>(row_inclusive_fn(t));
    if (lane < kW) {
      L.d0[lane] = te.d0;
      L.d1[lane] = te.d1;
      L.pp[lane] = te.p0 | (te.p1 << 1);
    }
  }
  __syncthreads();
  const ParityFn w{L.d0[wave], L.d1[wave], L.pp[wave] & 1u, L.pp[wave] >> 1};
  return fn_then(w, ex);
}

__device__ __forceinline__ uint32_t block_min(uint32_t v, ExactLds& L) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, d));
  if (lane == 0) L.first[wave] = v;
  __syncthreads();
  uint32_t m = L.first[0];
  for (int w = 1; w < kChainThreads / 64; ++w) m = min(m, L.first[w]);
  __syncthreads();
  return m;
}

__device__ inline bool is_checkpoint(uint32_t i, uint32_t n) { return (i % kGroup) == kGroup - 1 || i == n - 1; }

// a + stage[0] + stage[1] + ... in order (one lane), for the terms of global
// indices base .. base + len - 1, with their checkpoints.  The terms go in
// groups of kGroup aligned to the checkpoint groups, each group's LDS reads
// issued together, so the adds wait on each other only.
__device__ __forceinline__ float serial_add(float a, const float* stage, uint32_t base, uint32_t len, uint32_t n, float* prefix) {
  uint32_t j = 0;
  const uint32_t lead = min(len, (kGroup - base % kGroup) % kGroup);
  for (; j < lead; ++j) {
    a += stage[j];
    if (prefix && is_checkpoint(base + j, n)) prefix[(base + j) / kGroup] = a;
  }
  for (; j + kGroup <= len; j += kGroup) {
    float v[kGroup];
#pragma unroll
    for (uint32_t u = 0; u < kGroup; ++u) v[u] = stage[j + u];
#pragma unroll
    for (uint32_t u = 0; u < kGroup; ++u) a += v[u];
    if (prefix) prefix[(base + j) / kGroup] = a;  // base + j + kGroup - 1 ends a group
  }
  for (; j < len; ++j) {
    a += stage[j];
    if (prefix && is_checkpoint(base + j, n)) prefix[(base + j) / kGroup] = a;
  }
  return a;
}

// s + x[pos] + ... + x[end-1] in order (n: the sum's length, for its last
// checkpoint): lane 0 adds singly up to the next checkpoint-group boundary,
// then the serial chain (chain_body) runs the rest.
__device__ __forceinline__ float serial_rest(float s, uint32_t pos, uint32_t end, uint32_t n,
                                             const float* __restrict__ x, float* prefix, ExactLds& L) {
  static_assert(ChainLds<1, 0>::kFloats <= kExactTile, "the chain's staging fits the tile buffer");
  const uint32_t pa = min(end, (pos + kGroup - 1) / kGroup * kGroup);
  if (threadIdx.x == 0) L.s = serial_add(s, x + pos, pos, pa - pos, n, prefix);
  __syncthreads();
  s = L.s;
  __syncthreads();
  if (pa < end) {
    const float* const src[1] = {x + pa};
    float* const dst[1] = {&L.s};
    if (prefix)
      chain_body<1, 0>(end - pa, src, dst, prefix + pa / kGroup, L.stage, s);
    else
      chain_body<1, -1>(end - pa, src, dst, nullptr, L.stage, s);
    __syncthreads();
    s = L.s;
    __syncthreads();
  }
  return s;
}

// The sequential float sum s + x[pos] + ... + x[end-1], exactly (n: the
// whole sum's length); with prefix, the running sum after element i is
// stored at prefix[i / kGroup] at the end of every group of kGroup elements
// and after element n-1.  A range that starts at 0 (from s = +0) adds its
// first kExactHead terms serially.  Called by every thread of a
// kChainThreads workgroup; returns the sum in every thread.
__device__ __forceinline__ float exact_sum_range(float s, uint32_t pos, uint32_t end, uint32_t n,
                                                 const float* __restrict__ x, float* prefix, ExactLds& L) {
  const uint32_t tid = threadIdx.x;
  int dense = 0;
  if (pos == 0) {  // the serial head
    const uint32_t hend = min(end, kExactHead);
    for (uint32_t i = tid; i < hend; i += blockDim.x) L.stage[i] = x[i];
    __syncthreads();
    if (tid == 0) L.s = serial_add(s, L.stage, 0, hend, n, prefix);
    __syncthreads();
    s = L.s;
    pos = hend;
    __syncthreads();
  }
  while (pos < end) {  // uniform over the workgroup
    const uint32_t tend = min(end, pos + kExactTile);
    const uint32_t j0 = pos + tid * kExactPer;
    float xv[kExactPer];
#pragma unroll
    for (int e = 0; e < kExactPer; ++e) xv[e] = (j0 + e < tend) ? x[j0 + e] : 0.0f;
    if (dense >= kExactDense || !isfinite(s)) {
      s = serial_rest(s, pos, end, n, x, prefix, L);
      pos = end;
      continue;
    }
    if (s == 0.0f) {  // +0: the sum is the first non-zero term
      uint32_t first = tend;
#pragma unroll
      for (int e = kExactPer - 1; e >= 0; --e)
        if (j0 + e < tend && xv[e] != 0.0f) first = j0 + e;
      first = block_min(first, L);
      if (prefix) {
#pragma unroll
        for (int e = 0; e < kExactPer; ++e)
          if (j0 + e < first && is_checkpoint(j0 + e, n)) prefix[(j0 + e) / kGroup] = 0.0f;
      }
      if (first < tend) {
        s = x[first];  // 0 + x is x exactly
        if (prefix && tid == 0 && is_checkpoint(first, n)) prefix[first / kGroup] = s;
        pos = first + 1;
      } else {
        pos = tend;
      }
      continue;
    }
    // the binade of s: s = k*u, and a step is regular when the exact s + x
    // lies strictly inside (lo, hi) (or |s + x| < 2^-125 on the subnormal grid)
    const float as = fabsf(s);
    const bool sub = as < 0x1p-125f;
    const int ex = sub ? -126 : (int)((__float_as_uint(as) >> 23) & 0xffu) - 127;
    const double u = sub ? 0x1p-149 : ldexp(1.0, ex - 23), inv_u = sub ? 0x1p149 : ldexp(1.0, 23 - ex);
    const double lo = ldexp(1.0, ex), hi = ldexp(1.0, ex + 1);
    const bool neg = s < 0.0f;
    const int k0 = (int)((double)s * inv_u);  // |k0| <= 2^24
    ParityFn fe[kExactPer];
    bool hard[kExactPer];
    ParityFn mine{0u, 0u, 0u, 1u};
#pragma unroll
    for (int e = 0; e < kExactPer; ++e) {
      const double r = (double)xv[e] * inv_u;  // exact: a power-of-two scale
      hard[e] = !(fabs(r) < 0x1p25);           // leaves the binade whatever k is (or non-finite)
      ParityFn f{0u, 0u, 0u, 1u};
      if (!hard[e]) {
        const double qd = floor(r), fr = r - qd;
        const int q = (int)qd;
        if (fr == 0.5) {  // tie: k + t even
          f.d0 = (uint32_t)(q + (q & 1));
          f.d1 = (uint32_t)(q + ((q + 1) & 1));
          f.p0 = 0u;
          f.p1 = 0u;
        } else {
          const int t = q + (fr > 0.5 ? 1 : 0);
          f.d0 = f.d1 = (uint32_t)t;
          f.p0 = (uint32_t)t & 1u;
          f.p1 = ((uint32_t)t + 1u) & 1u;
        }
      }
      fe[e] = f;
      mine = fn_then(mine, f);
    }
    const ParityFn pre = block_exclusive_fn(mine, L);
    // k as a 32-bit word: exact (|k| <= 2^25) up to the first irregular
    // step; past it, wrapped garbage that is never used
    const uint32_t p0 = (uint32_t)k0 & 1u;
    uint32_t k = (uint32_t)k0 + (p0 ? pre.d1 : pre.d0);
    uint32_t par = p0 ? pre.p1 : pre.p0;
    uint32_t bad = tend;
    uint32_t kbad = 0u;  // k before the first irregular step of this thread
    uint32_t kv[kExactPer];
#pragma unroll
    for (int e = 0; e < kExactPer; ++e) {
      kv[e] = k;
      if (j0 + e >= tend || bad < tend) continue;
      const uint32_t kn = k + (par ? fe[e].d1 : fe[e].d0);
      const double vd = (double)(int)k * u + (double)xv[e];  // the exact s + x, rounded to double
      bool reg;
      if (sub)
        reg = fabs(vd) < 0x1p-125 && kn != 0u;
      else
        reg = neg ? (vd < -lo && vd > -hi) : (vd > lo && vd < hi);
      reg = !hard[e] && (reg || xv[e] == 0.0f);  // s + 0 = s
      if (!reg) {
        bad = j0 + e;
        kbad = k;
        continue;
      }
      k = kn;
      par = kn & 1u;
      kv[e] = kn;
    }
    // c, the first irregular step of the tile (tend if none), and k before
    // it: per wavefront the first lane with one, then the first such
    // wavefront; with none, k after the tile's last term
    uint32_t c = tend, kc = 0u;
    {
      const int lane = tid & 63, wave = tid >> 6;
      const uint64_t mb = __ballot(bad != tend);
      const int src = mb ? (int)__builtin_ctzll(mb) : 0;
      const uint32_t fb = (uint32_t)__builtin_amdgcn_readlane((int)bad, src);
      const uint32_t fk = (uint32_t)__builtin_amdgcn_readlane((int)kbad, src);
      if (lane == 0) {
        L.first[wave] = mb ? fb : tend;
        L.kfirst[wave] = fk;
      }
      if (tid == (tend - 1 - pos) / kExactPer) L.klast = k;
      __syncthreads();
      // the first of the 16 wavefronts with an irregular step (lane w reads
      // wavefront w's entry)
      const uint32_t fw = L.first[lane & 15];
      const uint64_t mw = __ballot(fw != tend) & 0xffffull;
      if (mw) {
        const int w = (int)__builtin_ctzll(mw);
        c = (uint32_t)__builtin_amdgcn_readlane((int)fw, w);
        kc = L.kfirst[w];
      } else {
        kc = L.klast;
      }
      __syncthreads();
    }
    // the terms before c are exact: s_i = k_i * u
    if (prefix) {
#pragma unroll
      for (int e = 0; e < kExactPer; ++e)
        if (j0 + e < c && is_checkpoint(j0 + e, n)) prefix[(j0 + e) / kGroup] = (float)((double)(int)kv[e] * u);
    }
    const float sp = (float)((double)(int)kc * u);
    if (c < tend) {
      s = sp + x[c];  // the step that may leave the binade: the reference's float add
      if (prefix && tid == 0 && is_checkpoint(c, n)) prefix[c / kGroup] = s;
      dense = (c - pos < kExactDenseRun) ? dense + 1 : 0;
      pos = c + 1;
    } else {
      s = sp;
      pos = tend;
    }
  }
  return s;
}

__device__ __forceinline__ float exact_sum(uint32_t n, const float* __restrict__ x, float* prefix, ExactLds& L) {
  return exact_sum_range(0.0f, 0, n, n, x, prefix, L);
}

// Whether any of w[0..n) is negative or not finite: the running sum c is
// then not monotone, and resample selects serially (the flag goes to
// sums[4]).  Called by every thread of the workgroup.
__device__ __forceinline__ float weights_irregular(uint32_t n, const float* w, ExactLds& L) {
  uint32_t bad = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) bad |= (w[i] < 0.0f || !isfinite(w[i])) ? 1u : 0u;
  return block_min(bad ? 0u : 1u, L) == 0u ? 1.0f : 0.0f;
}

// wtp and wtr (grid row y = 0, 1), wt, or resample's c with its checkpoints,
// by exact_sum: one workgroup per filter and sum.
__global__ __launch_bounds__(kChainThreads) void exact_chain_kernel(Filters fs, int which) {
  __shared__ ExactLds lds;
  const FilterDev& f = fsel(fs, blockIdx.x);
  if (which == kPrefixOnly && f.no_resample) return;
  const uint32_t n = f.n;
  float* sc = f.scratch;
  const float* src;
  float* dst;
  float* prefix = nullptr;
  if (which == kTailTotals) {
    src = sc + (size_t)blockIdx.y * n;
    dst = sums_of(f) + blockIdx.y;
  } else if (which == kWeightTotal) {
    src = f.w;
    dst = sums_of(f) + 2;
  } else {  // kPrefixOnly
    src = f.w;
    dst = sums_of(f) + 3;
    prefix = sc + sc_c32(n);
  }
  const float v = exact_sum(n, src, prefix, lds);
  const float irregular = prefix ? weights_irregular(n, src, lds) : 0.0f;
  if (threadIdx.x == 0) {
    *dst = v;
    if (prefix) sums_of(f)[4] = irregular;
  }
}

// The last sums of update, side by side: the four mean terms (grid rows
// y = 0..3) and resample's running sum c over the new weights with its
// checkpoints and the monotonicity flag (row 4).  The mean terms may have
// either sign: exact_sum handles them exactly, restarting at every binade
// crossing, and finishes serially if they cross too often.
__global__ __launch_bounds__(kChainThreads) void mean_prefix_kernel(Filters fs) {
  __shared__ ExactLds lds;
  const FilterDev& f = fsel(fs, blockIdx.x);
  const uint32_t n = f.n;
  float* sc = f.scratch;
  if (blockIdx.y < 4) {
    const float v = exact_sum(n, sc + (size_t)blockIdx.y * n, nullptr, lds);
    if (threadIdx.x == 0) f.mean[blockIdx.y] = v;
  } else {
    const float v = exact_sum(n, f.w, sc + sc_c32(n), lds);
    const float irregular = weights_irregular(n, f.w, lds);
    if (threadIdx.x == 0) {
      sums_of(f)[3] = v;
      sums_of(f)[4] = irregular;
    }
  }
}

// ------------------------------------------------------- tiled exact sums
// A single large filter's sums, spread over the chip instead of one CU.  The
// terms are cut into tiles of kTiledTile, each of 16 sub-tiles of 64 (one
// wavefront, one resample checkpoint group).  While a running sum s stays in
// one binade of spacing u (s = k*u), a run of terms acts on k as a two-state
// map (parity in -> increment, parity out) that depends on the terms and u
// only (the map of exact_sum, composed over the run).  So:
//   dsum  one workgroup per tile: the tile's terms summed in double, and
//         their signs (or the kernel that writes the terms does it);
//   map   one workgroup per tile: from the double sums of the earlier terms,
//         guesses the binade s is in where the tile starts, and where each
//         sub-tile starts; composes the tile's map under the tile's guess,
//         and each sub-tile's maps under its guess and the binade above;
//   walk  one workgroup per sum, from s = +0.  Wavefront 0 takes a run of
//         up to 64 tiles whose maps apply (s in the guessed binade, every
//         term of the sign of s, k below 2^24 at the run's end) in one
//         wavefront scan; a tile it cannot take whole, sub-tile by
//         sub-tile: runs of sub-tile maps the same way, any other sub-tile
//         as 64 float adds in order; a tile with terms of both signs, or of
//         the other sign than s, and a sum that is not finite go to
//         exact_sum_range, the one-CU algorithm, run by the whole workgroup
//         from the exact s;
//   ckpt  for resample's running sum c: the checkpoints (c after every 64th
//         weight) of the tiles the walk took whole, from the exact s it
//         recorded at their start (the walk writes the others').
// The guesses only decide which path a run of terms takes, never a bit of
// the result: every sum is the sequential float sum exactly.
constexpr uint32_t kTiledMin = 16384;  // particles from which a single filter's sums run tiled
constexpr uint32_t kWalkTiles = 1024;  // tile records the walk holds in LDS at a time
constexpr uint32_t kSubTiles = kTiledTile / 64;
static_assert(kSubTiles == 16, "a tile's sub-tiles fit one row of 16 lanes");

struct TileRec {
  double dsum;          // dsum: the tile's terms summed in double
  double asum;          // dsum: their magnitudes summed in double
  uint32_t flags;       // dsum: kPos, kNeg, kIrregularTerm; map: kNoMap
  uint32_t d0, d1, pp;  // map: the tile's map under binade e (pp = p0 | p1 << 1)
  int32_t e;            // map: the binade exponent the map is for
  uint32_t regular;     // walk: the tile was added whole from its map
  float s_in;           // walk: the exact running sum before the tile
  uint32_t pad[5];
};
static_assert(sizeof(TileRec) == 64, "16 floats per tile record");
// A sub-tile's maps under its guessed binade e (a) and e + 1 (b).
struct SubRec {
  uint32_t d0a, d1a, d0b, d1b;
  uint32_t pp;  // p0a | p1a << 1 | p0b << 2 | p1b << 3
  int32_t e;
  uint32_t flags;  // kNoMapA, kNoMapB, kPos, kNeg, kIrregularTerm (a NaN term)
  uint32_t pad;
};
static_assert(sizeof(SubRec) == 32, "8 floats per sub-tile record");
enum TileFlags : uint32_t { kNoMap = 1u, kPos = 2u, kNeg = 4u, kIrregularTerm = 8u, kNoMapA = 16u, kNoMapB = 32u };

__device__ inline TileRec* tiled_recs(const FilterDev& f, uint32_t k) {
  return reinterpret_cast<TileRec*>(f.scratch + sc_tiles(f.n)) + (size_t)k * tiled_tiles(f.n);
}
__device__ inline SubRec* tiled_subs(const FilterDev& f, uint32_t k) {
  const size_t T = tiled_tiles(f.n);
  return reinterpret_cast<SubRec*>(f.scratch + sc_tiles(f.n) + 16 * kTiledSums * T) + (size_t)k * T * kSubTiles;
}

// Sum k of a tiled launch: its terms, and where its result goes.
__device__ inline const float* tiled_src(const FilterDev& f, int which, uint32_t k) {
  if (which == kTailTotals || (which == kMeanPrefix && k < 4)) return f.scratch + (size_t)k * f.n;
  return f.w;
}
__device__ inline bool tiled_has_prefix(int which, uint32_t k) {
  return which == kPrefixOnly || (which == kMeanPrefix && k == 4);
}
__device__ inline float* tiled_dst(const FilterDev& f, int which, uint32_t k) {
  if (which == kTailTotals) return sums_of(f) + k;
  if (which == kWeightTotal) return sums_of(f) + 2;
  if (which == kMeanPrefix && k < 4) return f.mean + k;
  return sums_of(f) + 3;
}

__device__ inline int float_exponent(float a) { return (int)((__float_as_uint(a) >> 23) & 0xffu) - 127; }

// The map of adding |x| to k*u (r = |x|/u exactly, a power-of-two scale):
// t = r rounded to the nearest integer, ties making k + t even.  hard: |x| is
// too large for the run to stay in the binade (or not finite).
__device__ __forceinline__ ParityFn term_map(float x, double inv_u, bool& hard) {
  const double r = fabs((double)x) * inv_u;
  hard = !(r < 0x1p24);
  ParityFn f{0u, 0u, 0u, 1u};
  if (!hard) {
    const double qd = floor(r), fr = r - qd;
    const uint32_t q = (uint32_t)qd;
    if (fr == 0.5) {
      f.d0 = q + (q & 1u);
      f.d1 = q + ((q + 1u) & 1u);
      f.p0 = 0u;
      f.p1 = 0u;
    } else {
      const uint32_t t = q + (fr > 0.5 ? 1u : 0u);
      f.d0 = f.d1 = t;
      f.p0 = t & 1u;
      f.p1 = (t + 1u) & 1u;
    }
  }
  return f;
}

// The composition of the maps of lanes 0..l in lane order, in lane l (lane
// 63: the wavefront's).
__device__ __forceinline__ ParityFn wave_inclusive_fn(ParityFn f) {
  ParityFn inc = row_inclusive_fn(f);
  inc = fn_then(dpp_fn<0x142, 0xa>(inc), inc);
  inc = fn_then(dpp_fn<0x143, 0xc>(inc), inc);
  return inc;
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__device__ __forceinline__ double block_sum_d(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum_d(v);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kChainThreads / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

struct ScanLds {
  uint32_t d0[kChainThreads / 64], d1[kChainThreads / 64], pp[kChainThreads / 64];
};

// Tile t's record of sum k from its terms, one per thread of the workgroup
// (xi = 0 past n): the double sums of the terms and of their magnitudes, and
// their signs.
__device__ __forceinline__ void tile_dsum(const FilterDev& f, uint32_t k, uint32_t t, uint32_t i, float xi,
                                          double* red) {
  const uint32_t n = f.n;
  const double ds = block_sum_d((double)xi, red);
  const double as = block_sum_d(fabs((double)xi), red);
  const int pos = __syncthreads_or(xi > 0.0f ? 1 : 0);
  const int neg = __syncthreads_or(xi < 0.0f ? 1 : 0);
  const int irr = __syncthreads_or((i < n && (xi < 0.0f || !isfinite(xi))) ? 1 : 0);
  if (threadIdx.x == 0) {
    TileRec* R = tiled_recs(f, k) + t;
    R->dsum = ds;
    R->asum = as;
    R->flags = (pos ? kPos : 0u) | (neg ? kNeg : 0u) | (irr ? kIrregularTerm : 0u);
  }
}

// One workgroup per (tile, sum), one term per thread.
__global__ __launch_bounds__(kChainThreads) void tiled_dsum_kernel(Filters fs, int which) {
  __shared__ double red[kChainThreads / 64];
  const FilterDev& f = fs.one;
  const uint32_t t = blockIdx.x, k = blockIdx.y;
  const uint32_t i = t * kTiledTile + threadIdx.x;
  tile_dsum(f, k, t, i, i < f.n ? tiled_src(f, which, k)[i] : 0.0f, red);
}

__device__ inline bool normal_guess(double v, int& e) {
  const float g = fabsf((float)v);
  const bool ok = g >= 0x1p-126f && g <= 0x1p127f;  // normal (and finite)
  e = ok ? float_exponent(g) : 0;
  return ok;
}

__global__ __launch_bounds__(kChainThreads) void tiled_map_kernel(Filters fs, int which) {
  __shared__ ScanLds L;
  __shared__ double red[kChainThreads / 64];
  const FilterDev& f = fs.one;
  const uint32_t n = f.n, t = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  TileRec* recs = tiled_recs(f, k);
  // the double sum of every term before the tile, then before each sub-tile
  double before = 0.0;
  for (uint32_t j = tid; j < t; j += kChainThreads) before += recs[j].dsum;
  before = block_sum_d(before, red);
  const uint32_t i = t * kTiledTile + tid;
  const float xi = i < n ? tiled_src(f, which, k)[i] : 0.0f;
  const double ws = wave_sum_d((double)xi);
  if (lane == 0) red[wave] = ws;
  __syncthreads();
  double sub_before = before;
  for (int w = 0; w < wave; ++w) sub_before += red[w];
  __syncthreads();
  // the tile's map under the tile's guess
  int e;
  const bool guess = normal_guess(before, e);
  bool hard = !guess;
  ParityFn m{0u, 0u, 0u, 1u};
  if (guess) m = term_map(xi, ldexp(1.0, 23 - e), hard);
  // the sub-tile's maps under its guess (a) and the binade above (b)
  int es;
  const bool sguess = normal_guess(sub_before, es);
  bool hard_a = !sguess, hard_b = !sguess;
  ParityFn ma{0u, 0u, 0u, 1u}, mb{0u, 0u, 0u, 1u};
  if (sguess) {
    ma = term_map(xi, ldexp(1.0, 23 - es), hard_a);
    mb = term_map(xi, ldexp(1.0, 22 - es), hard_b);
  }
  const ParityFn ta = wave_inclusive_fn(ma), tb = wave_inclusive_fn(mb);
  const bool sub_hard_a = __ballot(hard_a) != 0, sub_hard_b = __ballot(hard_b) != 0;
  const bool sub_pos = __ballot(xi > 0.0f) != 0, sub_neg = __ballot(xi < 0.0f) != 0;
  const bool sub_nan = __ballot(xi != xi) != 0;
  if (lane == 63) {
    SubRec* S = tiled_subs(f, k) + (size_t)t * kSubTiles + wave;
    S->d0a = ta.d0;
    S->d1a = ta.d1;
    S->d0b = tb.d0;
    S->d1b = tb.d1;
    S->pp = ta.p0 | (ta.p1 << 1) | (tb.p0 << 2) | (tb.p1 << 3);
    S->e = es;
    S->flags = (sub_hard_a ? kNoMapA : 0u) | (sub_hard_b ? kNoMapB : 0u) | (sub_pos ? kPos : 0u) |
               (sub_neg ? kNeg : 0u) | (sub_nan ? kIrregularTerm : 0u);
  }
  const int any_hard = __syncthreads_or(hard ? 1 : 0);
  const ParityFn ex = block_exclusive_fn(m, L);
  if (tid == kChainThreads - 1) {
    TileRec* R = recs + t;
    const uint32_t fl = R->flags;
    // terms adding up to 2^24 u or more leave the binade whatever k is (and
    // the 32-bit increments could wrap); mixed signs have no map
    const bool nomap = any_hard || !(R->asum * ldexp(1.0, 23 - e) < 0x1p24) || ((fl & kPos) && (fl & kNeg));
    const ParityFn tot = fn_then(ex, m);
    R->d0 = tot.d0;
    R->d1 = tot.d1;
    R->pp = tot.p0 | (tot.p1 << 1);
    R->e = e;
    R->flags = fl | (nomap ? kNoMap : 0u);
  }
}

__device__ __forceinline__ uint32_t first_false(bool p) {  // the first lane where p is false (64: none)
  const uint64_t m = ~(uint64_t)__ballot(p);
  return m ? (uint32_t)__builtin_ctzll(m) : 64u;
}

__device__ __forceinline__ float from_k(uint32_t k, int e, bool neg) {
  const float v = (float)((double)k * ldexp(1.0, e - 23));
  return neg ? -v : v;
}

// Maps m, one per lane in order, applied from s = ±kin·2^(e-23) (s normal,
// in binade e): the number C of leading lanes whose maps apply (all valid,
// and k stays below 2^24: every term in the binade), with kb and ka, k
// before and after each lane's map.  A valid map adds less than 2^24, so 64
// of them cannot wrap k.
// ROW: the valid lanes are all in the first row of 16, whose composition
// needs only the row's four DPP levels.
template <bool ROW = false>
__device__ __forceinline__ uint32_t run_maps(float s, int e, ParityFn m, bool valid, uint32_t& kb, uint32_t& ka) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t L = first_false(valid);
  if (!valid) m = ParityFn{0u, 0u, 0u, 1u};
  const ParityFn inc = ROW ? row_inclusive_fn(m) : wave_inclusive_fn(m);
  const uint32_t kin = (uint32_t)(fabs((double)s) * ldexp(1.0, 23 - e));  // in [2^23, 2^24)
  ka = kin + ((kin & 1u) ? inc.d1 : inc.d0);
  kb = (uint32_t)__shfl_up((int)ka, 1);
  if (lane == 0) kb = kin;
  return first_false(lane < L && ka < (1u << 24));
}

__device__ __forceinline__ bool normal_finite(float s) { return fabsf(s) >= 0x1p-126f && isfinite(s); }

struct WalkLds {
  // per tile: the map's increments, and fe = flags | pp << 8 | (e + 256) << 10
  uint32_t d0[kWalkTiles], d1[kWalkTiles], fe[kWalkTiles];
  float s_in[kWalkTiles];
  uint32_t regular[kWalkTiles];
  float xt[kTiledTile];  // the terms of the tile being walked sub-tile by sub-tile (0 past n)
  float s;
  uint32_t stop, run_end;
};

// Tile t sub-tile by sub-tile (wavefront 0, every lane alike), from s of +0
// or of the sign of the tile's terms: runs of sub-tiles whose maps apply in
// one wavefront scan, any other sub-tile as 64 float adds in order; the
// checkpoint after every sub-tile goes to prefix (when given).
__device__ float walk_subtiles(float s, uint32_t t, uint32_t n, const float* x, const SubRec* subs, float* prefix,
                               float* xt) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t t0 = t * kTiledTile;
  // the sub-tile records and every term of the tile, loaded together
  SubRec r{};
  if (lane < kSubTiles) r = subs[(size_t)t * kSubTiles + lane];
  float xv[kSubTiles];
#pragma unroll
  for (uint32_t i = 0; i < kSubTiles; ++i) {
    const uint32_t idx = t0 + i * 64 + lane;
    xv[i] = idx < n ? x[idx] : 0.0f;
  }
#pragma unroll
  for (uint32_t i = 0; i < kSubTiles; ++i) xt[i * 64 + lane] = xv[i];  // only this wavefront reads them
  const uint32_t nq = min(kSubTiles, (n - t0 + 63) / 64);  // the sub-tiles holding terms
  for (uint32_t q = 0; q < nq;) {
    const uint32_t Q = q + lane;  // this lane's sub-tile
    const bool have = Q < nq;
    const int src = have ? (int)Q : 0;
    const uint32_t fl = (uint32_t)__shfl((int)r.flags, src);
    if (s == 0.0f) {  // +0 plus sub-tiles of zeros stays +0
      const uint32_t Z = first_false(have && !(fl & (kPos | kNeg | kIrregularTerm)));
      if (prefix && lane < Z) prefix[(t0 + Q * 64) / kGroup] = 0.0f;
      q += Z;
      if (Z > 0) continue;
    } else if (normal_finite(s)) {
      const int E = float_exponent(fabsf(s));
      const int e = __shfl(r.e, src);
      const uint32_t pp = (uint32_t)__shfl((int)r.pp, src);
      // every lane shuffles (a lane reads only lanes that run the shuffle)
      const uint32_t d0a = (uint32_t)__shfl((int)r.d0a, src), d1a = (uint32_t)__shfl((int)r.d1a, src);
      const uint32_t d0b = (uint32_t)__shfl((int)r.d0b, src), d1b = (uint32_t)__shfl((int)r.d1b, src);
      const bool use_a = e == E && !(fl & kNoMapA);
      const bool use_b = e + 1 == E && !(fl & kNoMapB);
      const ParityFn m = use_a ? ParityFn{d0a, d1a, pp & 1u, (pp >> 1) & 1u}
                               : ParityFn{d0b, d1b, (pp >> 2) & 1u, (pp >> 3) & 1u};
      const bool sign_ok = (s > 0.0f) ? !(fl & kNeg) : !(fl & kPos);
      const bool valid = have && (use_a || use_b) && sign_ok && m.d0 < (1u << 24) && m.d1 < (1u << 24);
      uint32_t kb, ka;
      const uint32_t C = run_maps<true>(s, E, m, valid, kb, ka);  // at most 16 sub-tiles: lanes 0..15
      if (C > 0) {
        const bool neg = s < 0.0f;
        if (prefix && lane < C) prefix[(t0 + Q * 64) / kGroup] = from_k(ka, E, neg);
        s = from_k((uint32_t)__builtin_amdgcn_readlane((int)ka, C - 1), E, neg);
        q += C;
        continue;
      }
    }
    // sub-tile q by the reference's float adds, in order: every lane reads
    // the same terms from LDS (a broadcast) and adds them (the zeros past n
    // leave the sum, never -0.0, unchanged)
    const uint32_t base = t0 + q * 64;
    const float4* xq = reinterpret_cast<const float4*>(xt + q * 64);
#pragma unroll
    for (int c = 0; c < 16; c += 4) {
      const float4 a = xq[c], b = xq[c + 1], d = xq[c + 2], e = xq[c + 3];
      s += a.x; s += a.y; s += a.z; s += a.w;
      s += b.x; s += b.y; s += b.z; s += b.w;
      s += d.x; s += d.y; s += d.z; s += d.w;
      s += e.x; s += e.y; s += e.z; s += e.w;
    }
    if (prefix && lane == 0) prefix[base / kGroup] = s;
    ++q;
  }
  return s;
}

// One workgroup per sum walks the tiles in order.  Wavefront 0 takes runs of
// tiles whose maps apply in one wavefront scan each (so a run of up to 64
// tiles in one binade costs one scan), a tile that cannot be taken whole
// (a binade crossing inside it, no map) sub-tile by sub-tile, and stops at a
// tile whose terms have both signs or the other sign than s, or at a sum
// that is not finite: the whole workgroup then runs exact_sum_range.
__global__ __launch_bounds__(kChainThreads) void tiled_walk_kernel(Filters fs, int which) {
  __shared__ ExactLds L;
  __shared__ WalkLds W;
  const FilterDev& f = fs.one;
  const uint32_t n = f.n, k = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
  const float* x = tiled_src(f, which, k);
  float* prefix = tiled_has_prefix(which, k) ? f.scratch + sc_c32(n) : nullptr;
  const uint32_t T = (uint32_t)tiled_tiles(n);
  TileRec* recs = tiled_recs(f, k);
  const SubRec* subs = tiled_subs(f, k);
  float s = 0.0f;
  int irr = 0;
  for (uint32_t c0 = 0; c0 < T; c0 += kWalkTiles) {  // a chunk of tile records in LDS
    const uint32_t c1 = min(T, c0 + kWalkTiles);
    uint32_t fl_any = 0u;
    for (uint32_t j = c0 + tid; j < c1; j += kChainThreads) {
      const TileRec& r = recs[j];
      W.d0[j - c0] = r.d0;
      W.d1[j - c0] = r.d1;
      W.fe[j - c0] = r.flags | (r.pp << 8) | ((uint32_t)(r.e + 256) << 10);
      W.regular[j - c0] = 0u;
      fl_any |= r.flags;
    }
    irr |= __syncthreads_or((fl_any & kIrregularTerm) ? 1 : 0);
    for (uint32_t t = c0; t < c1;) {  // uniform over the workgroup
      if (tid < 64) {
        uint32_t end = c1;  // where a stop's exact_sum_range ends
        while (t < c1) {
          const uint32_t j = t - c0 + lane;
          const bool have = t + lane < c1;
          const uint32_t fe = have ? W.fe[j] : (uint32_t)kNoMap;
          const uint32_t fl = fe & 0xffu;
          if (s == 0.0f) {  // +0 plus tiles of zeros stays +0
            const uint32_t Z = first_false(have && !(fl & (kPos | kNeg | kIrregularTerm)));
            if (lane < Z) {
              W.s_in[j] = 0.0f;
              W.regular[j] = 1u;
            }
            t += Z;
            if (Z > 0) continue;
          } else if (normal_finite(s)) {
            const int E = float_exponent(fabsf(s));
            const bool sign_ok = (s > 0.0f) ? !(fl & kNeg) : !(fl & kPos);
            const ParityFn m{have ? W.d0[j] : 0u, have ? W.d1[j] : 0u, (fe >> 8) & 1u, (fe >> 9) & 1u};
            const bool valid = have && !(fl & kNoMap) && (int)(fe >> 10) - 256 == E && sign_ok &&
                               m.d0 < (1u << 24) && m.d1 < (1u << 24);
            uint32_t kb, ka;
            const uint32_t C = run_maps(s, E, m, valid, kb, ka);
            if (C > 0) {
              const bool neg = s < 0.0f;
              if (lane < C) {
                W.s_in[j] = from_k(kb, E, neg);
                W.regular[j] = 1u;
              }
              s = from_k((uint32_t)__builtin_amdgcn_readlane((int)ka, C - 1), E, neg);
              t += C;
              continue;
            }
          } else if (!isfinite(s)) {
            break;  // the rest of the chunk by exact_sum_range
          }
          // tile t cannot be taken whole
          const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)fl);
          const bool mixed = (f0 & kPos) && (f0 & kNeg);
          const bool other_sign = (s > 0.0f && (f0 & kNeg)) || (s < 0.0f && (f0 & kPos));
          if (mixed || other_sign) {  // then while the tiles' terms have both signs
            end = t + 1;
            while (end < c1 && (W.fe[end - c0] & kPos) && (W.fe[end - c0] & kNeg)) ++end;
            break;
          }
          s = walk_subtiles(s, t, n, x, subs, prefix, W.xt);
          ++t;
        }
        if (tid == 0) {
          W.s = s;
          W.stop = t;
          W.run_end = end;
        }
      }
      __syncthreads();
      s = W.s;
      const uint32_t stop = W.stop, end = W.run_end;
      __syncthreads();
      if (stop < c1) s = exact_sum_range(s, stop * kTiledTile, min(n, end * kTiledTile), n, x, prefix, L);
      t = stop < c1 ? end : c1;
    }
    for (uint32_t j = c0 + tid; j < c1; j += kChainThreads) {
      recs[j].s_in = W.s_in[j - c0];
      recs[j].regular = W.regular[j - c0];
    }
    __syncthreads();
  }
  if (tid == 0) {
    *tiled_dst(f, which, k) = s;
    if (prefix) sums_of(f)[4] = irr ? 1.0f : 0.0f;
  }
}

// Checkpoints of resample's running sum in the tiles the walk added whole
// from their maps (the walk, or exact_sum_range, wrote the others').
__global__ __launch_bounds__(kChainThreads) void tiled_ckpt_kernel(Filters fs, int which) {
  __shared__ ScanLds L;
  const FilterDev& f = fs.one;
  const uint32_t n = f.n, t = blockIdx.x, tid = threadIdx.x;
  const uint32_t k = (which == kMeanPrefix) ? 4u : 0u;
  const TileRec* R = tiled_recs(f, k) + t;
  if (!R->regular) return;  // uniform over the workgroup
  float* prefix = f.scratch + sc_c32(n);
  const uint32_t i = t * kTiledTile + tid;
  const float xi = i < n ? tiled_src(f, which, k)[i] : 0.0f;
  const int e = R->e;
  bool hard;
  const ParityFn m = term_map(xi, ldexp(1.0, 23 - e), hard);
  const ParityFn ex = block_exclusive_fn(m, L);
  const float s_in = R->s_in;
  const uint32_t kin = (uint32_t)(fabs((double)s_in) * ldexp(1.0, 23 - e));
  const uint32_t p = kin & 1u;
  const uint32_t kb = kin + (p ? ex.d1 : ex.d0);
  const uint32_t pb = p ? ex.p1 : ex.p0;
  const uint32_t ka = kb + (pb ? m.d1 : m.d0);
  const double v = (double)ka * ldexp(1.0, e - 23);
  if (i < n && is_checkpoint(i, n)) prefix[i / kGroup] = (float)(s_in < 0.0f ? -v : v);
}

inline bool tiled_filter(const Filters& fs) { return !fs.arr && fs.one.n >= kTiledMin; }

// with_dsum false: the kernel that wrote the terms wrote the tile records too.
void launch_tiled(const Filters& fs, int which, uint32_t nsums, hipStream_t s, bool with_dsum = true) {
  const uint32_t T = (uint32_t)tiled_tiles(fs.one.n);
  if (with_dsum) tiled_dsum_kernel<<<dim3(T, nsums), kChainThreads, 0, s>>>(fs, which);
  tiled_map_kernel<<<dim3(T, nsums), kChainThreads, 0, s>>>(fs, which);
  tiled_walk_kernel<<<dim3(1, nsums), kChainThreads, 0, s>>>(fs, which);
  if (which == kPrefixOnly || which == kMeanPrefix) tiled_ckpt_kernel<<<T, kChainThreads, 0, s>>>(fs, which);
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
__device__ __forceinline__ void tail_a_elem(const GridDev& g, const FilterDev& f, uint32_t i, float& tp, float& tr) {
  const bool in = into_map(f.gdev ? *f.gdev : g, f.x[i], f.y[i], f.z[i]);
  f.inmap[i] = in ? 1 : 0;
  tp = in ? f.wp[i] : 0.f;
  tr = in ? f.wr[i] : 0.f;
  f.scratch[i] = tp;
  f.scratch[f.n + i] = tr;
}

// ParticleFilter.cpp:158-180
__device__ __forceinline__ float tail_c_elem(const FilterDev& f, uint32_t i) {
  const float* s = sums_of(f);
  const float wtp = s[0], wtr = s[1];
  const float wp = (wtp > 0) ? f.wp[i] / wtp : 0.f;
  const float wr = (wtr > 0) ? f.wr[i] / wtr : 0.f;
  f.wp[i] = wp;
  f.wr[i] = wr;
  const float w = f.inmap[i] ? (float)(wp * f.alpha + wr * (1 - f.alpha)) : 0.f;
  f.w[i] = w;
  return w;
}

// ParticleFilter.cpp:182-193: w /= wt and the weighted terms of the mean
__device__ __forceinline__ void tail_e_elem(const FilterDev& f, uint32_t i, float (&term)[5]) {
  const uint32_t n = f.n;
  const float wt = sums_of(f)[2];
  const float w = (wt > 0) ? f.w[i] / wt : 0.f;
  f.w[i] = w;
  float* sc = f.scratch;
  term[0] = w * f.x[i];
  term[1] = w * f.y[i];
  term[2] = w * f.z[i];
  term[3] = w * f.a[i];
  term[4] = w;
  for (int k = 0; k < 4; ++k) sc[k * (size_t)n + i] = term[k];
}

__global__ void tail_a_kernel(GridDev g, Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  float tp, tr;
  tail_a_elem(g, f, i, tp, tr);
}

__global__ void tail_c_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  (void)tail_c_elem(f, i);
}

__global__ void tail_e_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  float term[5];
  tail_e_elem(f, i, term);
}

// The same steps for a single filter whose sums run tiled: one workgroup per
// tile, which also writes the tile's records of the sums that follow (what
// tiled_dsum_kernel does from terms already written), saving that launch.
// The mean's step (five sums) keeps its own records kernel, which runs them
// in parallel workgroups.
__global__ __launch_bounds__(kChainThreads) void tail_a_tiled_kernel(GridDev g, Filters fs) {
  __shared__ double red[kChainThreads / 64];
  const FilterDev& f = fs.one;
  const uint32_t t = blockIdx.x, i = t * kTiledTile + threadIdx.x;
  float tp = 0.f, tr = 0.f;
  if (i < f.n) tail_a_elem(g, f, i, tp, tr);
  tile_dsum(f, 0, t, i, tp, red);
  tile_dsum(f, 1, t, i, tr, red);
}

__global__ __launch_bounds__(kChainThreads) void tail_c_tiled_kernel(Filters fs) {
  __shared__ double red[kChainThreads / 64];
  const FilterDev& f = fs.one;
  const uint32_t t = blockIdx.x, i = t * kTiledTile + threadIdx.x;
  const float w = i < f.n ? tail_c_elem(f, i) : 0.f;
  tile_dsum(f, 0, t, i, w, red);
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

__device__ inline void resample_gather(const FilterDev& f, uint32_t m, uint32_t i) {
  const uint32_t n = f.n;
  float* sc = f.scratch;
  if (f.idx_out) f.idx_out[m] = i;
  sc[4 * (size_t)n + m] = f.x[i];
  sc[5 * (size_t)n + m] = f.y[i];
  sc[6 * (size_t)n + m] = f.z[i];
  sc[7 * (size_t)n + m] = f.a[i];
  sc[8 * (size_t)n + m] = f.wp[i];
  sc[9 * (size_t)n + m] = f.wr[i];
}

// The reference's selection loop itself (ParticleFilter.cpp:203-217), one
// thread per filter, for weights whose running sum is not monotone (a
// negative or non-finite weight, flagged in sums[4] with the running sum):
// index i only moves forward, which the parallel search cannot express.
__global__ void resample_select_serial_kernel(Filters fs) {
  const FilterDev& f = fsel(fs, blockIdx.x);
  const uint32_t n = f.n;
  if (threadIdx.x != 0 || f.no_resample || sums_of(f)[4] == 0.0f) return;
  const float u01 = fs.u01 ? fs.u01[blockIdx.x] : f.u01;
  const float factor = 1.f / (float)n;
  const float r = factor * u01;
  float c = f.w[0];
  uint32_t i = 0;
  for (uint32_t m = 0; m < n; ++m) {
    const float u = r + factor * (float)m;
    while (u > c) {
      if (++i >= n) break;  // the reference reads past the end here: clamped below
      c += f.w[i];
    }
    resample_gather(f, m, min(i, n - 1));
  }
}

__global__ void resample_select_kernel(Filters fs, int mode) {
  const FilterDev& f = fsel(fs, blockIdx.y);
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = f.n;
  if (m >= n || f.no_resample) return;
  if (mode == A3D_RESAMPLE_REFERENCE && sums_of(f)[4] != 0.0f) return;  // the serial kernel selects
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
      // 16 weights loaded at a time (independent loads, one wait), then
      // added in order until the first c that u does not exceed
      bool found = false;
      for (uint32_t k0 = gi * kGroup; k0 < e && !found; k0 += 16) {
        float wv[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) wv[j] = k0 + j < e ? f.w[k0 + j] : 0.0f;
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) {
          if (!found && k0 + j < e) {
            c += wv[j];
            if (!(u > c)) {
              i = k0 + j;
              found = true;
            }
          }
        }
      }
    }
  } else {
    const double* c64 = reinterpret_cast<const double*>(sc + sc_c64(n));
    const double u = ((double)u01 + (double)m) * (1.0 / (double)n);
    i = first_not_below(c64, n, u);
  }
  resample_gather(f, m, i);
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
  exact_chain_kernel<<<fs.count, kChainThreads, 0, s>>>(fs, kWeightTotal);
  init_c_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  chain_kernel<4, -1><<<fs.count, kChainThreads, 0, s>>>(fs, kInitMean);
}

void launch_update_tail(const GridDev& g, const Filters& fs, hipStream_t s) {
  if (fs.count == 0 || fs.max_n == 0) return;
  if (tiled_filter(fs)) {
    const uint32_t T = (uint32_t)tiled_tiles(fs.one.n);
    tail_a_tiled_kernel<<<T, kChainThreads, 0, s>>>(g, fs);
    launch_tiled(fs, kTailTotals, 2, s, false);
    tail_c_tiled_kernel<<<T, kChainThreads, 0, s>>>(fs);
    launch_tiled(fs, kWeightTotal, 1, s, false);
    // (five sums: their records in parallel workgroups, not one after another)
    tail_e_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
    launch_tiled(fs, kMeanPrefix, 5, s);
    return;
  }
  tail_a_kernel<<<ew_grid(fs), kEw, 0, s>>>(g, fs);
  exact_chain_kernel<<<dim3(fs.count, 2), kChainThreads, 0, s>>>(fs, kTailTotals);
  tail_c_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  exact_chain_kernel<<<fs.count, kChainThreads, 0, s>>>(fs, kWeightTotal);
  tail_e_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
  // the mean and, beside it, resample's running sum c
  mean_prefix_kernel<<<dim3(fs.count, 5), kChainThreads, 0, s>>>(fs);
}

void launch_resample(const Filters& fs, int mode, bool have_prefix, hipStream_t s) {
  if (fs.count == 0 || fs.max_n == 0) return;
  if (mode == A3D_RESAMPLE_REFERENCE) {
    if (!have_prefix) {
      if (tiled_filter(fs))
        launch_tiled(fs, kPrefixOnly, 1, s);
      else
        exact_chain_kernel<<<fs.count, kChainThreads, 0, s>>>(fs, kPrefixOnly);
    }
  } else {
    prefix64_kernel<<<fs.count, 1024, 0, s>>>(fs);
  }
  resample_select_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs, mode);
  if (mode == A3D_RESAMPLE_REFERENCE) resample_select_serial_kernel<<<fs.count, 64, 0, s>>>(fs);
  resample_commit_kernel<<<ew_grid(fs), kEw, 0, s>>>(fs);
}

}  // namespace a3d
