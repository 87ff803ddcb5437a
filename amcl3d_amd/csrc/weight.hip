// Cloud-weight kernel: Grid3d::computeCloudWeight (Grid3d.cpp:133-199) for
// every particle of one or many filters, plus the range-beacon weight
// (ParticleFilter.cpp:224-244) of ParticleFilter::update's first loop
// (ParticleFilter.cpp:129-153).
//
// Two mappings of (particle, point) pairs to lanes give identical bits; the
// caller picks one with a3d_set_mapping (an autotuner times both).
//
// Mapping 0, particle lanes (lanes_*_kernel).  A wavefront carries P = 64/Q
// particles; lane l holds particle l div Q and, per step, point j0 + l mod Q,
// so the wavefront advances Q consecutive points of the cloud per step and
// gathers them for P nearby poses.  The Q values of one particle reach its
// first lane by quad permutes and are added in point order, so the sum is
// the reference's sequential float sum.
//
// Mapping 1, point lanes (weight_*_kernel).  One wavefront owns a block of up
// to PB particles of one filter and walks the whole cloud in chunks of 64
// points, lane j holding point c0+j.  For each particle i of the block the 64 lanes transform their points
// with particle i's pose (uniform values taken from lane i with readlane) and
// gather prob from the grid: 64 points of one particle per gather, so the
// addresses of one instruction are as close in the map as the points are in
// the cloud.  The 64 values land in row i of a per-wave LDS tile (0 for a
// point outside the map).  Lane i then adds its row in point order, so each
// particle's sum is the reference's sequential float sum over the cloud,
// bit for bit (a rejected point adds -0.0f, which leaves every sum
// unchanged).  The count of accepted points comes from a ballot.
#include "a3d_math.h"

// The structured buffer load (buffer_load_dword ... idxen), bound to the
// compiler's intrinsic by name.
__device__ float a3d_struct_load_f32(__amdgpu_buffer_rsrc_t r, int vindex, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.struct.ptr.buffer.load.f32");

namespace a3d {
namespace {

// Workgroups are dealt to the 8 XCDs round-robin by index.  Single-filter
// launches renumber them so that each XCD takes one contiguous run of the
// (spatially ordered) particle blocks, and its L2 holds the cells of one
// region of pose space.  A bijection of [0, nb).
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / 8, r = nb % 8, x = b % 8, s = b / 8;
  return (x < r) ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

// Work-queue kernels deal the (filter-major) item list to the XCDs, each
// share with its own counter: workgroup b starts on the share of its own XCD
// (b mod 8) and moves on to the other shares (in turn) when its own is empty.
// deal == 0: XCD x takes one contiguous range, so its L2 serves the particles
// of few filters at a time.  deal == k > 0: XCD x takes runs of k consecutive
// items, every 8th run, so all XCDs advance through the same part of the list
// together and the chip works on few filters at a time (whose cells then stay
// in the memory-side cache).
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_range(uint32_t n, uint32_t x) { return (uint32_t)((uint64_t)n * x / kXcds); }
__device__ inline uint32_t deal_len(uint32_t n, uint32_t x, uint32_t k) {
  if (k == 0) return xcd_range(n, x + 1) - xcd_range(n, x);
  const uint32_t rem = n % (kXcds * k), lo = x * k;
  return (n / (kXcds * k)) * k + (rem > lo ? min(k, rem - lo) : 0u);
}
// The t-th item (t < deal_len) of XCD x's share.
__device__ inline uint32_t deal_item(uint32_t n, uint32_t x, uint32_t k, uint32_t t) {
  if (k == 0) return xcd_range(n, x) + t;
  return ((t / k) * kXcds + x) * k + t % k;
}

constexpr int kTS = 68;        // tile row stride in floats (64 points + pad, 16-byte rows)
constexpr int kWaves = 4;      // wavefronts per workgroup
constexpr int kThreads = 64 * kWaves;
// Workgroups of the single-filter particle-lane kernels, whose wavefronts
// walk the cloud together.
constexpr int kBlockWaves = 4;
constexpr int kBlockThreads = 64 * kBlockWaves;

// Cell of one point under one pose: returns whether it is in the map and
// its index (0 when not, so that the gather is always in bounds).
template <bool RP0, bool IDX64>
__device__ __forceinline__ uint64_t cell_of(const GridDev& g, float px, float py, float pz, bool pv, const float* r,
                                            double ox, double oy, double oz, bool& valid) {
  float sx, sy, sz;
  if (RP0) {
    // roll = pitch = 0: r02 = r12 = +-0, r20 = r21 = +-0, r22 = 1, so the
    // reference's ((px*r00 + py*r01) + pz*r02) is (px*r00 + py*r01) and the
    // z sum is pz (finite coordinates; non-finite points are rejected).
    sx = px * r[0] + py * r[1];
    sy = px * r[3] + py * r[4];
    sz = pz;
  } else {
    sx = px * r[0] + py * r[1] + pz * r[2];
    sy = px * r[3] + py * r[4] + pz * r[5];
    sz = px * r[6] + py * r[7] + pz * r[8];
  }
  uint32_t ix = 0, iy = 0, iz = 0;
  bool ok = pv;
  ok = axis_cell(sx, ox, g.ext_hi[0], g.res, g.inv_res, g.near_w[0], g.size[0], ix) && ok;
  ok = axis_cell(sy, oy, g.ext_hi[1], g.res, g.inv_res, g.near_w[1], g.size[1], iy) && ok;
  ok = axis_cell(sz, oz, g.ext_hi[2], g.res, g.inv_res, g.near_w[2], g.size[2], iz) && ok;
  valid = ok;
  if (!ok) return 0;
  if (IDX64) return (uint64_t)ix + (uint64_t)iy * g.size[0] + (uint64_t)iz * g.step_z;
  return (uint64_t)(ix + iy * g.size[0] + iz * (uint32_t)g.step_z);
}

// Evaluates the 64 points held by the lanes for every particle of the block
// whose bit is set in `active`, writing tile rows and accumulating counts.
template <bool RP0, bool IDX64>
__device__ __forceinline__ void eval_chunk(const GridDev& g, float px, float py, float pz, bool pv,
                                           const float (&r)[9], double ox, double oy, double oz, uint32_t count,
                                           uint64_t active, float* tile, int& cnt) {
  // Four particles per iteration: their cells are computed first and the
  // four gathers issued together, so each wavefront keeps four loads in
  // flight.  Rows of particles outside the map are not written (never read).
  constexpr int U = 4;
  const int lane = lane_id();
  for (uint32_t i0 = 0; i0 < count; i0 += U) {
    uint64_t idx[U];
    bool val[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u;  // <= 63: count <= 64 and i0 is a multiple of U
      const bool act = i < count && ((active >> i) & 1ull);
      float ri[9];
      if (RP0) {
        ri[0] = readlane_f(r[0], i);
        ri[1] = readlane_f(r[1], i);
        ri[3] = readlane_f(r[3], i);
        ri[4] = readlane_f(r[4], i);
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) ri[k] = readlane_f(r[k], i);
      }
      const double oxi = readlane_d(ox, i), oyi = readlane_d(oy, i), ozi = readlane_d(oz, i);
      idx[u] = cell_of<RP0, IDX64>(g, px, py, pz, pv && act, ri, oxi, oyi, ozi, val[u]);
    }
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = g.prob[idx[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u;
      if (i < count && ((active >> i) & 1ull)) {
        tile[i * kTS + lane] = val[u] ? v[u] : -0.0f;  // x + -0.0f == x for every x
        const uint64_t vm = __ballot(val[u]);
        cnt += (lane == (int)i) ? (int)__popcll(vm) : 0;
      }
    }
  }
}

// KFILTER: particles of a filter (skip those outside the map, write wp and
// wr); otherwise explicit poses (weight every pose, write out[]).
template <int PB, bool IDX64, bool KFILTER>
__device__ void weight_block(const GridDev& g, const FilterDev& f, uint32_t p0, uint32_t count, float* tile,
                             float* out) {
  const int lane = lane_id();
  const bool has = (uint32_t)lane < count;
  const uint32_t pk = p0 + (has ? (uint32_t)lane : 0u);
  const uint32_t pi = f.order ? f.order[pk] : pk;
  float tx = 0.f, ty = 0.f, tz = 0.f, ta = 0.f;
  if (has) {
    tx = f.x[pi];
    ty = f.y[pi];
    tz = f.z[pi];
    ta = f.a[pi];
  }
  const bool act = KFILTER ? (has && into_map(g, tx, ty, tz)) : has;
  const uint64_t active = __ballot(act);
  if (active == 0) return;
  float r[9];
  rotation9(f.roll, f.pitch, ta, r);
  const double ox = (double)tx - g.min[0];
  const double oy = (double)ty - g.min[1];
  const double oz = (double)tz - g.min[2];
  float wr = 0.f;
  if (KFILTER && act) wr = range_weight(tx, ty, tz, f.ranges, f.nb, f.sigma);

  float wsum = 0.f;
  int cnt = 0;
  const uint32_t m = f.m;
  const float error_z = 0.f;
  for (uint32_t c0 = 0; c0 < m; c0 += 64) {
    const uint32_t j = c0 + (uint32_t)lane;
    bool pv = j < m;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (pv) {
      px = f.cx[j];
      py = f.cy[j];
      pz = f.cz[j] + error_z;
    }
    if (f.rp0) {
      pv = pv && isfinite(px) && isfinite(py) && isfinite(pz);
      eval_chunk<true, IDX64>(g, px, py, pz, pv, r, ox, oy, oz, count, active, tile, cnt);
    } else {
      eval_chunk<false, IDX64>(g, px, py, pz, pv, r, ox, oy, oz, count, active, tile, cnt);
    }
    wave_lds_sync();
    if (has) {
      const float4* row = reinterpret_cast<const float4*>(tile + lane * kTS);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const float4 v = row[k];
        wsum += v.x;
        wsum += v.y;
        wsum += v.z;
        wsum += v.w;
      }
    }
    wave_lds_sync();
  }
  if (act) {
    const float wp = (cnt <= 10) ? 0.f : wsum / (float)cnt;
    if (KFILTER) {
      f.wp[pi] = wp;
      f.wr[pi] = wr;
    } else {
      out[pi] = wp;
    }
  }
}

// ---------------------------------------------------------------------------
// Mapping 0, lanes = particles.  A wavefront carries P = 64/Q particles; lane
// l holds particle l / Q and, at each step, point (step*Q + l % Q) of the
// cloud, so the Q lanes of a particle are consecutive lanes of one quad.  A
// particle's sum takes the Q values of a step from its quad in point order
// (DPP quad permutes folded into the adds), so it is the reference's
// sequential float sum.  Particles that resampling made near-duplicates sit
// in neighbouring lanes: one gather then touches few cache lines.

// Fourth descriptor word of a buffer resource on the gfx9 family: 32-bit data
// format, no swizzle, no per-lane stride add.
constexpr int kBufWord3 = 0x00020000;

// How the cloud weight loads prob[i] (the grid's index kind, GridDev):
//   kIdxBuf  32-bit cell indices through a structured buffer of 4-byte records
//            (cells + 1 of them): no address arithmetic per lane.  The
//            hardware forms index * 4 in 32 bits, so only grids whose bytes,
//            the zero cell included, fit 2^32 take it (GridDev::buf32).
//   kIdx32   32-bit cell indices, 64-bit addresses (prob + index * 4): grids
//            of 2^30 to 2^32 - 2 cells.
//   kIdx64   64-bit cell indices (GridDev::idx64).
// Index `cells` is the zero cell.  Code for a wider kind reads every grid a
// narrower kind can (a batch over several maps runs the widest).  Kernels of
// the two wider kinds let a wavefront whose cells all lie in one 2^30-cell
// segment k (wave_segment) read them as
//   kIdxSeg  32-bit cell indices (modulo 2^32) through a structured buffer
//            based at prob + k * 2^30: the hardware's 32-bit byte offset
//            (index * 4 modulo 2^32) is the cell's offset inside the segment.
//            A point outside the map reads record 0xffffffff, past the
//            buffer's 0xfffffffe records, which loads 0 without a memory
//            access (adding 0 leaves every in-order sum unchanged: the sum
//            starts at +0 and a round-to-nearest sum is -0 only when both
//            addends are).
enum IndexKind { kIdxBuf = 0, kIdx32 = 1, kIdx64 = 2, kIdxSeg = 3 };

template <int IK>
struct ProbSrc;

template <>
struct ProbSrc<kIdxBuf> {
  using Idx = uint32_t;
  __amdgpu_buffer_rsrc_t r;
  Idx zero;
  __device__ explicit ProbSrc(const GridDev& g, uint32_t = 0)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)g.prob, (short)4, (int)(uint32_t)(g.cells + 1), kBufWord3)),
        zero((uint32_t)g.cells) {}
  __device__ float operator()(Idx i) const { return a3d_struct_load_f32(r, (int)i, 0, 0, 0); }
};

// prob as a global-memory pointer: a grid read through a descriptor in
// memory (FilterDev::gdev) would otherwise be loaded with flat instructions,
// whose outstanding loads also hold back every wait for the LDS reads of the
// cloud tile.
using GlobalFloat = const __attribute__((address_space(1))) float;

template <>
struct ProbSrc<kIdx32> {
  using Idx = uint32_t;
  GlobalFloat* p;
  Idx zero;
  __device__ explicit ProbSrc(const GridDev& g, uint32_t = 0) : p((GlobalFloat*)g.prob), zero((uint32_t)g.cells) {}
  __device__ float operator()(Idx i) const { return p[(uint64_t)i]; }
};

template <>
struct ProbSrc<kIdx64> {
  using Idx = uint64_t;
  GlobalFloat* p;
  Idx zero;
  __device__ explicit ProbSrc(const GridDev& g, uint32_t = 0) : p((GlobalFloat*)g.prob), zero(g.cells) {}
  __device__ float operator()(Idx i) const { return p[i]; }
};

// k: the wavefront's segment (uniform).
template <>
struct ProbSrc<kIdxSeg> {
  using Idx = uint32_t;
  static constexpr uint32_t kRecords = 0xfffffffeu;
  __amdgpu_buffer_rsrc_t r;
  Idx zero;
  __device__ explicit ProbSrc(const GridDev& g, uint32_t k)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)(g.prob + ((uint64_t)k << 30)), (short)4, (int)kRecords,
                                            kBufWord3)),
        zero(0xffffffffu) {}
  __device__ float operator()(Idx i) const { return a3d_struct_load_f32(r, (int)i, 0, 0, 0); }
};

// The pose of a lane's particle.  A lane without an active particle has
// NaN offsets, so every point fails its axis tests (q is NaN).  slow marks
// offsets within 2^-96 of zero, the only ones for which q can round to -0.0
// (the fast axis test needs a q that is not -0.0); such lanes take the exact
// path at every point.
struct LanePose {
  float r[9];
  double ox, oy, oz;
  bool slow;
};

// A cloud point as staged for the particle lanes: x and y as floats and z as
// a float in .z, or, with roll = pitch = 0 (where z enters the arithmetic
// only as (double)pz), as that double in (.z, .w).
template <bool RP0>
__device__ __forceinline__ float4 stage_point(float px, float py, float pz) {
  if (RP0) {
    const double d = (double)pz;
    return make_float4(px, py, __int_as_float(__double2loint(d)), __int_as_float(__double2hiint(d)));
  }
  return make_float4(px, py, pz, 0.f);
}

// The rotated point of Grid3d.cpp:163-172, widened to double.  With roll =
// pitch = 0 the terms of the reference's sums that are products with r02 =
// r12 = r20 = r21 = +-0 are dropped, which is exact for a finite point; a
// non-finite coordinate makes sz (pz itself) or one of sx, sy (px, py times
// cos(yaw), which is never 0 for a float yaw) non-finite, so the point fails
// an axis test, as in the reference.
template <bool RP0>
__device__ __forceinline__ void rotate(const float (&r)[9], const float4& p, double& dx, double& dy, double& dz) {
  if (RP0) {
    dx = (double)(p.x * r[0] + p.y * r[1]);
    dy = (double)(p.x * r[3] + p.y * r[4]);
    dz = __hiloint2double(__float_as_int(p.w), __float_as_int(p.z));
  } else {
    dx = (double)(p.x * r[0] + p.y * r[1] + p.z * r[2]);
    dy = (double)(p.x * r[3] + p.y * r[4] + p.z * r[5]);
    dz = (double)(p.x * r[6] + p.y * r[7] + p.z * r[8]);
  }
}

template <int IK>
__device__ __forceinline__ typename ProbSrc<IK>::Idx linear(const GridDev& g, uint32_t ix, uint32_t iy,
                                                               uint32_t iz) {
  if constexpr (IK == kIdx64) {
    return (uint64_t)ix + (uint64_t)iy * g.size[0] + (uint64_t)iz * g.step_z;
  } else {
    // (iz * sy + iy) * sx + ix in two 24-bit multiply-adds: the host takes
    // the 64-bit path unless sx < 2^24 and sy * sz < 2^24 (exact for in-map
    // cells; other lanes are replaced by the zero cell)
    uint32_t a, b;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(a) : "v"(iz), "s"(g.size[1]), "v"(iy));
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(b) : "v"(a), "s"(g.size[0]), "v"(ix));
    return b;
  }
}

// Lane-mask forms of a select and a count: the mask is the one a compare
// writes (inactive lanes 0), so each costs one vector instruction (the mask
// is the instruction's one scalar operand).
__device__ __forceinline__ uint32_t select_by_mask(uint64_t m, uint32_t if_set, uint32_t if_clear) {
  uint32_t r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(if_clear), "v"(if_set), "s"(m));
  return r;
}

__device__ __forceinline__ int add_mask_bit(int c, uint64_t m) {  // c + (this lane's bit of m)
  int r;
  uint64_t carry;
  asm("v_addc_co_u32_e64 %0, %1, 0, %2, %3" : "=v"(r), "=s"(carry) : "v"(c), "s"(m));
  return r;
}

// How the fast cell index is made exact:
//   kWindow     window test on every axis, flagged lanes (and slow lanes)
//               redone with the division;
//   kClean      a clean grid (the truncated product is always the
//               reference's index): no test;
//   kCleanZero  a clean grid and a wavefront with slow lanes: q = -0.0, the
//               one value the bit-pattern bound rejects wrongly, is accepted
//               by an explicit compare.
enum CellMode { kWindow = 0, kClean = 1, kCleanZero = 2 };

// Cell of one staged point under the lane's pose, fast form: vm, the lanes
// whose point is in the map; near (kWindow: some axis needs the exact
// division, or the lane is slow); and the index to load (the zero cell
// outside the map).
template <bool RP0, int IK, int MODE>
__device__ __forceinline__ typename ProbSrc<IK>::Idx cell_fast(const GridDev& g, const ProbSrc<IK>& src,
                                                                  const LanePose& lp, const float4& p, uint64_t& vm,
                                                                  bool& near) {
  double dx, dy, dz;
  rotate<RP0>(lp.r, p, dx, dy, dz);
  bool inx, iny, inz;
  uint32_t ix, iy, iz;
  if constexpr (MODE == kWindow) {
    bool nx, ny, nz;
    axis_fast(dx, lp.ox, g.lim[0], g.inv_res, g.near_w[0], g.near_w2[0], inx, nx, ix);
    axis_fast(dy, lp.oy, g.lim[1], g.inv_res, g.near_w[1], g.near_w2[1], iny, ny, iy);
    axis_fast(dz, lp.oz, g.lim[2], g.inv_res, g.near_w[2], g.near_w2[2], inz, nz, iz);
    near = nx | ny | nz | lp.slow;
  } else {
    constexpr bool kZero = MODE == kCleanZero;
    axis_clean<kZero>(dx, lp.ox, g.lim[0], g.inv_res, inx, ix);
    axis_clean<kZero>(dy, lp.oy, g.lim[1], g.inv_res, iny, iy);
    axis_clean<kZero>(dz, lp.oz, g.lim[2], g.inv_res, inz, iz);
    near = false;
  }
  const auto lin = linear<IK>(g, ix, iy, iz);
  if constexpr (IK == kIdx64) {
    const bool valid = inx & iny & inz;
    vm = __builtin_amdgcn_ballot_w64(valid);
    return valid ? lin : src.zero;
  } else {
    vm = __builtin_amdgcn_ballot_w64(inx) & __builtin_amdgcn_ballot_w64(iny) & __builtin_amdgcn_ballot_w64(inz);
    return select_by_mask(vm, lin, src.zero);
  }
}

// The same with the correctly rounded division on every axis.
template <bool RP0, int IK>
__device__ __forceinline__ typename ProbSrc<IK>::Idx cell_exact(const GridDev& g, const ProbSrc<IK>& src,
                                                                   const LanePose& lp, const float4& p, bool& valid) {
  double dx, dy, dz;
  rotate<RP0>(lp.r, p, dx, dy, dz);
  bool okx, oky, okz;
  uint32_t ix, iy, iz;
  axis_exact(dx, lp.ox, g.ext_hi[0], g.res, g.size[0], okx, ix);
  axis_exact(dy, lp.oy, g.ext_hi[1], g.res, g.size[1], oky, iy);
  axis_exact(dz, lp.oz, g.ext_hi[2], g.res, g.size[2], okz, iz);
  valid = okx & oky & okz;
  const auto lin = linear<IK>(g, ix, iy, iz);
  return valid ? lin : src.zero;
}

// Value of member K of this lane's group of Q consecutive lanes (a quad
// permute: Q = 4 broadcasts lane K of the quad, Q = 2 lane K of the pair).
template <int Q, int K>
__device__ __forceinline__ float group_member(float v) {
  constexpr int ctrl = (Q == 4) ? K * 0x55 : (K | (K << 2) | ((K + 2) << 4) | ((K + 2) << 6));
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, false));
}

// Points per cloud tile staged in LDS (the tile shape's second dimension,
// chosen at launch: 32, 64, 128 or 256): the workgroup loads a tile with one
// coalesced load per thread and array, then every step reads its point with
// one LDS read.  Per-wavefront staging (work items of different filters in
// one workgroup) uses tiles of 64.
constexpr int kMaxChunk = 256;
constexpr int kWaveChunk = 64;

// Steps whose gathers a wavefront issues together (a group).
template <int Q>
constexpr int kGroup = (Q == 1) ? 8 : 4;

// The steps of one staged tile of points: S = points/Q steps, in groups of U
// (S is a multiple of U for every tile size and Q).  The gathers run one
// group ahead of the sum: a group's values are added once the next group's
// cells are computed, so a wavefront computes those cells while its loads
// are in flight.  pend holds the previous group's values (-0.0f before the
// first group, which leaves the sum unchanged); the caller adds what is
// left pending after the last group, before any later point.
template <int Q, bool RP0, int IK, int MODE, typename Acc>
__device__ __forceinline__ void lanes_chunk(const GridDev& g, const ProbSrc<IK>& src, const LanePose& lp,
                                            const float4* cl, int po, int S, int& cnt,
                                            float (&pend)[kGroup<Q>], Acc&& accumulate) {
  using Idx = typename ProbSrc<IK>::Idx;
  constexpr int U = kGroup<Q>;
#pragma unroll 1
  for (int s0 = 0; s0 < S; s0 += U) {
    Idx idx[U];
    bool nr[U];
    float4 pt[U];  // all LDS reads of the group first
#pragma unroll
    for (int u = 0; u < U; ++u) pt[u] = cl[(s0 + u) * Q + po];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t vm;
      idx[u] = cell_fast<RP0, IK, MODE>(g, src, lp, pt[u], vm, nr[u]);
      cnt = add_mask_bit(cnt, vm);
    }
    bool any = false;
#pragma unroll
    for (int u = 0; u < U; ++u) any |= nr[u];
    if (MODE == kWindow && __builtin_expect(__builtin_amdgcn_ballot_w64(any) != 0, 0)) {
      // rare: redo the flagged lanes with the division, re-reading their
      // points (nothing of the fast form is kept for this path)
      const volatile float* vcl = reinterpret_cast<const volatile float*>(cl);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (nr[u]) {
          const int k = 4 * ((s0 + u) * Q + po);
          const float4 p = make_float4(vcl[k], vcl[k + 1], vcl[k + 2], vcl[k + 3]);
          bool ve;
          const Idx ie = cell_exact<RP0, IK>(g, src, lp, p, ve);
          cnt += (ve ? 1 : 0) - (idx[u] != src.zero ? 1 : 0);
          idx[u] = ie;
        }
      }
    }
    // the previous group's values (loaded while these cells were computed),
    // then this group's loads into the registers they free; the scheduling
    // barrier keeps the compiler from hoisting the adds (and their wait on
    // the loads) above the cell arithmetic
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) accumulate(pend[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) pend[u] = src(idx[u]);
  }
}

template <bool RP0>
__device__ inline float4 cloud_point(const FilterDev& f, uint32_t j) {
  return stage_point<RP0>(f.cx[j], f.cy[j], f.cz[j]);
}

// One lane's particle over the whole cloud.  BLOCK: the workgroup's
// wavefronts (all on the same filter) walk the cloud together: tiles of
// `chunk` points (a multiple of 32, at most kMaxChunk and at most the
// workgroup's thread count) are staged once per workgroup in LDS (double
// buffered, one barrier per tile), so the wavefronts' nearby poses gather the
// same points at the same time and share cache lines; every wavefront of the
// workgroup must call this (run = false skips the arithmetic of a wavefront
// with no active particle).  Otherwise each wavefront stages its own tiles of
// kWaveChunk points (work-queue kernels whose wavefronts hold different
// filters).
template <int Q, bool RP0, int IK, int MODE, bool BLOCK>
__device__ __forceinline__ void lanes_loop(const GridDev& g, const FilterDev& f, const LanePose& lp, float& wsum,
                                           int& cnt, float4* cl, bool run, uint32_t chunk, uint32_t segk) {
  using Idx = typename ProbSrc<IK>::Idx;
  const ProbSrc<IK> src(g, segk);
  const int lane = lane_id();
  const int po = lane % Q;
  const uint32_t m = f.m;
  // Each lane counts its own accepted points (an integer count is order-free;
  // the caller adds the Q lanes of a particle at the end).  The float sum is
  // not: the Q values of a step reach the particle's first lane and are added
  // in point order.  A rejected point reads the zero cell (or, through a
  // segment's buffer, an out-of-range record): +0, which leaves the sum
  // unchanged, since the sum starts at +0 and a round-to-nearest sum is -0
  // only when both addends are.
  auto accumulate = [&](float v) {
    if constexpr (Q == 1) {
      wsum += v;
    } else if constexpr (Q == 2) {
      wsum += group_member<2, 0>(v);
      wsum += group_member<2, 1>(v);
    } else {
      wsum += group_member<4, 0>(v);
      wsum += group_member<4, 1>(v);
      wsum += group_member<4, 2>(v);
      wsum += group_member<4, 3>(v);
    }
  };
  if (!BLOCK) chunk = kWaveChunk;
  const uint32_t mfull = m - m % chunk;
  const int steps = (int)(chunk / Q);
  float pend[kGroup<Q>];
#pragma unroll
  for (int u = 0; u < kGroup<Q>; ++u) pend[u] = -0.0f;
  if constexpr (BLOCK) {
    const uint32_t tid = threadIdx.x;
    if (mfull > 0) {
      if (tid < chunk) cl[tid] = cloud_point<RP0>(f, tid);
      __syncthreads();
    }
    uint32_t buf = 0;
    for (uint32_t c0 = 0; c0 < mfull; c0 += chunk, buf ^= 1u) {
      const float4* cur = cl + buf * chunk;
      float4* nxt = cl + (buf ^ 1u) * chunk;
      const bool more = c0 + chunk < mfull;
      float4 next = make_float4(0.f, 0.f, 0.f, 0.f);
      if (more && tid < chunk) next = cloud_point<RP0>(f, c0 + chunk + tid);
      if (run) lanes_chunk<Q, RP0, IK, MODE>(g, src, lp, cur, po, steps, cnt, pend, accumulate);
      // every wavefront is done with the buffer written next (it read it in
      // the previous tile, before the last barrier)
      if (more && tid < chunk) nxt[tid] = next;
      __syncthreads();
    }
  } else {
    for (uint32_t c0 = 0; c0 < mfull; c0 += chunk) {
      wave_lds_sync();  // the previous tile's reads precede these writes
      cl[lane] = cloud_point<RP0>(f, c0 + lane);
      wave_lds_sync();
      lanes_chunk<Q, RP0, IK, MODE>(g, src, lp, cl, po, steps, cnt, pend, accumulate);
    }
  }
  if (!run) return;
#pragma unroll
  for (int u = 0; u < kGroup<Q>; ++u) accumulate(pend[u]);  // the last group of the last tile
  for (uint32_t c0 = mfull; c0 < m; c0 += Q) {  // the last m mod chunk points
    const uint32_t j = c0 + po;
    const bool pin = j < m;
    const uint32_t js = pin ? j : 0u;
    // a lane past the cloud's end takes a NaN point, which every test rejects
    const float4 p = stage_point<RP0>(pin ? f.cx[js] : __builtin_nanf(""), f.cy[js], f.cz[js]);
    uint64_t vm;
    bool nr;
    Idx idx = cell_fast<RP0, IK, MODE>(g, src, lp, p, vm, nr);
    cnt = add_mask_bit(cnt, vm);
    if (MODE == kWindow && __builtin_expect(__builtin_amdgcn_ballot_w64(nr) != 0, 0)) {
      if (nr) {
        bool ve;
        const Idx ie = cell_exact<RP0, IK>(g, src, lp, p, ve);
        cnt += (ve ? 1 : 0) - (idx != src.zero ? 1 : 0);
        idx = ie;
      }
    }
    accumulate(src(idx));
  }
}

// The loop variant of a wavefront: uniform per wavefront; in BLOCK form
// wavefronts of one workgroup may take different ones, each passing the same
// barriers.
template <int Q, int IK, bool CLEAN, bool BLOCK>
__device__ __forceinline__ void lanes_variant_loop(const GridDev& g, const FilterDev& f, const LanePose& lp,
                                                   float& wsum, int& cnt, float4* cl, bool run, uint32_t chunk,
                                                   uint32_t segk) {
  if constexpr (CLEAN) {
    const bool zero = __builtin_amdgcn_ballot_w64(lp.slow) != 0;
    if (f.rp0) {
      if (zero)
        lanes_loop<Q, true, IK, kCleanZero, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
      else
        lanes_loop<Q, true, IK, kClean, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
    } else {
      if (zero)
        lanes_loop<Q, false, IK, kCleanZero, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
      else
        lanes_loop<Q, false, IK, kClean, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
    }
  } else {
    if (f.rp0)
      lanes_loop<Q, true, IK, kWindow, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
    else
      lanes_loop<Q, false, IK, kWindow, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
  }
}

// The range [lo, hi] of the finite rotated z of each of kZrParts parts of
// each filter's cloud, into f.zr[2p], f.zr[2p + 1] (filters without zr are
// skipped): a point's rotated z, sz = px*r20 + py*r21 + pz*r22 (pz itself
// with roll = pitch = 0, as the weight kernels stage it), is the same float
// under every particle of the filter, since the third row of the rotation
// depends on roll and pitch only.  A non-finite sz fails the z test under
// every pose and is left out (an empty part gives [inf, -inf]).  Grid
// (kZrParts, filters from base).
constexpr int kZrThreads = 256;
__global__ __launch_bounds__(kZrThreads) void cloud_zrange_kernel(Filters fs, uint32_t base) {
  const FilterDev& f = fs.arr ? fs.arr[base + blockIdx.y] : fs.one;
  if (!f.zr) return;  // uniform
  float r[9];
  rotation9(f.roll, f.pitch, 0.0f, r);
  const float inf = __builtin_inff();
  float lo = inf, hi = -inf;
  const uint32_t m = f.m, part = (m + kZrParts - 1) / kZrParts;
  const uint32_t j0 = blockIdx.x * part, j1 = min(m, j0 + part);
  const bool rp0 = f.rp0 != 0;
  constexpr int U = 8;  // loads in flight per thread
  for (uint32_t jb = j0 + threadIdx.x; jb < j1; jb += U * kZrThreads) {
    float z[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = jb + (uint32_t)u * kZrThreads;
      z[u] = inf;
      if (j < j1) z[u] = rp0 ? f.cz[j] : f.cx[j] * r[6] + f.cy[j] * r[7] + f.cz[j] * r[8];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (fabsf(z[u]) < inf) {
        lo = fminf(lo, z[u]);
        hi = fmaxf(hi, z[u]);
      }
    }
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    lo = fminf(lo, __shfl_xor(lo, o));
    hi = fmaxf(hi, __shfl_xor(hi, o));
  }
  __shared__ float red[2 * (kZrThreads / 64)];
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    red[2 * w] = lo;
    red[2 * w + 1] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 0; v < kZrThreads / 64; ++v) {
      lo = fminf(lo, red[2 * v]);
      hi = fmaxf(hi, red[2 * v + 1]);
    }
    f.zr[2 * blockIdx.x] = lo;
    f.zr[2 * blockIdx.x + 1] = hi;
  }
}

// Whether every cell this wavefront can gather over f's cloud lies in one
// 2^30-cell segment (then *k is its number), for a grid with seg_bits and a
// filter whose cloud z range [lo, hi] is in f.zr by parts (cloud_zrange_kernel).
//   q = (float)((double)sz + oz) does not decrease in sz nor in oz, nor does
//   the cell index floor(q / res) in q, so every in-map cell of the
//   wavefront has iz within one of [iz(lo + oz_min), iz(hi + oz_max)] (one
//   for the reciprocal product, which floors at most one apart from the
//   division), clamped to the map: its index lies in [izl * sx * sy,
//   (izh + 1) * sx * sy - 1].  An empty range (no finite point, no active
//   lane, or every q outside the map) means no in-map cell: any segment
//   serves.  The last check keeps the top index's low 32 bits below the
//   buffer's record count.
__device__ __forceinline__ bool wave_segment(const GridDev& g, const FilterDev& f, bool act, double oz,
                                             uint32_t* k) {
  float lo = f.zr[0], hi = f.zr[1];
#pragma unroll
  for (uint32_t p = 1; p < kZrParts; ++p) {
    lo = fminf(lo, f.zr[2 * p]);
    hi = fmaxf(hi, f.zr[2 * p + 1]);
  }
  const double inf = __builtin_inf();
  double ol = act ? oz : inf, oh = act ? oz : -inf;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    ol = fmin(ol, __shfl_xor(ol, o));
    oh = fmax(oh, __shfl_xor(oh, o));
  }
  const double qlo = (double)(float)((double)lo + ol), qhi = (double)(float)((double)hi + oh);
  double zl = floor(qlo * g.inv_res) - 1.0, zh = floor(qhi * g.inv_res) + 1.0;
  const double top = (double)g.size[2] - 1.0;
  if (!(zl >= 0.0)) zl = 0.0;
  if (!(zh <= top)) zh = top;
  uint32_t seg = 1u, kk = 0u;
  if (zl <= zh) {
    const uint64_t a = (uint64_t)zl * g.step_z, b = ((uint64_t)zh + 1) * g.step_z - 1;
    kk = (uint32_t)(a >> 30);
    seg = ((a >> g.seg_bits) == (b >> g.seg_bits) && (uint32_t)b < ProbSrc<kIdxSeg>::kRecords) ? 1u : 0u;
  }
  *k = (uint32_t)__builtin_amdgcn_readfirstlane((int)kk);
  return __builtin_amdgcn_readfirstlane((int)seg) != 0;
}

// CLEAN: the launch's grids are all clean (GridDev::clean).  BLOCK: the
// workgroup walks the cloud together (every wavefront calls this, count may
// be 0); cl then holds two chunks per workgroup, else one per wavefront.
template <int Q, int IK, bool KFILTER, bool CLEAN, bool BLOCK>
__device__ void weight_lanes(const GridDev& g, const FilterDev& f, uint32_t p0, uint32_t count, float* out,
                             float4* cl, uint32_t chunk) {
  const int lane = lane_id();
  const int ps = lane / Q, po = lane % Q;
  const bool has = (uint32_t)ps < count;
  const uint32_t pk = has ? p0 + (uint32_t)ps : 0u;  // f.n >= 1 whenever this runs
  const uint32_t pi = f.order ? f.order[pk] : pk;
  float tx = 0.f, ty = 0.f, tz = 0.f, ta = 0.f;
  if (has) {
    tx = f.x[pi];
    ty = f.y[pi];
    tz = f.z[pi];
    ta = f.a[pi];
  }
  const bool act = KFILTER ? (has && into_map(g, tx, ty, tz)) : has;
  const bool run = __ballot(act) != 0;
  if (!BLOCK && !run) return;
  LanePose lp;
  rotation9(f.roll, f.pitch, ta, lp.r);
  const double nan = __builtin_nan("");
  lp.ox = act ? (double)tx - g.min[0] : nan;
  lp.oy = act ? (double)ty - g.min[1] : nan;
  lp.oz = act ? (double)tz - g.min[2] : nan;
  constexpr double kTiny = 0x1p-96;
  lp.slow = fabs(lp.ox) < kTiny || fabs(lp.oy) < kTiny || fabs(lp.oz) < kTiny;
  float wr = 0.f;
  if (KFILTER && act && po == 0) wr = range_weight(tx, ty, tz, f.ranges, f.nb, f.sigma);
  float wsum = 0.f;
  int cnt = 0;
  if constexpr (IK != kIdxBuf) {
    uint32_t segk = 0;
    if (g.seg_bits && f.zr && wave_segment(g, f, act, lp.oz, &segk))
      lanes_variant_loop<Q, kIdxSeg, CLEAN, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, segk);
    else
      lanes_variant_loop<Q, IK, CLEAN, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, 0u);
  } else {
    lanes_variant_loop<Q, IK, CLEAN, BLOCK>(g, f, lp, wsum, cnt, cl, run, chunk, 0u);
  }
  if (Q > 1) {
    int c = cnt;
#pragma unroll
    for (int q = 1; q < Q; ++q) c += __shfl(cnt, ps * Q + q);
    cnt = c;
  }
  if (act && po == 0) {
    const float wp = (cnt <= 10) ? 0.f : wsum / (float)cnt;
    if (KFILTER) {
      f.wp[pi] = wp;
      f.wr[pi] = wr;
    } else {
      out[pi] = wp;
    }
  }
}

// One filter: the workgroup's wavefronts hold consecutive particles (of the
// visiting order) and walk the cloud together.
template <int Q, int IK, bool CLEAN>
__global__ __launch_bounds__(kBlockThreads) void lanes_single_kernel(GridDev g, FilterDev f, uint32_t chunk) {
  constexpr uint32_t P = 64 / Q;
  __shared__ float4 cloud[2 * kMaxChunk];
  const uint32_t wave = xcd_block(blockIdx.x, gridDim.x) * kBlockWaves + (threadIdx.x >> 6);
  const uint32_t p0 = wave * P;
  const uint32_t count = p0 < f.n ? min(P, f.n - p0) : 0u;
  weight_lanes<Q, IK, true, CLEAN, true>(g, f, p0, count, nullptr, cloud, chunk);
}

template <int Q, int IK, bool CLEAN>
__global__ __launch_bounds__(kBlockThreads) void lanes_poses_kernel(GridDev g, FilterDev f, float* out,
                                                                    uint32_t chunk) {
  constexpr uint32_t P = 64 / Q;
  __shared__ float4 cloud[2 * kMaxChunk];
  const uint32_t wave = blockIdx.x * kBlockWaves + (threadIdx.x >> 6);
  const uint32_t p0 = wave * P;
  const uint32_t count = p0 < f.n ? min(P, f.n - p0) : 0u;
  weight_lanes<Q, IK, false, CLEAN, true>(g, f, p0, count, out, cloud, chunk);
}

// One work item of up to kBlockWaves * P particles of one filter, taken by
// a whole workgroup: wavefront w holds particles p0 + w*P .. and the
// wavefronts walk the filter's cloud together.  A filter that names its own
// grid (a batch over several maps) is weighted on that grid, else on the
// launch's; the launch grid's seg_bits = 0 keeps every filter off segment
// loads.
template <int Q, int IK, bool CLEAN>
__device__ __forceinline__ void lanes_item(const GridDev& g, const FilterDev& f, const WorkItem& it, float4* cl,
                                           uint32_t chunk) {
  constexpr uint32_t P = 64 / Q;
  const uint32_t w = threadIdx.x >> 6, p0 = it.p0 + w * P;
  const uint32_t count = (w * P < it.count) ? min(P, it.count - w * P) : 0u;
  if (f.gdev) {
    GridDev gf = *f.gdev;
    if (!g.seg_bits) gf.seg_bits = 0u;  // the launch holds off segment loads
    weight_lanes<Q, IK, true, CLEAN, true>(gf, f, p0, count, nullptr, cl, chunk);
  } else {
    weight_lanes<Q, IK, true, CLEAN, true>(g, f, p0, count, nullptr, cl, chunk);
  }
}

// Clean grids: at most 128 registers, four wavefronts per SIMD (the loads
// of spread particle sets need the wavefronts to hide their latency).
template <int Q, int IK, bool CLEAN>
__global__ __launch_bounds__(kBlockThreads, CLEAN ? 4 : 1) void lanes_fleet_kernel(GridDev g, const FilterDev* fs,
                                                                    const WorkItem* items, uint32_t nitems,
                                                                    uint32_t* counter, int persistent,
                                                                    uint32_t chunk, uint32_t deal) {
  __shared__ float4 cloud[2 * kMaxChunk];
  __shared__ uint32_t taken;
  if (!persistent) {
    if (blockIdx.x >= nitems) return;
    const WorkItem it = items[blockIdx.x];
    lanes_item<Q, IK, CLEAN>(g, fs[it.filter], it, cloud, chunk);
    return;
  }
  for (uint32_t r = 0; r < kXcds; ++r) {
    const uint32_t x = (blockIdx.x + r) % kXcds, len = deal_len(nitems, x, deal);
    for (;;) {
      if (threadIdx.x == 0) taken = atomicAdd(counter + x, 1u);
      __syncthreads();
      const uint32_t t = taken;
      __syncthreads();  // every wavefront has read it before the next take
      if (t >= len) break;  // uniform over the workgroup
      const WorkItem it = items[deal_item(nitems, x, deal, t)];
      lanes_item<Q, IK, CLEAN>(g, fs[it.filter], it, cloud, chunk);
    }
  }
}

template <int PB, bool IDX64>
__global__ __launch_bounds__(kThreads) void weight_single_kernel(GridDev g, FilterDev f) {
  __shared__ float4 tiles4[kWaves * PB * kTS / 4];
  float* tile = reinterpret_cast<float*>(tiles4) + (threadIdx.x >> 6) * PB * kTS;
  const uint32_t wave = xcd_block(blockIdx.x, gridDim.x) * kWaves + (threadIdx.x >> 6);
  const uint32_t p0 = wave * PB;
  if (p0 >= f.n) return;
  const uint32_t count = min((uint32_t)PB, f.n - p0);
  weight_block<PB, IDX64, true>(g, f, p0, count, tile, nullptr);
}

template <int PB, bool IDX64>
__global__ __launch_bounds__(kThreads) void weight_poses_kernel(GridDev g, FilterDev f, float* out) {
  __shared__ float4 tiles4[kWaves * PB * kTS / 4];
  float* tile = reinterpret_cast<float*>(tiles4) + (threadIdx.x >> 6) * PB * kTS;
  const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const uint32_t p0 = wave * PB;
  if (p0 >= f.n) return;
  const uint32_t count = min((uint32_t)PB, f.n - p0);
  weight_block<PB, IDX64, false>(g, f, p0, count, tile, out);
}

template <int PB, bool IDX64>
__device__ __forceinline__ void block_item(const GridDev& g, const FilterDev& f, const WorkItem& it, float* tile) {
  const uint32_t count = min(it.count, (uint32_t)PB);
  if (f.gdev) {
    const GridDev gf = *f.gdev;
    weight_block<PB, IDX64, true>(gf, f, it.p0, count, tile, nullptr);
  } else {
    weight_block<PB, IDX64, true>(g, f, it.p0, count, tile, nullptr);
  }
}

// Many filters: work items (filter, first particle, count) are taken either
// one per wavefront (grid sized to the list) or, when persistent, from an
// atomic counter by resident wavefronts until the list is exhausted.  Items
// are listed in decreasing cost, so the persistent form is a
// longest-processing-time-first schedule.
template <int PB, bool IDX64>
__global__ __launch_bounds__(kThreads) void weight_fleet_kernel(GridDev g, const FilterDev* fs,
                                                                const WorkItem* items, uint32_t nitems,
                                                                uint32_t* counter, int persistent, uint32_t deal) {
  __shared__ float4 tiles4[kWaves * PB * kTS / 4];
  float* tile = reinterpret_cast<float*>(tiles4) + (threadIdx.x >> 6) * PB * kTS;
  const int lane = lane_id();
  uint32_t k = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (!persistent) {
    if (k >= nitems) return;
    block_item<PB, IDX64>(g, fs[items[k].filter], items[k], tile);
    return;
  }
  for (uint32_t r = 0; r < kXcds; ++r) {
    const uint32_t x = (blockIdx.x + r) % kXcds, len = deal_len(nitems, x, deal);
    for (;;) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(counter + x, 1u);
      t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
      if (t >= len) break;
      const WorkItem& it = items[deal_item(nitems, x, deal, t)];
      block_item<PB, IDX64>(g, fs[it.filter], it, tile);
    }
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <int PB>
void launch_single_pb(const GridDev& g, const FilterDev& f, hipStream_t s) {
  const uint32_t waves = (f.n + PB - 1) / PB;
  const uint32_t blocks = (waves + kWaves - 1) / kWaves;
  if (g.idx64)
    weight_single_kernel<PB, true><<<blocks, kThreads, 0, s>>>(g, f);
  else
    weight_single_kernel<PB, false><<<blocks, kThreads, 0, s>>>(g, f);
}

template <int PB>
void launch_poses_pb(const GridDev& g, const FilterDev& f, float* out, hipStream_t s) {
  const uint32_t waves = (f.n + PB - 1) / PB;
  const uint32_t blocks = (waves + kWaves - 1) / kWaves;
  if (g.idx64)
    weight_poses_kernel<PB, true><<<blocks, kThreads, 0, s>>>(g, f, out);
  else
    weight_poses_kernel<PB, false><<<blocks, kThreads, 0, s>>>(g, f, out);
}

template <int PB>
void launch_fleet_pb(const GridDev& g, const FilterDev* fs, const WorkItem* items, uint32_t nitems,
                     uint32_t* counter, int persistent, hipStream_t s) {
  uint32_t blocks;
  if (persistent) {
    int dev = 0, cus = 256, per_cu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (g.idx64)
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, weight_fleet_kernel<PB, true>, kThreads, 0);
    else
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, weight_fleet_kernel<PB, false>, kThreads, 0);
    if (per_cu < 1) per_cu = 1;
    blocks = (uint32_t)(cus * per_cu);
    const uint32_t need = (nitems + kWaves - 1) / kWaves;
    if (blocks > need) blocks = need;
    if (blocks == 0) blocks = 1;
    zero_u32_kernel<<<1, 64, 0, s>>>(counter, kXcds);
  } else {
    blocks = (nitems + kWaves - 1) / kWaves;
  }
  if (blocks == 0) return;
  const uint32_t deal = (uint32_t)current_deal();
  if (g.idx64)
    weight_fleet_kernel<PB, true><<<blocks, kThreads, 0, s>>>(g, fs, items, nitems, counter, persistent, deal);
  else
    weight_fleet_kernel<PB, false><<<blocks, kThreads, 0, s>>>(g, fs, items, nitems, counter, persistent, deal);
}

// The index kind a launch reads its grids with: the widest its maps need.
inline int index_kind(const GridDev& g) { return g.idx64 ? kIdx64 : (g.buf32 ? kIdxBuf : kIdx32); }

// Launches pick the kernel of the grid's index kind and, for a clean grid,
// the one without the window test: variant 2 * kind + clean.
#define A3D_LANES_VARIANT(v, KERNEL, Q, LAUNCH, ...)                                   \
  switch (v) {                                                                          \
    case 0: KERNEL<Q, kIdxBuf, false><<<LAUNCH>>>(__VA_ARGS__); break;                  \
    case 1: KERNEL<Q, kIdxBuf, true><<<LAUNCH>>>(__VA_ARGS__); break;                   \
    case 2: KERNEL<Q, kIdx32, false><<<LAUNCH>>>(__VA_ARGS__); break;                   \
    case 3: KERNEL<Q, kIdx32, true><<<LAUNCH>>>(__VA_ARGS__); break;                    \
    case 4: KERNEL<Q, kIdx64, false><<<LAUNCH>>>(__VA_ARGS__); break;                   \
    default: KERNEL<Q, kIdx64, true><<<LAUNCH>>>(__VA_ARGS__); break;                   \
  }

inline int lanes_variant(const GridDev& g) { return 2 * index_kind(g) + (g.clean ? 1 : 0); }

template <int Q>
void lanes_single_q(const GridDev& g, const FilterDev& f, uint32_t chunk, hipStream_t s) {
  constexpr uint32_t P = 64 / Q;
  const uint32_t waves = (f.n + P - 1) / P;
  const uint32_t blocks = (waves + kBlockWaves - 1) / kBlockWaves;
#define A3D_LAUNCH blocks, kBlockThreads, 0, s
  A3D_LANES_VARIANT(lanes_variant(g), lanes_single_kernel, Q, A3D_LAUNCH, g, f, chunk)
#undef A3D_LAUNCH
}

template <int Q>
void lanes_poses_q(const GridDev& g, const FilterDev& f, float* out, uint32_t chunk, hipStream_t s) {
  constexpr uint32_t P = 64 / Q;
  const uint32_t waves = (f.n + P - 1) / P;
  const uint32_t blocks = (waves + kBlockWaves - 1) / kBlockWaves;
#define A3D_LAUNCH blocks, kBlockThreads, 0, s
  A3D_LANES_VARIANT(lanes_variant(g), lanes_poses_kernel, Q, A3D_LAUNCH, g, f, out, chunk)
#undef A3D_LAUNCH
}

template <int Q>
int fleet_blocks_per_cu(int v) {
  int per_cu = 1;
  switch (v) {
    case 0: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanes_fleet_kernel<Q, kIdxBuf, false>,
                                                               kBlockThreads, 0); break;
    case 1: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanes_fleet_kernel<Q, kIdxBuf, true>,
                                                               kBlockThreads, 0); break;
    case 2: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanes_fleet_kernel<Q, kIdx32, false>,
                                                               kBlockThreads, 0); break;
    case 3: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanes_fleet_kernel<Q, kIdx32, true>,
                                                               kBlockThreads, 0); break;
    case 4: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanes_fleet_kernel<Q, kIdx64, false>,
                                                               kBlockThreads, 0); break;
    default: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanes_fleet_kernel<Q, kIdx64, true>,
                                                                kBlockThreads, 0); break;
  }
  return per_cu;
}

// g.clean must hold for every grid the filters name (the caller clears it
// otherwise), and g's index kind must be at least as wide as each grid's.
template <int Q>
void lanes_fleet_q(const GridDev& g, const FilterDev* fs, const WorkItem* items, uint32_t nitems, uint32_t* counter,
                   int persistent, uint32_t chunk, hipStream_t s) {
  const int v = lanes_variant(g);
  uint32_t blocks;
  if (persistent) {
    static int cus = 0, per_cu[6] = {0, 0, 0, 0, 0, 0};
    if (!cus) {
      int dev = 0;
      (void)hipGetDevice(&dev);
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      for (int k = 0; k < 6; ++k) per_cu[k] = fleet_blocks_per_cu<Q>(k);
      if (cus < 1) cus = 256;
    }
    blocks = (uint32_t)(cus * std::max(1, per_cu[v]));
    if (blocks > nitems) blocks = nitems;  // one item per workgroup
    if (blocks == 0) blocks = 1;
    zero_u32_kernel<<<1, 64, 0, s>>>(counter, kXcds);
  } else {
    blocks = nitems;
  }
  if (blocks == 0) return;
  const dim3 b(blocks), t(kBlockThreads);
  const uint32_t deal = (uint32_t)current_deal();
#define A3D_LAUNCH b, t, 0, s
  A3D_LANES_VARIANT(v, lanes_fleet_kernel, Q, A3D_LAUNCH, g, fs, items, nitems, counter, persistent, chunk, deal)
#undef A3D_LAUNCH
}

#undef A3D_LANES_VARIANT

}  // namespace

int item_particles(int mapping, int ppw) {
  ppw = clamp_ppw(ppw);
  return mapping == 0 ? ppw * kBlockWaves : ppw;
}

int clamp_ppw(int ppw) {
  if (ppw <= 16) return 16;
  if (ppw <= 32) return 32;
  return 64;
}

int clamp_chunk(int points) {
  if (points <= 32) return 32;
  if (points <= 64) return 64;
  if (points <= 128) return 128;
  return kMaxChunk;
}

void launch_weight(const GridDev& g, const Filters& fs, const WorkItem* d_items, uint32_t nitems,
                   uint32_t* d_counter, int ppw, int persistent, hipStream_t s) {
  ppw = clamp_ppw(ppw);
  if (current_mapping() == 0) {  // lanes = particles, ppw particles per wavefront
    const uint32_t chunk = (uint32_t)clamp_chunk(current_chunk());
    // the clouds' z ranges for the kernels' segment loads (a launch of a
    // wider index kind that allows them)
    if (index_kind(g) != kIdxBuf && g.seg_bits && (fs.arr || fs.one.n > 0))
      for (uint32_t base = 0; base < fs.count; base += 65535u)
        cloud_zrange_kernel<<<dim3(kZrParts, std::min(fs.count - base, 65535u)), kZrThreads, 0, s>>>(fs, base);
    if (!fs.arr) {
      if (fs.one.n == 0) return;
      if (ppw == 64) lanes_single_q<1>(g, fs.one, chunk, s);
      else if (ppw == 32) lanes_single_q<2>(g, fs.one, chunk, s);
      else lanes_single_q<4>(g, fs.one, chunk, s);
      return;
    }
    if (nitems == 0) return;
    if (ppw == 64) lanes_fleet_q<1>(g, fs.arr, d_items, nitems, d_counter, persistent, chunk, s);
    else if (ppw == 32) lanes_fleet_q<2>(g, fs.arr, d_items, nitems, d_counter, persistent, chunk, s);
    else lanes_fleet_q<4>(g, fs.arr, d_items, nitems, d_counter, persistent, chunk, s);
    return;
  }
  if (!fs.arr) {
    if (fs.one.n == 0) return;
    if (ppw == 16) launch_single_pb<16>(g, fs.one, s);
    else if (ppw == 32) launch_single_pb<32>(g, fs.one, s);
    else launch_single_pb<64>(g, fs.one, s);
    return;
  }
  if (nitems == 0) return;
  if (ppw == 16) launch_fleet_pb<16>(g, fs.arr, d_items, nitems, d_counter, persistent, s);
  else if (ppw == 32) launch_fleet_pb<32>(g, fs.arr, d_items, nitems, d_counter, persistent, s);
  else launch_fleet_pb<64>(g, fs.arr, d_items, nitems, d_counter, persistent, s);
}

void launch_weight_poses(const GridDev& g, const float* cx, const float* cy, const float* cz, uint32_t m,
                         const float* tx, const float* ty, const float* tz, const float* yaw, uint32_t n, float roll,
                         float pitch, float* out, float* zr, int ppw, hipStream_t s) {
  if (n == 0) return;
  FilterDev f{};
  f.zr = zr;
  f.x = const_cast<float*>(tx);
  f.y = const_cast<float*>(ty);
  f.z = const_cast<float*>(tz);
  f.a = const_cast<float*>(yaw);
  f.n = n;
  f.cx = cx;
  f.cy = cy;
  f.cz = cz;
  f.m = m;
  f.roll = roll;
  f.pitch = pitch;
  f.rp0 = (roll == 0.0f && pitch == 0.0f) ? 1u : 0u;
  ppw = clamp_ppw(ppw);
  if (current_mapping() == 0) {
    const uint32_t chunk = (uint32_t)clamp_chunk(current_chunk());
    if (index_kind(g) != kIdxBuf && g.seg_bits) cloud_zrange_kernel<<<dim3(kZrParts, 1), kZrThreads, 0, s>>>(single(f), 0u);
    if (ppw == 64) lanes_poses_q<1>(g, f, out, chunk, s);
    else if (ppw == 32) lanes_poses_q<2>(g, f, out, chunk, s);
    else lanes_poses_q<4>(g, f, out, chunk, s);
    return;
  }
  if (ppw == 16) launch_poses_pb<16>(g, f, out, s);
  else if (ppw == 32) launch_poses_pb<32>(g, f, out, s);
  else launch_poses_pb<64>(g, f, out, s);
}

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  zero_u32_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

}  // namespace a3d
