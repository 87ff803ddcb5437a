// Arithmetic of the reference's hot path, restated for the device.
//
// Every helper reproduces the reference's float/double mixing exactly: float
// products and sums where the reference has float operands, double where a
// double operand promotes the expression, one rounding to float on
// assignment to a float.  All HIP sources are compiled with
// -ffp-contract=off, so no multiply-add is fused, and with IEEE division and
// square root.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "a3d_internal.h"

namespace a3d {

constexpr double kPi = 3.14159265358979323846;

// Grid3d.cpp:139-149.  The reference calls unqualified sin/cos on float
// arguments; with GCC 5 these are the double functions, so each rotation
// entry is a double expression rounded once to float.
__device__ inline void rotation9(float roll, float pitch, float yaw, float r[9]) {
  const double sr = sin((double)roll), cr = cos((double)roll);
  const double sp = sin((double)pitch), cp = cos((double)pitch);
  const double sy = sin((double)yaw), cy = cos((double)yaw);
  r[0] = (float)(cy * cp);
  r[1] = (float)(cy * sp * sr - sy * cr);
  r[2] = (float)(cy * sp * cr + sy * sr);
  r[3] = (float)(sy * cp);
  r[4] = (float)(sy * sp * sr + cy * cr);
  r[5] = (float)(sy * sp * cr - cy * sr);
  r[6] = (float)(-sp);
  r[7] = (float)(cp * sr);
  r[8] = (float)(cp * cr);
}

// One axis of Grid3d.cpp:174-185: s is the rotated point (float, summed in
// the reference's order), off = (double)t - min.  Returns whether the point
// is inside along this axis and its cell index ix.
//   q  = (float)((double)s + off)                          Grid3d.cpp:174-176
//   q >= 0 && q < ext  (in double)                          Grid3d.cpp:178-179
//   ix = (uint32)floor((double)q / res), ix < size          Grid3d.cpp:181-185
// The division is replaced by a multiplication with the double reciprocal;
// whenever the product lies within w (GridDev::near_w) of an integer, where
// the two could floor differently, the correctly rounded division is
// evaluated instead, so ix always equals the reference's.
__device__ inline bool axis_cell(float s, double off, float ext_hi, double res, double inv_res, double w,
                                 uint32_t size, uint32_t& ix) {
  const float q = (float)((double)s + off);
  if (!(q >= 0.0f && q < ext_hi)) return false;
  const double qd = (double)q;
  const double d = qd * inv_res;
  double k = floor(d);
  const double fr = d - k;
  if (fr < w || fr > 1.0 - w) k = floor(qd / res);
  if (!(k < (double)size)) return false;
  ix = (uint32_t)k;
  return true;
}

// double -> int32 rounding toward zero with saturation (v_cvt_i32_f64: values
// out of range clamp to INT32_MIN / INT32_MAX, NaN gives 0): defined for
// every input, unlike a C conversion.
__device__ inline int cvt_i32_sat(double k) {
  int r;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(k));
  return r;
}

// The axis forms below take the rotated coordinate already widened to double
// (sd = (double)s, exact).
//
// The same axis without branches, for a q that is not -0.0 (the caller
// routes the lanes where q could be -0.0 to axis_exact):
//   in   = bits(q) < lim, i.e. 0 <= q < ext && floor(q/res) < size, exactly;
//   ix   = trunc(q * inv_res), which is floor for the q >= 0 that pass;
//   near = q * inv_res within w of an integer.
// Away from the window, |q*inv_res - q/res| <= size * 2^-51 < w / 8 keeps
// both on the same side of every integer, so ix is the reference's index;
// the caller recomputes the near lanes with axis_exact.
__device__ inline void axis_fast(double sd, double off, uint32_t lim, double inv_res, double w, double w2, bool& in,
                                 bool& near, uint32_t& ix) {
  const float q = (float)(sd + off);
  const double d = (double)q * inv_res;
  ix = (uint32_t)cvt_i32_sat(d);
  near = __builtin_amdgcn_fract(d + w) < w2;  // w2 = 2w
  in = __float_as_uint(q) < lim;
}

// axis_fast on a clean grid (GridDev::clean: the truncated product is the
// reference's index for every q in range), with no window test.  ZERO also
// accepts q = -0.0 (for lanes whose offset is within 2^-96 of zero), which
// lies in cell 0 whenever the axis has a cell (lim > 0).
template <bool ZERO>
__device__ inline void axis_clean(double sd, double off, uint32_t lim, double inv_res, bool& in, uint32_t& ix) {
  const float q = (float)(sd + off);
  ix = (uint32_t)cvt_i32_sat((double)q * inv_res);
  in = __float_as_uint(q) < lim;
  if (ZERO) in = in | ((q == 0.0f) & (lim != 0u));
}

__device__ inline void axis_exact(double sd, double off, float ext_hi, double res, uint32_t size, bool& ok,
                                  uint32_t& ix) {
  const float q = (float)(sd + off);
  const double k = floor((double)q / res);
  ok = (q >= 0.0f) & (q < ext_hi) & (k < (double)size);
  ix = ok ? (uint32_t)k : 0u;
}

// Grid3d::isIntoMap, Grid3d.cpp:201-208 (float vs double compares).
__host__ __device__ inline bool into_map(const GridDev& g, float x, float y, float z) {
  return x >= g.min[0] && x < g.max[0] && y >= g.min[1] && y < g.max[1] && z >= g.min[2] && z < g.max[2];
}

// ParticleFilter::computeRangeWeight, ParticleFilter.cpp:224-244.  exp is the
// double function (GCC 5 overload of exp(float)); sqrt of a float sum is
// evaluated in float, which equals (float)sqrt((double)v) exactly.
__device__ inline float range_weight(float x, float y, float z, const float* rng, uint32_t nb, double sigma) {
  if (nb == 0) return 0.f;
  float w = 1.f;
  const float k1 = (float)(1.f / (sigma * sqrt(2 * kPi)));
  const float k2 = (float)(0.5f / (sigma * sigma));
  for (uint32_t i = 0; i < nb; ++i) {
    const float rr = rng[4 * i + 0], ax = rng[4 * i + 1], ay = rng[4 * i + 2], az = rng[4 * i + 3];
    const float v = (x - ax) * (x - ax) + (y - ay) * (y - ay) + (z - az) * (z - az);
    const float r = sqrtf(v);
    const float e = -k2 * (r - rr) * (r - rr);
    w = (float)((double)w * ((double)k1 * exp((double)e)));
  }
  return w;
}

// normal_distribution<float>(0, sigma) applied to the standard normal z:
// ret * stddev + mean in float (ParticleFilter.cpp:246-250).
__host__ __device__ inline float gauss_draw(float z, double sigma) {
  const float s = (float)sigma;
  const float mean = 0.0f;
  return z * s + mean;
}

// --------------------------------------------------------- wave utilities
__device__ inline int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ inline float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ inline double readlane_d(double v, int lane) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Orders the LDS accesses of one wavefront (LDS operations of a wave execute
// in order; this keeps the compiler from moving them across the point).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class F>
__device__ inline const FilterDev& filter_at(const Filters& fs, uint32_t i) {
  return fs.arr ? fs.arr[i] : fs.one;
}

}  // namespace a3d
