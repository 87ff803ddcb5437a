// Counter-based draws for served sessions.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): a 4x32-bit counter and a 2x32-bit key, ten
// rounds of two 32x32->64 multiplications with key bumps by the Weyl
// constants.  The key is the session seed; the counter is
// (particle, step lo, step hi, 0) for the four predict normals of a particle
// and (0xffffffff, step lo, step hi, 1) for the resample uniform.  Normals
// come from the Box-Muller transform of two uniforms in (0, 1).  A session's
// draws therefore depend only on (seed, step, particle), never on the batch
// the session was served in.
#include "a3d_math.h"

namespace a3d {
namespace {

struct U4 {
  uint32_t v[4];
};

__device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.v[0]), lo0 = M0 * c.v[0];
    const uint32_t hi1 = __umulhi(M1, c.v[2]), lo1 = M1 * c.v[2];
    U4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1;
    o.v[3] = lo0;
    c = o;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// (0, 1): 24 random bits, offset by half a step
__device__ inline float u01_of(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }

__device__ inline void box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
  const float u1 = u01_of(a), u2 = u01_of(b);
  const float r = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincospif(2.0f * u2, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

__global__ void draws_kernel(Filters fs, const uint64_t* seed, const uint64_t* step, float* u01) {
  const uint32_t fi = blockIdx.y;
  const FilterDev& f = fs.arr ? fs.arr[fi] : fs.one;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t sd = seed[fi], st = step[fi];
  const uint32_t k0 = (uint32_t)sd, k1 = (uint32_t)(sd >> 32);
  if (i < f.n) {
    U4 c;
    c.v[0] = i;
    c.v[1] = (uint32_t)st;
    c.v[2] = (uint32_t)(st >> 32);
    c.v[3] = 0u;
    const U4 r = philox4x32_10(c, k0, k1);
    float* z = const_cast<float*>(f.z4) + 4 * (size_t)i;
    box_muller(r.v[0], r.v[1], z[0], z[1]);
    box_muller(r.v[2], r.v[3], z[2], z[3]);
  }
  if (i == 0 && u01) {
    U4 c;
    c.v[0] = 0xffffffffu;
    c.v[1] = (uint32_t)st;
    c.v[2] = (uint32_t)(st >> 32);
    c.v[3] = 1u;
    const U4 r = philox4x32_10(c, k0, k1);
    // uniform_real_distribution<float>(0, 1) is in [0, 1)
    u01[fi] = (float)(r.v[0] >> 8) * (1.0f / 16777216.0f);
  }
}

}  // namespace

void launch_draws(const Filters& fs, const uint64_t* d_seed, const uint64_t* d_step, float* d_u01, hipStream_t s) {
  if (fs.count == 0) return;
  const dim3 grid((fs.max_n + 255) / 256 + 1, fs.count);
  draws_kernel<<<grid, 256, 0, s>>>(fs, d_seed, d_step, d_u01);
}

}  // namespace a3d

extern "C" int a3d_session_draws(uint64_t seed, uint64_t step, int num_particles, float* z4, float* u01) {
  using namespace a3d;
  if (num_particles <= 0 || (!z4 && !u01)) return fail(A3D_ERR_ARG, "a3d_session_draws: bad argument");
  const uint32_t n = (uint32_t)num_particles;
  float* dz = nullptr;
  uint64_t* dp = nullptr;
  float* du = nullptr;
  int rc = A3D_OK;
  if (hipMalloc(&dz, 4 * (size_t)n * 4) != hipSuccess || hipMalloc(&dp, 16) != hipSuccess ||
      hipMalloc(&du, 4) != hipSuccess) {
    rc = fail(A3D_ERR_NOMEM, "a3d_session_draws: allocation");
  } else {
    const uint64_t h[2] = {seed, step};
    FilterDev f{};
    f.n = n;
    f.z4 = dz;
    if (hip_ok(hipMemcpy(dp, h, 16, hipMemcpyHostToDevice), "upload")) {
      launch_draws(single(f), dp, dp + 1, du, nullptr);
      if (!hip_ok(hipGetLastError(), "draws kernel") || !hip_ok(hipDeviceSynchronize(), "draws") ||
          (z4 && !hip_ok(hipMemcpy(z4, dz, 4 * (size_t)n * 4, hipMemcpyDeviceToHost), "download")) ||
          (u01 && !hip_ok(hipMemcpy(u01, du, 4, hipMemcpyDeviceToHost), "download")))
        rc = A3D_ERR_HIP;
    } else {
      rc = A3D_ERR_HIP;
    }
  }
  (void)hipFree(dz);
  (void)hipFree(dp);
  (void)hipFree(du);
  return rc;
}
