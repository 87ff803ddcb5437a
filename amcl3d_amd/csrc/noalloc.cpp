// ParticleFilter::update (ParticleFilter.cpp:121-196) for callers that cannot
// allocate device memory: code under HIP graph capture, serving loops and
// harnesses that forbid hipMalloc.  The caller owns the particles (the seven
// structure-of-arrays fields), the cloud, the beacons and the mean; the
// working memory the update needs (tail sums, the resample checkpoints the
// tail leaves, the in-map flags and the particle visiting order) lives in one
// file-scope device array, `scratch`, of a fixed size for at most
// kNoAllocMaxParticles particles.  Its device address is taken once per
// device with hipGetSymbolAddress (the array's name in host code is only a
// stand-in), and every call on that device shares the array: calls that
// overlap on one device race, and each call leaves its contents behind (every
// kernel of the update writes what it later reads, so nothing is cleared).
#include <hip/hip_runtime.h>

#include <mutex>

#include "a3d_handles.h"

using namespace a3d;

namespace {

constexpr uint32_t kNoAllocMaxParticles = 131072;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// [0, kSums): per-filter scratch floats; then the in-map flags (one byte per
// particle); then the visiting-order buffers
constexpr size_t kSumsBytes = align256(scratch_floats(kNoAllocMaxParticles) * sizeof(float));
constexpr size_t kInmapBytes = align256(kNoAllocMaxParticles);
constexpr size_t kOrderBytes = align256(order_bytes(kNoAllocMaxParticles));
constexpr size_t kScratchBytes = kSumsBytes + kInmapBytes + kOrderBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the
// first call there).
int scratch_on_current_device(unsigned char** out) {
  int dev = 0;
  A3D_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(A3D_ERR_HIP, std::string("a3d_pf_update_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return A3D_OK;
}

}  // namespace

extern "C" int a3d_pf_update_noalloc(const a3d_grid* g, float* d_particles, size_t n, const float* d_x,
                                     const float* d_y, const float* d_z, size_t m, const float* d_ranges, size_t nb,
                                     double alpha, double sigma, double roll, double pitch, float* d_mean,
                                     void* stream) {
  if (!g || !d_particles || !d_mean) return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: null argument");
  if (n == 0) return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: no particles");
  if (n > kNoAllocMaxParticles)
    return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: at most 131072 particles (the size of its static working memory)");
  if (m && (!d_x || !d_y || !d_z)) return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: null cloud");
  if (nb && !d_ranges) return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: null ranges");
  if (m > 0xffffffffull || nb > 0xffffffffull) return fail(A3D_ERR_ARG, "a3d_pf_update_noalloc: too large");
  unsigned char* base = nullptr;
  if (int rc = scratch_on_current_device(&base)) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t N = (uint32_t)n;
  FilterDev f{};
  f.x = d_particles;
  f.y = d_particles + n;
  f.z = d_particles + 2 * n;
  f.a = d_particles + 3 * n;
  f.w = d_particles + 4 * n;
  f.wp = d_particles + 5 * n;
  f.wr = d_particles + 6 * n;
  f.scratch = reinterpret_cast<float*>(base);
  f.zr = f.scratch + sc_sums(N) + kSumZr;
  f.inmap = base + kSumsBytes;
  f.mean = d_mean;
  f.n = N;
  f.cx = d_x;
  f.cy = d_y;
  f.cz = d_z;
  f.m = (uint32_t)m;
  f.ranges = nb ? d_ranges : nullptr;
  f.nb = (uint32_t)nb;
  f.alpha = alpha;
  f.sigma = sigma;
  f.roll = (float)roll;  // double -> float parameter of computeCloudWeight
  f.pitch = (float)pitch;
  f.rp0 = (f.roll == 0.0f && f.pitch == 0.0f) ? 1u : 0u;
  if (current_order() == A3D_ORDER_SPATIAL) {
    const OrderDev o = order_view(base + kSumsBytes + kInmapBytes, N);
    launch_order(f, o, s);
    f.order = o.order;
  }
  launch_weight(g->g, single(f), nullptr, 0, nullptr, current_ppw(), 0, s);
  launch_update_tail(g->g, single(f), s);
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}
