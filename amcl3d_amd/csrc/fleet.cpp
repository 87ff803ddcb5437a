// Fleet: many independent particle filters on one resident grid.
//
// All filters' particles live in one structure-of-arrays allocation (field k
// of filter f at state + k*total + offset[f]).  An update stages every
// filter's cloud and ranges, then runs ONE cloud-weight launch over a work
// list of (filter, block of particles) items in decreasing cost, taken by
// persistent wavefronts from an atomic counter, and one tail launch with a
// workgroup per filter.  Each filter's arithmetic is exactly that of the
// single-filter path, so results do not depend on the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "a3d_handles.h"

using namespace a3d;

struct a3d_fleet {
  hipStream_t stream = nullptr;
  uint32_t F = 0;
  std::vector<uint32_t> n;
  std::vector<uint64_t> poff;  // F + 1
  std::vector<uint64_t> soff;  // scratch offsets (floats)
  uint64_t total = 0;
  uint32_t max_n = 0;
  float* state = nullptr;
  float* scratch = nullptr;
  uint8_t* inmap = nullptr;
  float* mean = nullptr;
  FilterDev* d_filters = nullptr;
  std::vector<FilterDev> h_filters;
  // staging
  float* stage = nullptr;
  size_t stage_cap = 0;
  // staged update
  float* cloud = nullptr;
  size_t cloud_cap = 0;
  WorkItem* d_items = nullptr;
  size_t items_cap = 0;
  uint32_t nitems = 0;
  int items_ppw = 32;      // kernel shape the staged list was cut for
  int items_mapping = 0;
  bool items_spatial = false;  // the staged filters visit particles in spatial order
  // particle visiting orders: one view per filter in one allocation
  uint8_t* ord = nullptr;
  std::vector<uint64_t> ooff;  // F + 1, bytes
  OrderDev* d_ord = nullptr;
  std::vector<OrderDev> h_ord;
  uint32_t* d_counter = nullptr;
  bool staged = false;
  bool prefix_valid = false;  // the last update left resample's running sum c
  // device-parameter (_dev) forms: the layout fixed by a3d_fleet_set_layout,
  // per filter {cloud offset, beacon offset, particle offset}
  bool layout = false;
  uint64_t* d_lay = nullptr;
  bool capturing = false;
  std::vector<uint32_t> st_m, st_nb;
  std::vector<double> st_alpha, st_sigma, st_roll, st_pitch;
  std::vector<uint64_t> st_moff, st_boff;
  TimingPool timing;

  FilterDev base(uint32_t f) const {
    FilterDev d{};
    const uint64_t o = poff[f];
    d.x = state + 0 * total + o;
    d.y = state + 1 * total + o;
    d.z = state + 2 * total + o;
    d.a = state + 3 * total + o;
    d.w = state + 4 * total + o;
    d.wp = state + 5 * total + o;
    d.wr = state + 6 * total + o;
    d.scratch = scratch + soff[f];
    d.zr = scratch ? d.scratch + sc_sums(n[f]) + kSumZr : nullptr;
    d.inmap = inmap + o;
    d.mean = mean + 4 * (size_t)f;
    d.n = n[f];
    return d;
  }
  Filters device_filters() const {
    Filters s{};
    s.arr = d_filters;
    s.count = F;
    s.max_n = max_n;
    return s;
  }
};

static int grow(float*& buf, size_t& cap, size_t floats, hipStream_t s) {
  if (floats <= cap) return A3D_OK;
  if (buf) {
    A3D_HIP(hipStreamSynchronize(s));
    (void)hipFree(buf);
  }
  buf = nullptr;
  cap = 0;
  const size_t want = std::max(floats, (size_t)1 << 16);
  if (hipMalloc(&buf, want * sizeof(float)) != hipSuccess) return fail(A3D_ERR_NOMEM, "fleet: staging allocation");
  cap = want;
  return A3D_OK;
}

static int upload_filters(a3d_fleet* fl) {
  A3D_HIP(hipMemcpyAsync(fl->d_filters, fl->h_filters.data(), fl->F * sizeof(FilterDev), hipMemcpyHostToDevice,
                         fl->stream));
  return A3D_OK;
}

// Work list of an update: blocks of particles of the kernel shape current
// now, filters in decreasing cost n*m (a longest-first schedule).
static int build_items(a3d_fleet* fl, const uint32_t* m) {
  const int ppw = clamp_ppw(current_ppw()), mapping = current_mapping();
  const uint32_t per_item = (uint32_t)item_particles(mapping, ppw);
  std::vector<uint32_t> order(fl->F);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)fl->n[a] * m[a] > (uint64_t)fl->n[b] * m[b];
  });
  std::vector<WorkItem> items;
  items.reserve(fl->total / per_item + fl->F);
  for (uint32_t f : order)
    for (uint32_t p0 = 0; p0 < fl->n[f]; p0 += per_item)
      items.push_back({f, p0, std::min<uint32_t>(per_item, fl->n[f] - p0), 0});
  if (items.size() > fl->items_cap) {
    A3D_HIP(hipStreamSynchronize(fl->stream));
    (void)hipFree(fl->d_items);
    fl->d_items = nullptr;
    fl->items_cap = 0;
    if (hipMalloc(&fl->d_items, items.size() * sizeof(WorkItem)) != hipSuccess)
      return fail(A3D_ERR_NOMEM, "fleet: work list allocation");
    fl->items_cap = items.size();
  }
  A3D_HIP(hipMemcpyAsync(fl->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                         fl->stream));
  A3D_HIP(hipStreamSynchronize(fl->stream));  // host vectors go out of scope
  fl->nitems = (uint32_t)items.size();
  fl->items_ppw = ppw;
  fl->items_mapping = mapping;
  fl->st_m.assign(m, m + fl->F);
  return A3D_OK;
}

// The _dev forms take each step's parameters from device arrays, written into
// the filters' descriptors by these kernels (one thread per filter), so a
// step enqueues kernels only and a HIP graph can capture it.
// ParticleFilter.cpp:101-104: sigma_k = fabs(delta_k * mod_k) in double.
__global__ void fleet_patch_predict_kernel(FilterDev* fs, uint32_t F, const double* mods, const double* deltas,
                                           const float* z4, const uint64_t* lay) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FilterDev& d = fs[f];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double dl = deltas[4 * (size_t)f + k];
    d.delta[k] = dl;
    d.dev[k] = fabs(dl * mods[4 * (size_t)f + k]);
  }
  d.z4 = z4 + 4 * lay[3 * (size_t)f + 2];
}

// params: alpha, sigma, roll, pitch per filter (roll and pitch are narrowed
// to float, as computeCloudWeight's parameters, Grid3d.h:103-104).
__global__ void fleet_patch_update_kernel(FilterDev* fs, uint32_t F, const float* x, const float* y, const float* z,
                                          const float* ranges, const double* params, const uint64_t* lay) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  FilterDev& d = fs[f];
  const uint64_t mo = lay[3 * (size_t)f], bo = lay[3 * (size_t)f + 1];
  d.cx = x + mo;
  d.cy = y + mo;
  d.cz = z + mo;
  d.ranges = d.nb ? ranges + 4 * bo : nullptr;
  d.alpha = params[4 * (size_t)f];
  d.sigma = params[4 * (size_t)f + 1];
  d.roll = (float)params[4 * (size_t)f + 2];
  d.pitch = (float)params[4 * (size_t)f + 3];
  d.rp0 = (d.roll == 0.0f && d.pitch == 0.0f) ? 1u : 0u;
}

static int refuse_capture(const a3d_fleet* fl, const char* what) {
  return fl->capturing ? fail(A3D_ERR_STATE, std::string(what) + ": not allowed during capture (use the _dev forms)")
                       : A3D_OK;
}

extern "C" {

int a3d_fleet_create(size_t nfilters, const uint32_t* num_particles, a3d_fleet** out) {
  if (!out || !num_particles || nfilters == 0 || nfilters > 0xffffffull)
    return fail(A3D_ERR_ARG, "a3d_fleet_create: bad argument");
  *out = nullptr;
  a3d_fleet* fl = new a3d_fleet();
  fl->F = (uint32_t)nfilters;
  fl->n.assign(num_particles, num_particles + nfilters);
  fl->poff.assign(nfilters + 1, 0);
  fl->soff.assign(nfilters + 1, 0);
  for (size_t f = 0; f < nfilters; ++f) {
    if (fl->n[f] == 0) {
      delete fl;
      return fail(A3D_ERR_ARG, "a3d_fleet_create: every filter needs at least one particle");
    }
    fl->poff[f + 1] = fl->poff[f] + fl->n[f];
    fl->soff[f + 1] = fl->soff[f] + ((scratch_floats(fl->n[f]) + 3) & ~(size_t)3);  // keep doubles aligned
    fl->max_n = std::max(fl->max_n, fl->n[f]);
  }
  fl->total = fl->poff[nfilters];
  if (fl->total > 0xffffffffull) {
    delete fl;
    return fail(A3D_ERR_ARG, "a3d_fleet_create: more than 2^32 particles");
  }
  bool ok = hipStreamCreateWithFlags(&fl->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&fl->state, 7 * fl->total * 4) == hipSuccess &&
            hipMalloc(&fl->scratch, fl->soff[nfilters] * 4) == hipSuccess &&
            hipMalloc(&fl->inmap, fl->total) == hipSuccess && hipMalloc(&fl->mean, 16 * nfilters) == hipSuccess &&
            hipMalloc(&fl->d_filters, sizeof(FilterDev) * nfilters) == hipSuccess &&
            hipMalloc(&fl->d_counter, 64) == hipSuccess;
  fl->ooff.assign(nfilters + 1, 0);
  for (size_t f = 0; f < nfilters; ++f) fl->ooff[f + 1] = fl->ooff[f] + order_stride(fl->n[f]);
  if (ok)
    ok = hipMalloc(&fl->ord, fl->ooff[nfilters]) == hipSuccess &&
         hipMalloc(&fl->d_ord, sizeof(OrderDev) * nfilters) == hipSuccess;
  if (ok) {
    fl->h_ord.resize(nfilters);
    for (size_t f = 0; f < nfilters; ++f) fl->h_ord[f] = order_view(fl->ord + fl->ooff[f], fl->n[f]);
    ok = hipMemcpyAsync(fl->d_ord, fl->h_ord.data(), sizeof(OrderDev) * nfilters, hipMemcpyHostToDevice,
                        fl->stream) == hipSuccess;
  }
  if (ok)
    ok = hipMemsetAsync(fl->state, 0, 7 * fl->total * 4, fl->stream) == hipSuccess &&
         hipMemsetAsync(fl->mean, 0, 16 * nfilters, fl->stream) == hipSuccess &&
         hipStreamSynchronize(fl->stream) == hipSuccess;
  if (!ok) {
    a3d_fleet_destroy(fl);
    return fail(A3D_ERR_NOMEM, "a3d_fleet_create: device allocation failed");
  }
  fl->h_filters.resize(nfilters);
  *out = fl;
  return A3D_OK;
}

int a3d_fleet_destroy(a3d_fleet* fl) {
  if (!fl) return A3D_OK;
  if (fl->capturing) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(fl->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
  }
  if (fl->stream) (void)hipStreamSynchronize(fl->stream);
  fl->timing.destroy();
  (void)hipFree(fl->state);
  (void)hipFree(fl->scratch);
  (void)hipFree(fl->inmap);
  (void)hipFree(fl->mean);
  (void)hipFree(fl->d_filters);
  (void)hipFree(fl->d_counter);
  (void)hipFree(fl->ord);
  (void)hipFree(fl->d_ord);
  (void)hipFree(fl->stage);
  (void)hipFree(fl->cloud);
  (void)hipFree(fl->d_items);
  (void)hipFree(fl->d_lay);
  if (fl->stream) (void)hipStreamDestroy(fl->stream);
  delete fl;
  return A3D_OK;
}

int a3d_fleet_set_particles(a3d_fleet* fl, const float* p7) {
  if (!fl || !p7) return fail(A3D_ERR_ARG, "a3d_fleet_set_particles: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_set_particles")) return rc;
  std::vector<float> soa(7 * fl->total);
  for (uint64_t i = 0; i < fl->total; ++i)
    for (int k = 0; k < 7; ++k) soa[k * fl->total + i] = p7[7 * i + k];
  A3D_HIP(hipMemcpyAsync(fl->state, soa.data(), soa.size() * 4, hipMemcpyHostToDevice, fl->stream));
  A3D_HIP(hipStreamSynchronize(fl->stream));
  fl->prefix_valid = false;
  return A3D_OK;
}

int a3d_fleet_get_particles(const a3d_fleet* fl, float* p7) {
  if (!fl || !p7) return fail(A3D_ERR_ARG, "a3d_fleet_get_particles: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_get_particles")) return rc;
  std::vector<float> soa(7 * fl->total);
  A3D_HIP(hipMemcpyAsync(soa.data(), fl->state, soa.size() * 4, hipMemcpyDeviceToHost, fl->stream));
  A3D_HIP(hipStreamSynchronize(fl->stream));
  for (uint64_t i = 0; i < fl->total; ++i)
    for (int k = 0; k < 7; ++k) p7[7 * i + k] = soa[k * fl->total + i];
  return A3D_OK;
}

int a3d_fleet_init(a3d_fleet* fl, const float* poses, const float* devs, const float* z4) {
  if (!fl || !poses || !devs || !z4) return fail(A3D_ERR_ARG, "a3d_fleet_init: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_init")) return rc;
  fl->layout = false;  // the descriptors below replace the layout's
  fl->staged = false;  // and the staged update's
  int rc = grow(fl->stage, fl->stage_cap, 4 * fl->total, fl->stream);
  if (rc) return rc;
  A3D_HIP(hipMemcpyAsync(fl->stage, z4, 4 * fl->total * 4, hipMemcpyHostToDevice, fl->stream));
  for (uint32_t f = 0; f < fl->F; ++f) {
    FilterDev d = fl->base(f);
    const float* dv = devs + 4 * (size_t)f;
    const float dev = std::max(std::max(dv[0], dv[1]), dv[2]);  // ParticleFilter.cpp:54-56
    d.dev[0] = (float)(1. / (dev * std::sqrt(2 * M_PI)));
    d.dev[1] = (float)(1. / (2 * dev * dev));
    for (int k = 0; k < 4; ++k) {
      d.delta[k] = dv[k];
      d.pose[k] = poses[4 * (size_t)f + k];
    }
    d.z4 = fl->stage + 4 * fl->poff[f];
    fl->h_filters[f] = d;
  }
  rc = upload_filters(fl);
  if (rc) return rc;
  launch_init(fl->device_filters(), fl->stream);
  fl->prefix_valid = false;
  A3D_HIP(hipGetLastError());
  A3D_HIP(hipStreamSynchronize(fl->stream));
  return A3D_OK;
}

int a3d_fleet_predict(a3d_fleet* fl, const double* mods, const double* deltas, const float* z4) {
  if (!fl || !mods || !deltas || !z4) return fail(A3D_ERR_ARG, "a3d_fleet_predict: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_predict")) return rc;
  fl->layout = false;  // the descriptors below replace the layout's
  fl->staged = false;  // and the staged update's
  int rc = grow(fl->stage, fl->stage_cap, 4 * fl->total, fl->stream);
  if (rc) return rc;
  A3D_HIP(hipMemcpyAsync(fl->stage, z4, 4 * fl->total * 4, hipMemcpyHostToDevice, fl->stream));
  for (uint32_t f = 0; f < fl->F; ++f) {
    FilterDev d = fl->base(f);
    for (int k = 0; k < 4; ++k) {
      d.delta[k] = deltas[4 * (size_t)f + k];
      d.dev[k] = std::fabs(deltas[4 * (size_t)f + k] * mods[4 * (size_t)f + k]);
    }
    d.z4 = fl->stage + 4 * fl->poff[f];
    fl->h_filters[f] = d;
  }
  rc = upload_filters(fl);
  if (rc) return rc;
  launch_predict(fl->device_filters(), fl->stream);
  A3D_HIP(hipGetLastError());
  A3D_HIP(hipStreamSynchronize(fl->stream));
  return A3D_OK;
}

int a3d_fleet_stage_update(a3d_fleet* fl, const float* x, const float* y, const float* z, const uint32_t* m,
                           const float* ranges, const uint32_t* nb, const double* alpha, const double* sigma,
                           const double* roll, const double* pitch) {
  if (!fl || !m || !nb || !alpha || !sigma || !roll || !pitch)
    return fail(A3D_ERR_ARG, "a3d_fleet_stage_update: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_stage_update")) return rc;
  fl->layout = false;
  std::vector<uint64_t> moff(fl->F + 1, 0), boff(fl->F + 1, 0);
  for (uint32_t f = 0; f < fl->F; ++f) {
    moff[f + 1] = moff[f] + m[f];
    boff[f + 1] = boff[f] + nb[f];
  }
  const uint64_t M = moff[fl->F], B = boff[fl->F];
  if ((M && (!x || !y || !z)) || (B && !ranges)) return fail(A3D_ERR_ARG, "a3d_fleet_stage_update: null data");
  int rc = grow(fl->cloud, fl->cloud_cap, 3 * M + 4 * B + 16, fl->stream);
  if (rc) return rc;
  float* cx = fl->cloud;
  float* cy = cx + M;
  float* cz = cy + M;
  float* rg = cz + M;
  if (M) {
    A3D_HIP(hipMemcpyAsync(cx, x, M * 4, hipMemcpyHostToDevice, fl->stream));
    A3D_HIP(hipMemcpyAsync(cy, y, M * 4, hipMemcpyHostToDevice, fl->stream));
    A3D_HIP(hipMemcpyAsync(cz, z, M * 4, hipMemcpyHostToDevice, fl->stream));
  }
  if (B) A3D_HIP(hipMemcpyAsync(rg, ranges, B * 16, hipMemcpyHostToDevice, fl->stream));
  for (uint32_t f = 0; f < fl->F; ++f) {
    FilterDev d = fl->base(f);
    d.cx = cx + moff[f];
    d.cy = cy + moff[f];
    d.cz = cz + moff[f];
    d.m = m[f];
    d.ranges = nb[f] ? rg + 4 * boff[f] : nullptr;
    d.nb = nb[f];
    d.alpha = alpha[f];
    d.sigma = sigma[f];
    d.roll = (float)roll[f];
    d.pitch = (float)pitch[f];
    d.rp0 = (d.roll == 0.0f && d.pitch == 0.0f) ? 1u : 0u;
    fl->h_filters[f] = d;
  }
  // particles visited in spatial order (computed per update, before the weight launch)
  fl->items_spatial = current_order() == A3D_ORDER_SPATIAL;
  for (uint32_t f = 0; f < fl->F; ++f) fl->h_filters[f].order = fl->items_spatial ? fl->h_ord[f].order : nullptr;
  rc = upload_filters(fl);
  if (rc) return rc;
  rc = build_items(fl, m);
  if (rc) return rc;
  fl->staged = true;
  return A3D_OK;
}

int a3d_fleet_update_staged(a3d_fleet* fl, const a3d_grid* g, int persistent) {
  if (!fl || !g) return fail(A3D_ERR_ARG, "a3d_fleet_update_staged: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_update_staged")) return rc;
  if (!fl->staged) return fail(A3D_ERR_STATE, "a3d_fleet_update_staged: call a3d_fleet_stage_update first");
  if (persistent >= 0 && current_mapping() != fl->items_mapping)
    return fail(A3D_ERR_STATE, "a3d_fleet_update_staged: the mapping changed since staging; stage again");
  auto ev = fl->timing.take();
  if (ev.first) (void)hipEventRecord(ev.first, fl->stream);
  if (persistent >= 0) {
    // the list was cut for the tile current at staging time; launch with that tile
    if (fl->items_spatial) launch_order_many(fl->d_filters, fl->d_ord, fl->F, fl->max_n, fl->stream);
    launch_weight(g->g, fl->device_filters(), fl->d_items, fl->nitems, fl->d_counter, fl->items_ppw, persistent,
                  fl->stream);
  } else {
    // the plain single-filter path (a3d_pf_update's launches), once per filter
    for (uint32_t f = 0; f < fl->F; ++f) {
      if (fl->items_spatial) launch_order(fl->h_filters[f], fl->h_ord[f], fl->stream);
      launch_weight(g->g, single(fl->h_filters[f]), nullptr, 0, nullptr, current_ppw(), 0, fl->stream);
    }
  }
  if (ev.first) {
    (void)hipEventRecord(ev.second, fl->stream);
    fl->timing.pending.push_back(ev);
    fl->timing.launches++;
    uint64_t e = 0;
    for (uint32_t f = 0; f < fl->F; ++f) e += (uint64_t)fl->n[f] * fl->st_m[f];
    fl->timing.evals += e;
  }
  if (persistent >= 0) {
    launch_update_tail(g->g, fl->device_filters(), fl->stream);
  } else {
    for (uint32_t f = 0; f < fl->F; ++f) launch_update_tail(g->g, single(fl->h_filters[f]), fl->stream);
  }
  fl->prefix_valid = true;
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_fleet_update(a3d_fleet* fl, const a3d_grid* g, const float* x, const float* y, const float* z,
                     const uint32_t* m, const float* ranges, const uint32_t* nb, const double* alpha,
                     const double* sigma, const double* roll, const double* pitch) {
  if (!g) return fail(A3D_ERR_ARG, "a3d_fleet_update: null grid");
  if (fl && fl->capturing) return refuse_capture(fl, "a3d_fleet_update");
  int rc = a3d_fleet_stage_update(fl, x, y, z, m, ranges, nb, alpha, sigma, roll, pitch);
  if (rc) return rc;
  rc = a3d_fleet_update_staged(fl, g, 1);
  if (rc) return rc;
  A3D_HIP(hipStreamSynchronize(fl->stream));
  return A3D_OK;
}

int a3d_fleet_resample(a3d_fleet* fl, const float* u01, int mode) {
  if (!fl || !u01) return fail(A3D_ERR_ARG, "a3d_fleet_resample: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_resample")) return rc;
  if (mode != A3D_RESAMPLE_REFERENCE && mode != A3D_RESAMPLE_FP64)
    return fail(A3D_ERR_ARG, "a3d_fleet_resample: unknown mode");
  for (uint32_t f = 0; f < fl->F; ++f) {
    FilterDev d = fl->h_filters[f].x ? fl->h_filters[f] : fl->base(f);
    d.u01 = u01[f];
    d.idx_out = nullptr;
    fl->h_filters[f] = d;
  }
  int rc = upload_filters(fl);
  if (rc) return rc;
  launch_resample(fl->device_filters(), mode, fl->prefix_valid, fl->stream);
  fl->prefix_valid = false;
  A3D_HIP(hipGetLastError());
  A3D_HIP(hipStreamSynchronize(fl->stream));
  return A3D_OK;
}

int a3d_fleet_get_means(const a3d_fleet* fl, float* means4) {
  if (!fl || !means4) return fail(A3D_ERR_ARG, "a3d_fleet_get_means: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_get_means")) return rc;
  A3D_HIP(hipMemcpyAsync(means4, fl->mean, 16 * (size_t)fl->F, hipMemcpyDeviceToHost, fl->stream));
  A3D_HIP(hipStreamSynchronize(fl->stream));
  return A3D_OK;
}

int a3d_fleet_synchronize(a3d_fleet* fl) {
  if (!fl) return fail(A3D_ERR_ARG, "a3d_fleet_synchronize: null fleet");
  if (int rc = refuse_capture(fl, "a3d_fleet_synchronize")) return rc;
  A3D_HIP(hipStreamSynchronize(fl->stream));
  return A3D_OK;
}

int a3d_fleet_timing(a3d_fleet* fl, int reset, double* weight_ms_avg, uint64_t* launches, uint64_t* evals) {
  if (!fl) return fail(A3D_ERR_ARG, "a3d_fleet_timing: null fleet");
  if (int rc = refuse_capture(fl, "a3d_fleet_timing")) return rc;
  A3D_HIP(hipStreamSynchronize(fl->stream));
  fl->timing.resolve();
  if (weight_ms_avg) *weight_ms_avg = fl->timing.launches ? fl->timing.acc_ms / fl->timing.launches : 0.0;
  if (launches) *launches = fl->timing.launches;
  if (evals) *evals = fl->timing.evals;
  if (reset) {
    fl->timing.acc_ms = 0.0;
    fl->timing.launches = 0;
    fl->timing.evals = 0;
  }
  return A3D_OK;
}

// ------------------------------------------------ device-parameter forms
int a3d_fleet_set_layout(a3d_fleet* fl, const uint32_t* m, const uint32_t* nb) {
  if (!fl || !m || !nb) return fail(A3D_ERR_ARG, "a3d_fleet_set_layout: null argument");
  if (int rc = refuse_capture(fl, "a3d_fleet_set_layout")) return rc;
  std::vector<uint64_t> lay(3 * (size_t)fl->F);
  uint64_t mo = 0, bo = 0;
  for (uint32_t f = 0; f < fl->F; ++f) {
    lay[3 * (size_t)f] = mo;
    lay[3 * (size_t)f + 1] = bo;
    lay[3 * (size_t)f + 2] = fl->poff[f];
    mo += m[f];
    bo += nb[f];
  }
  fl->layout = false;
  fl->staged = false;
  if (!fl->d_lay && hipMalloc(&fl->d_lay, lay.size() * sizeof(uint64_t)) != hipSuccess) {
    fl->d_lay = nullptr;
    return fail(A3D_ERR_NOMEM, "a3d_fleet_set_layout: device allocation failed");
  }
  fl->items_spatial = current_order() == A3D_ORDER_SPATIAL;
  for (uint32_t f = 0; f < fl->F; ++f) {
    FilterDev d = fl->base(f);
    d.m = m[f];
    d.nb = nb[f];
    d.order = fl->items_spatial ? fl->h_ord[f].order : nullptr;
    fl->h_filters[f] = d;
  }
  int rc = upload_filters(fl);
  if (rc) return rc;
  A3D_HIP(hipMemcpyAsync(fl->d_lay, lay.data(), lay.size() * sizeof(uint64_t), hipMemcpyHostToDevice, fl->stream));
  rc = build_items(fl, m);  // synchronises: the host arrays above are no longer read
  if (rc) return rc;
  fl->st_nb.assign(nb, nb + fl->F);
  fl->layout = true;
  return A3D_OK;
}

static int need_layout(const a3d_fleet* fl, const char* what) {
  if (!fl) return fail(A3D_ERR_ARG, std::string(what) + ": null fleet");
  if (!fl->layout) return fail(A3D_ERR_STATE, std::string(what) + ": call a3d_fleet_set_layout first");
  return A3D_OK;
}

int a3d_fleet_predict_dev(a3d_fleet* fl, const double* d_mods, const double* d_deltas, const float* d_z4) {
  if (int rc = need_layout(fl, "a3d_fleet_predict_dev")) return rc;
  if (!d_mods || !d_deltas || !d_z4) return fail(A3D_ERR_ARG, "a3d_fleet_predict_dev: null argument");
  fleet_patch_predict_kernel<<<(fl->F + 63) / 64, 64, 0, fl->stream>>>(fl->d_filters, fl->F, d_mods, d_deltas,
                                                                      d_z4, fl->d_lay);
  launch_predict(fl->device_filters(), fl->stream);
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_fleet_update_dev(a3d_fleet* fl, const a3d_grid* g, const float* d_x, const float* d_y, const float* d_z,
                         const float* d_ranges, const double* d_params) {
  if (int rc = need_layout(fl, "a3d_fleet_update_dev")) return rc;
  if (!g || !d_params) return fail(A3D_ERR_ARG, "a3d_fleet_update_dev: null argument");
  if (current_mapping() != fl->items_mapping)
    return fail(A3D_ERR_STATE, "a3d_fleet_update_dev: the mapping changed since a3d_fleet_set_layout; set it again");
  uint64_t M = 0, B = 0, evals = 0;
  for (uint32_t f = 0; f < fl->F; ++f) {
    M += fl->st_m[f];
    B += fl->st_nb[f];
    evals += (uint64_t)fl->n[f] * fl->st_m[f];
  }
  if ((M && (!d_x || !d_y || !d_z)) || (B && !d_ranges)) return fail(A3D_ERR_ARG, "a3d_fleet_update_dev: null data");
  fleet_patch_update_kernel<<<(fl->F + 63) / 64, 64, 0, fl->stream>>>(fl->d_filters, fl->F, d_x, d_y, d_z,
                                                                     d_ranges, d_params, fl->d_lay);
  if (fl->items_spatial) launch_order_many(fl->d_filters, fl->d_ord, fl->F, fl->max_n, fl->stream);
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (!fl->capturing) ev = fl->timing.take();
  if (ev.first) (void)hipEventRecord(ev.first, fl->stream);
  launch_weight(g->g, fl->device_filters(), fl->d_items, fl->nitems, fl->d_counter, fl->items_ppw, 1, fl->stream);
  if (ev.first) {
    (void)hipEventRecord(ev.second, fl->stream);
    fl->timing.pending.push_back(ev);
    fl->timing.launches++;
    fl->timing.evals += evals;
  }
  launch_update_tail(g->g, fl->device_filters(), fl->stream);
  fl->prefix_valid = true;
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_fleet_resample_dev(a3d_fleet* fl, const float* d_u01, int mode) {
  if (int rc = need_layout(fl, "a3d_fleet_resample_dev")) return rc;
  if (!d_u01) return fail(A3D_ERR_ARG, "a3d_fleet_resample_dev: null argument");
  if (mode != A3D_RESAMPLE_REFERENCE && mode != A3D_RESAMPLE_FP64)
    return fail(A3D_ERR_ARG, "a3d_fleet_resample_dev: unknown mode");
  Filters fs = fl->device_filters();
  fs.u01 = d_u01;
  launch_resample(fs, mode, fl->prefix_valid, fl->stream);
  fl->prefix_valid = false;
  A3D_HIP(hipGetLastError());
  return A3D_OK;
}

int a3d_fleet_capture_begin(a3d_fleet* fl) {
  if (int rc = need_layout(fl, "a3d_fleet_capture_begin")) return rc;
  if (fl->capturing) return fail(A3D_ERR_STATE, "a3d_fleet_capture_begin: already capturing");
  A3D_HIP(hipStreamSynchronize(fl->stream));
  A3D_HIP(hipStreamBeginCapture(fl->stream, hipStreamCaptureModeThreadLocal));
  fl->capturing = true;
  return A3D_OK;
}

int a3d_fleet_capture_end(a3d_fleet* fl, a3d_graph** out) {
  if (!fl || !out) return fail(A3D_ERR_ARG, "a3d_fleet_capture_end: null argument");
  if (!fl->capturing) return fail(A3D_ERR_STATE, "a3d_fleet_capture_end: not capturing");
  fl->capturing = false;
  a3d_graph* G = new a3d_graph();
  G->stream = fl->stream;
  if (!hip_ok(hipStreamEndCapture(fl->stream, &G->graph), "hipStreamEndCapture") ||
      !hip_ok(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
    if (G->graph) (void)hipGraphDestroy(G->graph);
    delete G;
    return A3D_ERR_HIP;
  }
  *out = G;
  return A3D_OK;
}

}  // extern "C"
