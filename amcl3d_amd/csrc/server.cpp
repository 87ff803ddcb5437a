// Serving loop: many localisation sessions on resident maps, continuous
// batching of their update requests.
//
// A session owns device-resident particles (structure of arrays, as a
// filter), its scratch, its mean and an RNG counter; its draws come from
// Philox keyed by the session seed and indexed by the session's step, so a
// session's results are the same whichever requests share its batches.
// Sessions are admitted while their device bytes fit in the HBM budget.
// With latest_only, a session keeps at most one queued request: a newer scan
// supersedes the queued one (as the reference's queue-size-1 subscriber
// drops stale scans, Node.cpp:60), and the two odometry increments compose.
//
// A worker thread forms batches from the queue (arrival order, at most one
// request per session in flight, bounded by sessions and evaluations per
// batch), stages their clouds through pinned host memory, and launches on
// one of `max_inflight` streams: draws, predict, ONE cloud-weight launch over
// all requests' particle blocks, the update tail, resample.  Requests that
// arrive while batches run join the next batch as soon as a stream frees, so
// launches overlap on the device.  Latency is submit -> mean on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "a3d_handles.h"

using namespace a3d;

namespace {

// Under a backlog a batch may grow up to this many times the configured
// bound (max_batch_evals).
constexpr uint64_t kBacklogGrowth = 8;

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Session {
  bool open = false;
  int map = 0;
  uint32_t n = 0;
  uint64_t seed = 0;
  uint64_t step = 0;  // draws index of the next predict (0 was init)
  float* state = nullptr;
  float* scratch = nullptr;
  uint8_t* inmap = nullptr;
  float* mean = nullptr;
  uint8_t* ord = nullptr;  // visiting order buffers (order_view)
  size_t bytes = 0;
  bool busy = false;

  OrderDev ov() const { return order_view(ord, n); }

  FilterDev dev() const {
    FilterDev d{};
    d.x = state;
    d.y = state + n;
    d.z = state + 2 * (size_t)n;
    d.a = state + 3 * (size_t)n;
    d.w = state + 4 * (size_t)n;
    d.wp = state + 5 * (size_t)n;
    d.wr = state + 6 * (size_t)n;
    d.scratch = scratch;
    d.zr = scratch ? scratch + sc_sums(n) + kSumZr : nullptr;
    d.inmap = inmap;
    d.mean = mean;
    d.n = n;
    return d;
  }
};

struct Request {
  uint64_t ticket = 0;
  int session = 0;
  double mods[4] = {0, 0, 0, 0};
  double deltas[4] = {0, 0, 0, 0};
  std::vector<float> cloud;  // x[m], y[m], z[m]
  std::vector<float> ranges;
  uint32_t m = 0, nb = 0;
  double alpha = 0, sigma = 0, roll = 0, pitch = 0;
  int resample = 0;
  double submit_ms = 0;
};

// Odometry increment a followed by b (both relative to the pose before them),
// as Node.cpp:147-151 would read it from lastupdatebase^-1 * base.
void compose_increment(const double a[4], const double b[4], double out[4]) {
  const double c = std::cos(a[3]), s = std::sin(a[3]);
  out[0] = a[0] + (c * b[0] - s * b[1]);
  out[1] = a[1] + (s * b[0] + c * b[1]);
  out[2] = a[2] + b[2];
  out[3] = a[3] + b[3];
}

// Device and pinned buffers of one in-flight batch.
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request> reqs;
  // pinned host staging
  float* h_data = nullptr;
  size_t h_cap = 0;  // floats
  float* h_mean = nullptr;
  size_t h_mean_cap = 0;
  // device
  float* d_data = nullptr;  // clouds, ranges
  size_t d_cap = 0;
  float* d_z4 = nullptr;
  size_t z4_cap = 0;
  FilterDev* d_filters = nullptr;
  size_t filt_cap = 0;
  OrderDev* d_ord = nullptr;
  size_t ord_cap = 0;
  float* d_means = nullptr;
  size_t means_cap = 0;
  WorkItem* d_items = nullptr;
  size_t items_cap = 0;
  uint64_t* d_seedstep = nullptr;
  size_t ss_cap = 0;
  float* d_u01 = nullptr;
  uint32_t* d_counter = nullptr;
  std::vector<FilterDev> h_filters;
  std::vector<OrderDev> h_ord;
  std::vector<uint64_t> h_seedstep;
  std::vector<WorkItem> h_items;
  uint64_t evals = 0;
};

template <typename T>
bool grow_dev(T*& p, size_t& cap, size_t count) {
  if (count <= cap) return true;
  (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = std::max(count, (size_t)1024) * 3 / 2;
  if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) return false;
  cap = want;
  return true;
}

bool grow_pinned(float*& p, size_t& cap, size_t count) {
  if (count <= cap) return true;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = std::max(count, (size_t)1 << 16) * 3 / 2;
  if (hipHostMalloc((void**)&p, want * sizeof(float), hipHostMallocDefault) != hipSuccess) return false;
  cap = want;
  return true;
}

}  // namespace

struct a3d_server {
  a3d_server_config cfg{};
  int device = 0;
  std::vector<const a3d_grid*> maps;
  std::vector<GridDev*> d_maps;  // device copy of each map's view (mixed-map batches)
  std::vector<Session> sessions;
  size_t hbm_used = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> pending;
  std::deque<a3d_result> results;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  std::string worker_error;
  uint64_t next_ticket = 1;
  uint64_t inflight_requests = 0;
  uint64_t st_requests = 0, st_batches = 0, st_evals = 0;
  double st_busy_ms = 0;
  hipStream_t admin = nullptr;  // session init / readback

  void run();
  bool launch(Slot& s);
  void retire(Slot& s);
  // Waits for every batch enqueued on the slot streams.
  void sync_slots() {
    for (auto& s : slots)
      if (s.stream) (void)hipStreamSynchronize(s.stream);
  }
};

static int alloc_session(a3d_server* S, Session& se) {
  const size_t n = se.n;
  se.bytes = 7 * n * 4 + scratch_floats((uint32_t)n) * 4 + n + 64 + order_stride((uint32_t)n);
  if (S->hbm_used + se.bytes > S->cfg.hbm_budget_bytes)
    return fail(A3D_ERR_NOMEM, "serving: session does not fit in the HBM budget");
  if (hipMalloc(&se.state, 7 * n * 4) != hipSuccess || hipMalloc(&se.scratch, scratch_floats((uint32_t)n) * 4) != hipSuccess ||
      hipMalloc(&se.inmap, n) != hipSuccess || hipMalloc(&se.mean, 64) != hipSuccess ||
      hipMalloc(&se.ord, order_stride((uint32_t)n)) != hipSuccess) {
    (void)hipFree(se.state);
    (void)hipFree(se.scratch);
    (void)hipFree(se.inmap);
    (void)hipFree(se.mean);
    (void)hipFree(se.ord);
    se.state = se.scratch = se.mean = nullptr;
    se.inmap = se.ord = nullptr;
    return fail(A3D_ERR_NOMEM, "serving: device allocation failed");
  }
  S->hbm_used += se.bytes;
  return A3D_OK;
}

static void free_session(a3d_server* S, Session& se) {
  (void)hipFree(se.state);
  (void)hipFree(se.scratch);
  (void)hipFree(se.inmap);
  (void)hipFree(se.mean);
  (void)hipFree(se.ord);
  se.state = se.scratch = se.mean = nullptr;
  se.inmap = se.ord = nullptr;
  S->hbm_used -= std::min(S->hbm_used, se.bytes);
  se.open = false;
}

bool a3d_server::launch(Slot& s) {
  // sizes
  size_t nf = s.reqs.size(), data = 0, z4 = 0;
  for (auto& r : s.reqs) {
    data += 3 * (size_t)r.m + 4 * (size_t)r.nb;
    z4 += 4 * (size_t)sessions[r.session].n;
  }
  if (!grow_pinned(s.h_data, s.h_cap, data + 16) || !grow_dev(s.d_data, s.d_cap, data + 16) ||
      !grow_dev(s.d_z4, s.z4_cap, z4 + 16) || !grow_dev(s.d_filters, s.filt_cap, nf) ||
      !grow_dev(s.d_ord, s.ord_cap, nf) ||
      !grow_dev(s.d_seedstep, s.ss_cap, 2 * nf) || !grow_pinned(s.h_mean, s.h_mean_cap, 4 * nf)) {
    worker_error = "serving: staging allocation failed";
    return false;
  }
  if (!grow_dev(s.d_means, s.means_cap, 4 * nf)) {
    worker_error = "serving: staging allocation failed";
    return false;
  }
  s.h_filters.clear();
  s.h_ord.clear();
  s.h_seedstep.clear();
  s.h_items.clear();
  const int ppw = clamp_ppw(current_ppw());
  const uint32_t per_item = (uint32_t)item_particles(current_mapping(), ppw);
  const bool spatial = current_order() == A3D_ORDER_SPATIAL;
  size_t off = 0, zoff = 0;
  uint32_t max_n = 0;
  s.evals = 0;
  for (size_t i = 0; i < nf; ++i) {
    Request& r = s.reqs[i];
    Session& se = sessions[r.session];
    FilterDev d = se.dev();
    std::memcpy(s.h_data + off, r.cloud.data(), 3 * (size_t)r.m * 4);
    d.cx = s.d_data + off;
    d.cy = d.cx + r.m;
    d.cz = d.cy + r.m;
    d.m = r.m;
    off += 3 * (size_t)r.m;
    if (r.nb) std::memcpy(s.h_data + off, r.ranges.data(), 4 * (size_t)r.nb * 4);
    d.ranges = r.nb ? s.d_data + off : nullptr;
    d.nb = r.nb;
    off += 4 * (size_t)r.nb;
    d.alpha = r.alpha;
    d.sigma = r.sigma;
    d.roll = (float)r.roll;
    d.pitch = (float)r.pitch;
    d.rp0 = (d.roll == 0.0f && d.pitch == 0.0f) ? 1u : 0u;
    for (int k = 0; k < 4; ++k) {
      d.delta[k] = r.deltas[k];
      d.dev[k] = std::fabs(r.deltas[k] * r.mods[k]);
    }
    d.z4 = s.d_z4 + zoff;
    d.no_resample = r.resample ? 0u : 1u;
    d.order = spatial ? se.ov().order : nullptr;
    zoff += 4 * (size_t)se.n;
    s.h_filters.push_back(d);
    s.h_ord.push_back(se.ov());
    s.h_seedstep.push_back(se.seed);
    max_n = std::max(max_n, se.n);
    for (uint32_t p0 = 0; p0 < se.n; p0 += per_item)
      s.h_items.push_back({(uint32_t)i, p0, std::min<uint32_t>(per_item, se.n - p0), 0});
    s.evals += (uint64_t)se.n * r.m;
  }
  for (size_t i = 0; i < nf; ++i) s.h_seedstep.push_back(sessions[s.reqs[i].session].step);
  // a batch over several maps: every filter names its grid, and one launch
  // serves them all
  const int map0 = sessions[s.reqs[0].session].map;
  bool mixed = false;
  for (auto& r : s.reqs) mixed |= sessions[r.session].map != map0;
  if (mixed)
    for (size_t i = 0; i < nf; ++i) s.h_filters[i].gdev = d_maps[sessions[s.reqs[i].session].map];
  if (!grow_dev(s.d_items, s.items_cap, s.h_items.size())) {
    worker_error = "serving: staging allocation failed";
    return false;
  }
  hipStream_t st = s.stream;
  bool ok = hipMemcpyAsync(s.d_data, s.h_data, off * 4, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(s.d_filters, s.h_filters.data(), nf * sizeof(FilterDev), hipMemcpyHostToDevice, st) ==
                hipSuccess &&
            hipMemcpyAsync(s.d_ord, s.h_ord.data(), nf * sizeof(OrderDev), hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(s.d_seedstep, s.h_seedstep.data(), 2 * nf * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(s.d_items, s.h_items.data(), s.h_items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                           st) == hipSuccess;
  if (!ok) {
    worker_error = "serving: staging copy failed";
    return false;
  }
  Filters fs{};
  fs.arr = s.d_filters;
  fs.count = (uint32_t)nf;
  fs.max_n = max_n;
  launch_draws(fs, s.d_seedstep, s.d_seedstep + nf, s.d_u01, st);
  launch_predict(fs, st);
  // the launch's grid: the widest index kind any map of the batch needs
  // (64-bit indices, or 32-bit ones through 64-bit addresses), clean only
  // if every map is
  GridDev gl = maps[map0]->g;
  for (auto& r : s.reqs) {
    gl.idx64 |= maps[sessions[r.session].map]->g.idx64;
    gl.buf32 &= maps[sessions[r.session].map]->g.buf32;
    gl.clean &= maps[sessions[r.session].map]->g.clean;
  }
  // No segment loads in served batches: in batches that mix a large map with
  // smaller ones they measured slower than 64-bit addresses (the
  // 1,024-session serving mix: 9.4e11 against 1.04e12 evals/s), on one map
  // about as fast (+1.4%), and their cloud z-range kernel is one more launch
  // in each batch's chain.
  gl.seg_bits = 0u;
  if (spatial) launch_order_many(s.d_filters, s.d_ord, (uint32_t)nf, max_n, st);  // after predict moved them
  launch_weight(gl, fs, s.d_items, (uint32_t)s.h_items.size(), s.d_counter, ppw, 1, st);
  launch_update_tail(maps[map0]->g, fs, st);
  // means to the host before resample (getMean is read after update, Node.cpp:167)
  launch_gather_means(fs, s.d_means, st);
  (void)hipMemcpyAsync(s.h_mean, s.d_means, 16 * nf, hipMemcpyDeviceToHost, st);
  // resample the requests that ask for it, each with the uniform of its draws
  bool any_rs = false;
  for (auto& r : s.reqs) any_rs |= r.resample != 0;
  if (any_rs) {
    Filters fr = fs;
    fr.u01 = s.d_u01;
    launch_resample(fr, A3D_RESAMPLE_REFERENCE, true, st);
  }
  (void)hipEventRecord(s.done, st);
  if (hipGetLastError() != hipSuccess) {
    worker_error = "serving: launch failed";
    return false;
  }
  for (auto& r : s.reqs) sessions[r.session].step++;
  s.busy = true;
  st_batches++;
  return true;
}

void a3d_server::retire(Slot& s) {
  const double t = now_ms();
  for (size_t i = 0; i < s.reqs.size(); ++i) {
    Request& r = s.reqs[i];
    a3d_result res{};
    res.ticket = r.ticket;
    res.session = r.session;
    std::memcpy(res.mean, s.h_mean + 4 * i, 16);
    res.latency_ms = t - r.submit_ms;
    res.submit_ms = r.submit_ms;
    res.evals = (uint64_t)sessions[r.session].n * r.m;
    std::memcpy(res.deltas, r.deltas, 32);
    res.superseded = 0;
    results.push_back(res);
    sessions[r.session].busy = false;
    st_evals += res.evals;
    st_requests++;
  }
  inflight_requests -= s.reqs.size();
  s.reqs.clear();
  s.busy = false;
}

void a3d_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  double busy_since = -1;
  while (!stop) {
    bool progress = false;
    for (auto& s : slots) {
      if (s.busy && hipEventQuery(s.done) == hipSuccess) {
        retire(s);
        progress = true;
      }
    }
    // the batch bound: the configured one, or under a backlog the queued
    // evaluations shared among the free slots (at most kBacklogGrowth x)
    uint64_t bound = cfg.max_batch_evals;
    {
      int free_slots = 0;
      for (auto& s : slots) free_slots += s.busy ? 0 : 1;
      if (free_slots > 0 && !pending.empty()) {
        uint64_t backlog = 0;
        for (auto& r : pending)
          if (!sessions[r.session].busy) backlog += (uint64_t)sessions[r.session].n * r.m;
        bound = std::max(bound, std::min(kBacklogGrowth * cfg.max_batch_evals, backlog / (uint64_t)free_slots));
      }
    }
    for (auto& s : slots) {
      if (s.busy || pending.empty()) continue;
      // form a batch in arrival order: one request per session, bounded size
      uint64_t evals = 0;
      for (auto it = pending.begin(); it != pending.end();) {
        if ((int)s.reqs.size() >= cfg.max_batch_sessions) break;
        Session& se = sessions[it->session];
        bool dup = se.busy;
        if (!dup) {
          const uint64_t e = (uint64_t)se.n * it->m;
          if (!s.reqs.empty() && evals + e > bound) break;
          evals += e;
          se.busy = true;
          s.reqs.push_back(std::move(*it));
          it = pending.erase(it);
        } else {
          ++it;
        }
      }
      if (s.reqs.empty()) continue;
      if (!launch(s)) {
        if (worker_error.empty()) worker_error = "serving: launch failed";
        stop = true;
        break;
      }
      progress = true;
    }
    bool any_busy = false;
    for (auto& s : slots) any_busy |= s.busy;
    const double t = now_ms();
    if (any_busy && busy_since < 0) busy_since = t;
    if (!any_busy && busy_since >= 0) {
      st_busy_ms += t - busy_since;
      busy_since = -1;
    }
    cv.notify_all();
    if (!progress) {
      if (any_busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        cv.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
  }
  if (busy_since >= 0) st_busy_ms += now_ms() - busy_since;
  cv.notify_all();
}

extern "C" {

int a3d_server_create(const a3d_server_config* cfg, a3d_server** out) {
  if (!cfg || !out) return fail(A3D_ERR_ARG, "a3d_server_create: null argument");
  if (cfg->max_inflight < 1 || cfg->max_inflight > 16 || cfg->max_batch_sessions < 1 ||
      cfg->max_batch_sessions > 4096 || cfg->max_batch_evals == 0)
    return fail(A3D_ERR_ARG, "a3d_server_create: bad configuration");
  a3d_server* S = new a3d_server();
  S->cfg = *cfg;
  (void)hipGetDevice(&S->device);
  S->slots.resize(cfg->max_inflight);
  bool ok = hipStreamCreateWithFlags(&S->admin, hipStreamNonBlocking) == hipSuccess;
  for (auto& s : S->slots) {
    ok = ok && hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&s.done, hipEventDisableTiming) == hipSuccess &&
         hipMalloc(&s.d_counter, 64) == hipSuccess && hipMalloc(&s.d_u01, 4 * 4096) == hipSuccess;
  }
  if (!ok) {
    a3d_server_destroy(S);
    return fail(A3D_ERR_HIP, "a3d_server_create: stream or buffer creation failed");
  }
  S->worker = std::thread([S]() { S->run(); });
  *out = S;
  return A3D_OK;
}

int a3d_server_destroy(a3d_server* S) {
  if (!S) return A3D_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (auto& s : S->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    (void)hipHostFree(s.h_data);
    (void)hipHostFree(s.h_mean);
    (void)hipFree(s.d_data);
    (void)hipFree(s.d_z4);
    (void)hipFree(s.d_filters);
    (void)hipFree(s.d_ord);
    (void)hipFree(s.d_means);
    (void)hipFree(s.d_items);
    (void)hipFree(s.d_seedstep);
    (void)hipFree(s.d_u01);
    (void)hipFree(s.d_counter);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
  }
  for (auto& se : S->sessions)
    if (se.open) free_session(S, se);
  for (GridDev* d : S->d_maps) (void)hipFree(d);
  if (S->admin) (void)hipStreamDestroy(S->admin);
  delete S;
  return A3D_OK;
}

int a3d_server_add_map(a3d_server* S, const a3d_grid* g, int* map_id) {
  if (!S || !g || !map_id) return fail(A3D_ERR_ARG, "a3d_server_add_map: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  GridDev* d = nullptr;
  if (hipMalloc(&d, sizeof(GridDev)) != hipSuccess) return fail(A3D_ERR_NOMEM, "a3d_server_add_map: allocation");
  if (hipMemcpy(d, &g->g, sizeof(GridDev), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return fail(A3D_ERR_HIP, "a3d_server_add_map: copy");
  }
  S->maps.push_back(g);
  S->d_maps.push_back(d);
  *map_id = (int)S->maps.size() - 1;
  return A3D_OK;
}

int a3d_server_open_session(a3d_server* S, int map_id, int num_particles, const float pose[4], const float devs[4],
                            uint64_t seed, int* session) {
  if (!S || !pose || !devs || !session || num_particles <= 0)
    return fail(A3D_ERR_ARG, "a3d_server_open_session: bad argument");
  std::lock_guard<std::mutex> lk(S->mu);
  if (map_id < 0 || map_id >= (int)S->maps.size()) return fail(A3D_ERR_ARG, "a3d_server_open_session: bad map");
  Session se;
  se.map = map_id;
  se.n = (uint32_t)num_particles;
  se.seed = seed;
  int rc = alloc_session(S, se);
  if (rc) return rc;
  // ParticleFilter::init with the session's step-0 draws
  float* z4 = nullptr;
  uint64_t* ss = nullptr;
  if (hipMalloc(&z4, 4 * (size_t)se.n * 4) != hipSuccess || hipMalloc(&ss, 16) != hipSuccess) {
    (void)hipFree(z4);
    free_session(S, se);
    return fail(A3D_ERR_NOMEM, "a3d_server_open_session: init staging");
  }
  const uint64_t h[2] = {seed, 0};
  FilterDev d = se.dev();
  const float dev = std::max(std::max(devs[0], devs[1]), devs[2]);
  d.dev[0] = (float)(1. / (dev * std::sqrt(2 * M_PI)));
  d.dev[1] = (float)(1. / (2 * dev * dev));
  for (int k = 0; k < 4; ++k) {
    d.delta[k] = devs[k];
    d.pose[k] = pose[k];
  }
  d.z4 = z4;
  bool ok = hipMemsetAsync(se.state, 0, 7 * (size_t)se.n * 4, S->admin) == hipSuccess &&
            hipMemcpyAsync(ss, h, 16, hipMemcpyHostToDevice, S->admin) == hipSuccess;
  if (ok) {
    launch_draws(single(d), ss, ss + 1, nullptr, S->admin);
    launch_init(single(d), S->admin);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(S->admin) == hipSuccess;
  }
  (void)hipFree(z4);
  (void)hipFree(ss);
  if (!ok) {
    free_session(S, se);
    return fail(A3D_ERR_HIP, "a3d_server_open_session: init failed");
  }
  se.open = true;
  se.step = 1;
  S->sessions.push_back(se);
  *session = (int)S->sessions.size() - 1;
  return A3D_OK;
}

int a3d_server_open_session_particles(a3d_server* S, int map_id, const float* p7, size_t num_particles,
                                      uint64_t seed, int* session) {
  if (!S || !p7 || !session || num_particles == 0 || num_particles > 0x7fffffffull)
    return fail(A3D_ERR_ARG, "a3d_server_open_session_particles: bad argument");
  std::lock_guard<std::mutex> lk(S->mu);
  if (map_id < 0 || map_id >= (int)S->maps.size())
    return fail(A3D_ERR_ARG, "a3d_server_open_session_particles: bad map");
  Session se;
  se.map = map_id;
  se.n = (uint32_t)num_particles;
  se.seed = seed;
  int rc = alloc_session(S, se);
  if (rc) return rc;
  const size_t n = se.n;
  std::vector<float> soa(7 * n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 7; ++k) soa[k * n + i] = p7[7 * i + k];
  const float zero[4] = {0.f, 0.f, 0.f, 0.f};
  const bool ok = hipMemcpyAsync(se.state, soa.data(), 7 * n * 4, hipMemcpyHostToDevice, S->admin) == hipSuccess &&
                  hipMemcpyAsync(se.mean, zero, 16, hipMemcpyHostToDevice, S->admin) == hipSuccess &&
                  hipStreamSynchronize(S->admin) == hipSuccess;
  if (!ok) {
    free_session(S, se);
    return fail(A3D_ERR_HIP, "a3d_server_open_session_particles: upload failed");
  }
  se.open = true;
  se.step = 1;  // step 0's draws belong to init, which the caller's particles replace
  S->sessions.push_back(se);
  *session = (int)S->sessions.size() - 1;
  return A3D_OK;
}

int a3d_server_close_session(a3d_server* S, int session) {
  if (!S) return fail(A3D_ERR_ARG, "a3d_server_close_session: null server");
  std::unique_lock<std::mutex> lk(S->mu);
  if (session < 0 || session >= (int)S->sessions.size() || !S->sessions[session].open)
    return fail(A3D_ERR_ARG, "a3d_server_close_session: bad session");
  for (auto& r : S->pending)
    if (r.session == session) return fail(A3D_ERR_BUSY, "a3d_server_close_session: requests pending");
  S->cv.wait(lk, [&]() { return !S->sessions[session].busy || S->stop; });
  // still busy: the worker stopped on an error with the session's batch
  // enqueued; its stream may still read or write the session's state
  if (S->sessions[session].busy) S->sync_slots();
  free_session(S, S->sessions[session]);
  return A3D_OK;
}

int a3d_server_submit(a3d_server* S, int session, const double mods[4], const double deltas[4], const float* x,
                      const float* y, const float* z, size_t m, const float* ranges, size_t nb, double alpha,
                      double sigma, double roll, double pitch, int resample, uint64_t* ticket) {
  if (!S || !mods || !deltas || (m && (!x || !y || !z)) || (nb && !ranges) || m > 0xffffffffull)
    return fail(A3D_ERR_ARG, "a3d_server_submit: bad argument");
  Request r;
  r.session = session;
  std::memcpy(r.mods, mods, 32);
  std::memcpy(r.deltas, deltas, 32);
  r.m = (uint32_t)m;
  r.nb = (uint32_t)nb;
  r.cloud.resize(3 * m);
  if (m) {
    std::memcpy(r.cloud.data(), x, m * 4);
    std::memcpy(r.cloud.data() + m, y, m * 4);
    std::memcpy(r.cloud.data() + 2 * m, z, m * 4);
  }
  if (nb) r.ranges.assign(ranges, ranges + 4 * nb);
  r.alpha = alpha;
  r.sigma = sigma;
  r.roll = roll;
  r.pitch = pitch;
  r.resample = resample;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->worker_error.empty()) return fail(A3D_ERR_HIP, S->worker_error);
    if (session < 0 || session >= (int)S->sessions.size() || !S->sessions[session].open)
      return fail(A3D_ERR_ARG, "a3d_server_submit: bad session");
    r.ticket = S->next_ticket++;
    r.submit_ms = now_ms();
    if (ticket) *ticket = r.ticket;
    if (S->cfg.latest_only) {
      for (auto it = S->pending.begin(); it != S->pending.end(); ++it) {
        if (it->session != session) continue;
        // the queued request never runs: report it and fold its motion in
        a3d_result old{};
        old.ticket = it->ticket;
        old.session = session;
        old.latency_ms = r.submit_ms - it->submit_ms;
        old.submit_ms = it->submit_ms;
        std::memcpy(old.deltas, it->deltas, 32);
        old.superseded = 1;
        S->results.push_back(old);
        double d[4];
        compose_increment(it->deltas, r.deltas, d);
        std::memcpy(r.deltas, d, 32);
        S->pending.erase(it);
        S->inflight_requests--;
        break;  // at most one queued request per session
      }
    }
    S->inflight_requests++;
    S->pending.push_back(std::move(r));
  }
  S->cv.notify_all();
  return A3D_OK;
}

int a3d_server_poll(a3d_server* S, a3d_result* out, size_t max, size_t* n) {
  if (!S || !n || (max && !out)) return fail(A3D_ERR_ARG, "a3d_server_poll: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  size_t k = 0;
  while (k < max && !S->results.empty()) {
    out[k++] = S->results.front();
    S->results.pop_front();
  }
  *n = k;
  if (!S->worker_error.empty()) return fail(A3D_ERR_HIP, S->worker_error);
  return A3D_OK;
}

int a3d_server_drain(a3d_server* S) {
  if (!S) return fail(A3D_ERR_ARG, "a3d_server_drain: null server");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv.wait(lk, [&]() { return S->inflight_requests == 0 || S->stop; });
  if (!S->worker_error.empty()) return fail(A3D_ERR_HIP, S->worker_error);
  return A3D_OK;
}

int a3d_server_get_particles(a3d_server* S, int session, float* p7, size_t n) {
  if (!S || !p7) return fail(A3D_ERR_ARG, "a3d_server_get_particles: null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (session < 0 || session >= (int)S->sessions.size() || !S->sessions[session].open)
    return fail(A3D_ERR_ARG, "a3d_server_get_particles: bad session");
  S->cv.wait(lk, [&]() { return !S->sessions[session].busy || S->stop; });
  if (S->sessions[session].busy) {
    // the worker stopped with this session's batch enqueued: its state is
    // not the result of a completed update
    S->sync_slots();
    return fail(A3D_ERR_HIP, S->worker_error.empty() ? "a3d_server_get_particles: server stopped" : S->worker_error);
  }
  Session& se = S->sessions[session];
  if (n != se.n) return fail(A3D_ERR_ARG, "a3d_server_get_particles: n must equal the session's particle count");
  std::vector<float> soa(7 * n);
  A3D_HIP(hipMemcpyAsync(soa.data(), se.state, 7 * n * 4, hipMemcpyDeviceToHost, S->admin));
  A3D_HIP(hipStreamSynchronize(S->admin));
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 7; ++k) p7[7 * i + k] = soa[k * n + i];
  return A3D_OK;
}

int a3d_server_stats(a3d_server* S, uint64_t* requests, uint64_t* batches, uint64_t* evals, double* busy_ms) {
  if (!S) return fail(A3D_ERR_ARG, "a3d_server_stats: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  if (requests) *requests = S->st_requests;
  if (batches) *batches = S->st_batches;
  if (evals) *evals = S->st_evals;
  if (busy_ms) *busy_ms = S->st_busy_ms;
  return A3D_OK;
}

}  // extern "C"
