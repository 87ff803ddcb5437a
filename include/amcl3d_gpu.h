/*
 * amcl3d_gpu.h -- C ABI of the MI355X (gfx950) particle-filter measurement
 * update.  The library behind it is hand-written HIP; this header is plain C.
 *
 * It replaces the C++ classes of the reference's library target `amcl3d`
 * (amcl3d/CMakeLists.txt:44-52) on the path the ROS node calls
 * (amcl3d/src/Node.cpp:41,154,162,167,181,387):
 *
 *   Grid3d         amcl3d/src/Grid3d.h:34-159
 *   ParticleFilter amcl3d/src/ParticleFilter.h:67-213
 *   computeGrid    amcl3d/src/PointCloudTools.h:96 (PointCloudTools.cpp:84-149)
 *
 * Conventions
 *  - Every function returns 0 (A3D_OK) or an A3D_ERR_* code; a3d_last_error()
 *    gives the message of the last failure on the calling thread.  No C++
 *    exception crosses this boundary.
 *  - Handles own device memory.  Host arrays passed in are read during the
 *    call only.  Functions taking host arrays are synchronous.  Functions with
 *    the `_dev` suffix take device pointers (from a3d_buffer_alloc), enqueue
 *    on the handle's HIP stream and return without waiting.
 *  - A handle is not thread-safe.  A grid is read-only after it is built and
 *    may be shared by any number of filters, fleets and servers.
 *  - Random draws are inputs: standard normals z (4 per particle, order
 *    x,y,z,yaw) and one uniform u01 per resample.  A draw with standard
 *    deviation s is (float)(z * (float)s), the value normal_distribution<float>
 *    produces from z (ParticleFilter.cpp:246-256).
 *  - Particles are read and written as the reference's struct Particle
 *    {x, y, z, a, w, wp, wr}: 7 floats, 28 bytes (ParticleFilter.h:35-49).
 *    Ranges are the reference's struct Range {r, ax, ay, az}: 4 floats
 *    (ParticleFilter.h:53-63).  Clouds are three float arrays x, y, z (the
 *    coordinates of pcl::PointXYZ, in cloud order).
 */
#ifndef AMCL3D_GPU_H
#define AMCL3D_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  A3D_OK = 0,
  A3D_ERR_ARG = 1,    /* invalid argument (null pointer, size 0, bad mode) */
  A3D_ERR_HIP = 2,    /* HIP runtime failure, or no usable gfx950 device */
  A3D_ERR_STATE = 3,  /* call not valid in the handle's state */
  A3D_ERR_NOMEM = 4,  /* device allocation failed / over the HBM budget */
  A3D_ERR_IO = 5,     /* file could not be read or written */
  A3D_ERR_BUSY = 6    /* serving: no slot free for the request right now */
};

enum {
  A3D_RESAMPLE_REFERENCE = 0, /* float running sum, ParticleFilter.cpp:198-222 */
  A3D_RESAMPLE_FP64 = 1       /* FP64 prefix sum: sequential double sums in blocks
                                 of 1024 weights, block totals summed in order;
                                 u_m = (u01 + m) / N in double */
};

typedef struct a3d_grid a3d_grid;   /* Grid3d: device-resident likelihood grid */
typedef struct a3d_pf a3d_pf;       /* ParticleFilter: device-resident particles */
typedef struct a3d_fleet a3d_fleet; /* many independent filters updated per launch */
typedef struct a3d_graph a3d_graph; /* a captured chain of _dev calls */
typedef struct a3d_server a3d_server;

/* ---------------------------------------------------------------- runtime */
const char* a3d_last_error(void);
const char* a3d_version(void);
/* Selects the current HIP device for subsequent handle creation. */
int a3d_set_device(int device);
/* Name, CU count and HBM bytes of the current device. */
int a3d_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes);
int a3d_device_synchronize(void);

/* Device buffers for inputs kept resident in HBM (clouds, draws, ranges). */
int a3d_buffer_alloc(size_t bytes, void** dptr);
int a3d_buffer_free(void* dptr);
int a3d_buffer_upload(void* dptr, const void* host, size_t bytes);
int a3d_buffer_download(void* host, const void* dptr, size_t bytes);
/* Device-to-device copy (also the reference copy kernel for HBM-byte
 * counter calibration). */
int a3d_buffer_copy(void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------- Grid3d */
/* Grid3d::open minus the octomap/.grid handling (Grid3d.cpp:25-78): builds the
 * grid on the device with computeGrid (PointCloudTools.cpp:84-149) from the
 * map's occupied points xyz[3n].  Sizes are ceil((max-min)/res) per axis; the
 * reference's 250M-cell cap is lifted (64-bit cell indices).  keep_dist != 0
 * also keeps the squared-distance field (Grid3dCell::dist). */
int a3d_grid_create_from_points(const float* xyz, size_t n, const double min[3], const double max[3], double res,
                                double sensor_dev, int keep_dist, a3d_grid** out);
/* Grid3d::loadGrid without the file (Grid3d.cpp:239-275): cells given by the
 * caller, x fastest.  size[] must equal ceil((max-min)/res).  dist may be NULL. */
int a3d_grid_create_from_cells(const float* prob, const float* dist, const uint32_t size[3], const double min[3],
                               const double max[3], double res, double sensor_dev, a3d_grid** out);
/* Grid3d::loadGrid / saveGrid with the reference's .grid file layout
 * (Grid3d.cpp:210-275): u32 sx, sy, sz, f64 sensor_dev, then sx*sy*sz cells of
 * {f32 dist, f32 prob}.  Bounds and resolution are not in the file and are
 * given by the caller, as the reference takes them from the octomap.
 * Loading fails (A3D_ERR_STATE) when |file sensor_dev - sensor_dev| >= DBL_EPSILON. */
int a3d_grid_load(const char* path, const double min[3], const double max[3], double res, double sensor_dev,
                  a3d_grid** out);
int a3d_grid_save(const a3d_grid* g, const char* path);
int a3d_grid_destroy(a3d_grid* g);
int a3d_grid_get_info(const a3d_grid* g, uint32_t size[3], double min[3], double max[3], double* res,
                      double* sensor_dev);
/* Copies cells [first, first+count) of prob and (if kept and non-NULL) dist. */
int a3d_grid_read_cells(const a3d_grid* g, uint64_t first, uint64_t count, float* prob, float* dist);
/* Grid3d::isIntoMap (Grid3d.cpp:201-208). */
int a3d_grid_is_into_map(const a3d_grid* g, float x, float y, float z, int* inside);
/* Grid3d::computeCloudWeight (Grid3d.cpp:133-199) for n poses at once:
 * out[i] = computeCloudWeight(cloud, tx[i], ty[i], tz[i], roll, pitch, yaw[i]). */
int a3d_grid_cloud_weight(const a3d_grid* g, const float* x, const float* y, const float* z, size_t m,
                          const float* tx, const float* ty, const float* tz, const float* yaw, size_t n, float roll,
                          float pitch, float* out);

/* ------------------------------------------------------------ OcTree maps */
/* openOcTree + computePointCloud (PointCloudTools.cpp:26-82) on the host:
 * reads an octomap OcTree file, binary (.bt, OcTree::readBinary) or full
 * (.ot, AbstractOcTree::read, tree type OcTree), and keeps the centres of its
 * occupied leaves as floats, in the reference's leaf-iterator order (depth
 * first, children in octomap's order), with the metric bounds of all leaves
 * (getMetricMin/Max) and the resolution.  Errors (A3D_ERR_IO) carry the
 * reference's messages: "Cannot find file <path>", "OcTree cannot be read"
 * (.bt), "OcTree cannot be created from file <path>" (.ot, other extensions).
 * Truncated data and nodes below depth 16 are errors. */
typedef struct a3d_octree a3d_octree;
int a3d_octree_read(const char* path, a3d_octree** out);
/* nodes is OcTree::size(); occupied the number of occupied leaf centres. */
int a3d_octree_get_info(const a3d_octree* t, double min[3], double max[3], double* res, uint64_t* nodes,
                        uint64_t* occupied);
/* Centres [first, first+count) as x, y, z floats (host memory). */
int a3d_octree_get_points(const a3d_octree* t, uint64_t first, uint64_t count, float* xyz);
int a3d_octree_destroy(a3d_octree* t);
/* Grid3d::open(map_path, sensor_dev) (Grid3d.cpp:25-78): the OcTree file
 * (A3D_ERR_IO "OcTree is empty" when it has at most one node), then the .grid
 * cache named like the reference's (the map path up to the first ".bt" /
 * ".ot" in it, plus ".grid") when it loads with this sensor_dev and the map's
 * sizes, else computeGrid on the device and the cache written (a failed write
 * is ignored, as in the reference).  keep_dist as in
 * a3d_grid_create_from_points. */
int a3d_grid_open(const char* map_path, double sensor_dev, int keep_dist, a3d_grid** out);

/* ------------------------------------------------------------ VoxelGrid */
/* pcl::VoxelGrid<pcl::PointXYZ> as the node applies it to each scan before
 * update (Node.cpp:130-141: setLeafSize(voxel_size) on all axes, PCL
 * defaults otherwise): one point per occupied voxel of side `leaf`, at the
 * float centroid of the voxel's finite points, in increasing voxel index
 * (the order PCL outputs).  points: n points `stride` bytes apart with x, y,
 * z as their first three floats (12 for packed xyz, 16 for pcl::PointXYZ).
 * Output: out_x/out_y/out_z with capacity n, *out_n points.  The points of a
 * voxel are summed in point order; PCL sums them in the order its std::sort
 * leaves, so a centroid of three or more points may differ in the last bits. */
int a3d_voxel_grid(const void* points, size_t n, size_t stride, float leaf, float* out_x, float* out_y, float* out_z,
                   size_t* out_n);
/* The same on device buffers (synchronous): d_points in, d_x/d_y/d_z out. */
int a3d_voxel_grid_dev(const void* d_points, size_t n, size_t stride, float leaf, float* d_x, float* d_y, float* d_z,
                       size_t* out_n);

/* ----------------------------------------------------------- ParticleFilter */
int a3d_pf_create(a3d_pf** out);
int a3d_pf_destroy(a3d_pf* pf);
/* ParticleFilter::init (ParticleFilter.cpp:46-95).  z4: 4*num_particles
 * standard normals; the first four are unused (particle 0 is the pose). */
int a3d_pf_init(a3d_pf* pf, int num_particles, float x_init, float y_init, float z_init, float a_init, float x_dev,
                float y_dev, float z_dev, float a_dev, const float* z4);
/* ParticleFilter::predict (ParticleFilter.cpp:97-119).  z4: 4*N normals. */
int a3d_pf_predict(a3d_pf* pf, double odom_x_mod, double odom_y_mod, double odom_z_mod, double odom_a_mod,
                   double delta_x, double delta_y, double delta_z, double delta_a, const float* z4);
/* ParticleFilter::update (ParticleFilter.cpp:121-196).  ranges: nb x {r,ax,ay,az}. */
int a3d_pf_update(a3d_pf* pf, const a3d_grid* g, const float* x, const float* y, const float* z, size_t m,
                  const float* ranges, size_t nb, double alpha, double sigma, double roll, double pitch);
/* ParticleFilter::resample (ParticleFilter.cpp:198-222).  idx_out (optional):
 * N ancestor indices. */
int a3d_pf_resample(a3d_pf* pf, float u01, int mode, uint32_t* idx_out);
/* ParticleFilter::getMean (ParticleFilter.h:94-97): {x, y, z, a}. */
int a3d_pf_get_mean(const a3d_pf* pf, float mean[4]);
int a3d_pf_is_initialized(const a3d_pf* pf, int* initialized);
int a3d_pf_size(const a3d_pf* pf, size_t* n);
/* Particle state as struct Particle[n] (replaces buildParticlesPoseMsg). */
int a3d_pf_get_particles(const a3d_pf* pf, float* p7, size_t n);
int a3d_pf_set_particles(a3d_pf* pf, const float* p7, size_t n);

/* Asynchronous variants on device buffers, for inputs resident in HBM and for
 * graph capture.  d_x, d_y, d_z: cloud arrays; d_ranges: nb x 4 floats. */
int a3d_pf_predict_dev(a3d_pf* pf, const double mods[4], const double deltas[4], const float* d_z4);
int a3d_pf_update_dev(a3d_pf* pf, const a3d_grid* g, const float* d_x, const float* d_y, const float* d_z,
                      size_t m, const float* d_ranges, size_t nb, double alpha, double sigma, double roll,
                      double pitch);
int a3d_pf_resample_dev(a3d_pf* pf, float u01, int mode);
/* No counterpart in the reference: a3d_pf_update_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  It allocates nothing, not even on its
 * first call: the particles are the caller's, seven device arrays of n floats
 * in one block (x, y, z, a, w, wp, wr; field k at d_particles + k*n), updated
 * in place as ParticleFilter::update (ParticleFilter.cpp:121-196) does, and
 * d_mean (4 floats) receives the mean.  Its working memory is a fixed static
 * device array for at most 131072 particles (about 8 MB per device, held
 * until the process ends); larger n fails with A3D_ERR_ARG.  Enqueued on
 * `stream` (a hipStream_t; NULL: the null stream) without waiting.  Every
 * call on one device shares that array, so calls that overlap on a device
 * race: serialise them (one stream, or wait between calls). */
int a3d_pf_update_noalloc(const a3d_grid* g, float* d_particles, size_t n, const float* d_x, const float* d_y,
                          const float* d_z, size_t m, const float* d_ranges, size_t nb, double alpha, double sigma,
                          double roll, double pitch, float* d_mean, void* stream);
int a3d_pf_synchronize(a3d_pf* pf);
/* Captures the _dev calls made between begin and end into a HIP graph. */
int a3d_pf_capture_begin(a3d_pf* pf);
int a3d_pf_capture_end(a3d_pf* pf, a3d_graph** out);
int a3d_graph_launch(a3d_graph* graph);
int a3d_graph_destroy(a3d_graph* graph);
/* Average duration (ms) of the cloud-weight kernel over the _dev updates made
 * since the last reset, measured with HIP events on the handle's stream.
 * Not available inside a capture. */
int a3d_pf_timing(a3d_pf* pf, int reset, double* weight_ms_avg, uint64_t* launches, uint64_t* evals);

/* Shape of the cloud-weight kernel; every shape gives identical results.
 * Mapping A3D_MAP_PARTICLE_LANES: a wavefront carries `particles_per_wave`
 * particles (64, 32 or 16), each lane one particle and 64/particles_per_wave
 * consecutive points per step.  A3D_MAP_POINT_LANES: lanes are 64
 * consecutive points, a wavefront walks the cloud for `particles_per_wave`
 * particles through an LDS tile. */
enum { A3D_MAP_PARTICLE_LANES = 0, A3D_MAP_POINT_LANES = 1 };
int a3d_set_tile(int particles_per_wave);
int a3d_get_tile(int* particles_per_wave);
/* Points of the cloud staged per LDS tile by the A3D_MAP_PARTICLE_LANES
 * kernels (32, 64, 128 or 256; default 64): the workgroup's wavefronts walk
 * the cloud one tile at a time, one barrier per tile.  Point-lane kernels
 * always stage 64 (one point per lane). */
int a3d_set_cloud_tile(int points);
int a3d_get_cloud_tile(int* points);
/* How persistent launches (fleets, served batches) deal their work list of
 * particle blocks to the chip's 8 XCDs (accelerator dies, each with its own
 * L2): 0 (default), one contiguous range per XCD, so an L2 serves few filters
 * at a time; run > 0, runs of `run` consecutive blocks dealt round-robin, so
 * the whole chip works on few filters at a time and their cells stay in the
 * memory-side cache.  Results are identical either way. */
int a3d_set_deal(int run);
int a3d_get_deal(int* run);
int a3d_set_mapping(int mapping);
int a3d_get_mapping(int* mapping);
/* Order in which the cloud-weight kernel visits a filter's particles (the
 * particles themselves stay in index order; results are identical):
 * A3D_ORDER_INDEX, or A3D_ORDER_SPATIAL, a counting sort by a Morton code of
 * (x, y, z, yaw) over the particles' bounding box, computed on the device
 * before each single-filter update. */
enum { A3D_ORDER_INDEX = 0, A3D_ORDER_SPATIAL = 1 };
int a3d_set_order(int order);
int a3d_get_order(int* order);

/* -------------------------------------------------------------------- Fleet */
/* Many independent filters on one grid, updated by one persistent launch.
 * Per filter f: n[f] particles; all arrays below are concatenations in filter
 * order, with offsets given by prefix sums of the counts. */
int a3d_fleet_create(size_t nfilters, const uint32_t* num_particles, a3d_fleet** out);
int a3d_fleet_destroy(a3d_fleet* fl);
int a3d_fleet_set_particles(a3d_fleet* fl, const float* p7);
int a3d_fleet_get_particles(const a3d_fleet* fl, float* p7);
/* init for every filter: poses/devs: 4 floats per filter. */
int a3d_fleet_init(a3d_fleet* fl, const float* poses, const float* devs, const float* z4);
/* mods/deltas: 4 doubles per filter. */
int a3d_fleet_predict(a3d_fleet* fl, const double* mods, const double* deltas, const float* z4);
/* Clouds: m[f] points each, concatenated; ranges: nb[f] each, concatenated.
 * alpha, sigma, roll, pitch: one per filter. */
int a3d_fleet_update(a3d_fleet* fl, const a3d_grid* g, const float* x, const float* y, const float* z,
                     const uint32_t* m, const float* ranges, const uint32_t* nb, const double* alpha,
                     const double* sigma, const double* roll, const double* pitch);
int a3d_fleet_resample(a3d_fleet* fl, const float* u01, int mode);
int a3d_fleet_get_means(const a3d_fleet* fl, float* means4);
/* Same update with clouds, ranges and parameters already on the device
 * (uploaded once with a3d_fleet_stage_update); runs asynchronously.  A
 * staged update stays valid across updates and resamples; a3d_fleet_init,
 * a3d_fleet_predict and a3d_fleet_set_layout replace the descriptors it
 * staged, so stage again after them (a3d_fleet_update_staged refuses until
 * then). */
int a3d_fleet_stage_update(a3d_fleet* fl, const float* x, const float* y, const float* z, const uint32_t* m,
                           const float* ranges, const uint32_t* nb, const double* alpha, const double* sigma,
                           const double* roll, const double* pitch);
int a3d_fleet_update_staged(a3d_fleet* fl, const a3d_grid* g, int persistent);
int a3d_fleet_synchronize(a3d_fleet* fl);
/* Cloud-weight kernel timing for the fleet (as a3d_pf_timing). */
int a3d_fleet_timing(a3d_fleet* fl, int reset, double* weight_ms_avg, uint64_t* launches, uint64_t* evals);

/* Device-parameter forms: every per-step input is a device array, so a step
 * only enqueues kernels on the fleet's stream (no host synchronisation) and
 * one predict -> update -> resample of all filters can be captured in a HIP
 * graph and replayed; the graph reads the arrays' contents at each replay.
 * Each is ParticleFilter::predict / update / resample (ParticleFilter.cpp:
 * 97-222) for every filter, as a3d_fleet_predict / _update / _resample.
 *
 * a3d_fleet_set_layout fixes the cloud size m[f] and beacon count nb[f] of
 * every filter for the _dev calls and cuts the work list for the current
 * kernel shape (host call, synchronises).  a3d_fleet_init and
 * a3d_fleet_predict (host forms) and a3d_fleet_stage_update clear it.
 *   predict_dev: d_mods, d_deltas 4 doubles per filter; d_z4 4 normals per
 *                particle, filters concatenated.
 *   update_dev:  d_x, d_y, d_z the filters' clouds concatenated in filter
 *                order (m[f] points each); d_ranges nb[f] x {r, ax, ay, az}
 *                each, concatenated; d_params {alpha, sigma, roll, pitch}
 *                per filter (doubles; roll and pitch are narrowed to float
 *                as computeCloudWeight's parameters).
 *   resample_dev: d_u01 one uniform per filter. */
int a3d_fleet_set_layout(a3d_fleet* fl, const uint32_t* m, const uint32_t* nb);
int a3d_fleet_predict_dev(a3d_fleet* fl, const double* d_mods, const double* d_deltas, const float* d_z4);
int a3d_fleet_update_dev(a3d_fleet* fl, const a3d_grid* g, const float* d_x, const float* d_y, const float* d_z,
                         const float* d_ranges, const double* d_params);
int a3d_fleet_resample_dev(a3d_fleet* fl, const float* d_u01, int mode);
/* Capture of _dev calls on the fleet's stream into a graph (a3d_graph_launch
 * replays it; a3d_graph_destroy frees it).  Host forms are refused while
 * capturing. */
int a3d_fleet_capture_begin(a3d_fleet* fl);
int a3d_fleet_capture_end(a3d_fleet* fl, a3d_graph** out);

/* ------------------------------------------------------------------ Serving */
/* A request loop that keeps every session's particles, weights and RNG state
 * on the device, admits sessions by an HBM budget and batches updates as they
 * arrive (continuous batching).  Draws come from a counter-based generator
 * (Philox4x32-10) keyed by the session seed, so a session's results do not
 * depend on which other sessions share its batches. */
typedef struct {
  size_t hbm_budget_bytes;    /* session state + staging admitted at most */
  uint64_t max_batch_evals;   /* particle.point evaluations per launch; under a
                                 backlog the bound grows to the queued
                                 evaluations over the free launch slots, up to
                                 8x this, so an overloaded server runs fewer,
                                 larger launches */
  int max_batch_sessions;     /* sessions per launch */
  int max_inflight;           /* launches in flight (streams) */
  int latest_only;            /* != 0: a session keeps at most one queued request; a
                                 newer one supersedes it (the reference's scan
                                 subscriber has queue size 1, Node.cpp:60), and
                                 the superseded odometry increment is composed
                                 into the newer one */
} a3d_server_config;

typedef struct {
  uint64_t ticket;
  int session;
  float mean[4];
  double latency_ms;    /* submit -> result on the host */
  double submit_ms;     /* on the server clock */
  uint64_t evals;       /* 0 for a superseded request */
  double deltas[4];     /* odometry increment the update applied (its own, composed
                           after those of the requests it superseded) */
  int superseded;       /* 1: replaced by a newer request of the session before
                           it ran; mean is not set */
} a3d_result;

int a3d_server_create(const a3d_server_config* cfg, a3d_server** out);
int a3d_server_destroy(a3d_server* s);
/* Adds a grid the sessions may use; the server does not own it. */
int a3d_server_add_map(a3d_server* s, const a3d_grid* g, int* map_id);
/* Opens a session (init on the device from the seed).  A3D_ERR_NOMEM when the
 * HBM budget is exhausted. */
int a3d_server_open_session(a3d_server* s, int map_id, int num_particles, const float pose[4], const float devs[4],
                            uint64_t seed, int* session);
/* Opens a session from given particles (struct Particle[n]), e.g. a global
 * localisation set drawn uniformly over the map's free space; the session's
 * draws then start at step 1, as after a3d_server_open_session. */
int a3d_server_open_session_particles(a3d_server* s, int map_id, const float* p7, size_t num_particles,
                                      uint64_t seed, int* session);
int a3d_server_close_session(a3d_server* s, int session);
/* Queues predict + update (+ resample if resample != 0) for a session.  The
 * cloud and ranges are copied at submit.  deltas is the odometry increment
 * since the session's previous request: (x, y, z) in the frame of the
 * previous pose and the yaw change.  When a request supersedes a queued one
 * (latest_only) the increments compose as poses do: x = a.x + cos(a.yaw)*b.x
 * - sin(a.yaw)*b.y, y = a.y + sin(a.yaw)*b.x + cos(a.yaw)*b.y, z = a.z + b.z,
 * yaw = a.yaw + b.yaw. */
int a3d_server_submit(a3d_server* s, int session, const double mods[4], const double deltas[4], const float* x,
                      const float* y, const float* z, size_t m, const float* ranges, size_t nb, double alpha,
                      double sigma, double roll, double pitch, int resample, uint64_t* ticket);
/* Completed results, oldest first. */
int a3d_server_poll(a3d_server* s, a3d_result* out, size_t max, size_t* n);
/* Waits until every submitted request has completed. */
int a3d_server_drain(a3d_server* s);
int a3d_server_get_particles(a3d_server* s, int session, float* p7, size_t n);
/* Counters: requests done, batches launched, evals done, mean batch size. */
int a3d_server_stats(a3d_server* s, uint64_t* requests, uint64_t* batches, uint64_t* evals, double* busy_ms);
/* ------------------------------------------------------------------ Node */
/* The ROS node's glue around the filter (amcl3d/src/Node.cpp), without ROS:
 * the caller delivers the node's three inputs (odometry transforms, raw
 * point clouds with the time they arrive, radio ranges) and reads back what
 * the node publishes.  Parameters as amcl3d/src/Parameters.h:33-73. */
typedef struct a3d_node a3d_node;
typedef struct {
  int num_particles;
  int resample_interval;
  int set_initial_pose;
  double init_x, init_y, init_z, init_a;
  double init_x_dev, init_y_dev, init_z_dev, init_a_dev;
  double sensor_range; /* range sigma of update */
  double voxel_size;
  double odom_x_mod, odom_y_mod, odom_z_mod, odom_a_mod;
  double update_rate; /* Hz */
  double d_th, a_th;
  double take_off_height;
  double alpha;
} a3d_node_params;
/* What odomCallback publishes (Node.cpp:306-321): the base pose in the world
 * (lastbase_2_world) and the odom frame in the world (lastodom_2_world), as
 * translation + quaternion (x, y, z, w); published = 0 when the filter was
 * not initialised yet and nothing was published. */
typedef struct {
  int published;
  int amcl_out;
  int has_takenoff;
  double base_t[3], base_q[4];
  double odom_t[3], odom_q[4];
} a3d_node_output;
/* The node's filter draws from the counter-based generator of served sessions
 * keyed by `seed`: step 0 for init, then one step per predict (its u01 drives
 * that update's resample); a3d_session_draws replays them. */
int a3d_node_create(const a3d_node_params* p, const a3d_grid* grid, uint64_t seed, a3d_node** out);
int a3d_node_destroy(a3d_node* nd);
/* Node::odomCallback (Node.cpp:193-324): q = (x, y, z, w). */
int a3d_node_odometry(a3d_node* nd, const double t[3], const double q[4], a3d_node_output* out);
/* Node::rangeCallback (Node.cpp:326-345). */
int a3d_node_range(a3d_node* nd, float range, double ax, double ay, double az);
/* Node::pointcloudCallback (Node.cpp:112-191) with checkUpdateThresholds
 * (Node.cpp:347-373): now_ns stands for ros::Time::now(); points as in
 * a3d_voxel_grid.  *updated = 1 when predict -> update (-> resample) ran. */
int a3d_node_point_cloud(a3d_node* nd, int64_t now_ns, const void* points, size_t n, size_t stride, int* updated);
/* Node::setInitialPose (Node.cpp:375-398). */
int a3d_node_set_initial_pose(a3d_node* nd, const double t[3], const double q[4], float x_dev, float y_dev,
                              float z_dev, float a_dev);
/* mean_p_, the odometry deltas of the last predict, roll_/pitch_, is_odom_,
 * amcl_out_ and the next draw step. */
int a3d_node_get_state(const a3d_node* nd, float mean[4], double deltas[4], double* roll, double* pitch, int* is_odom,
                       int* amcl_out, uint64_t* draw_step);
/* The node's filter (for its particles). */
int a3d_node_get_filter(a3d_node* nd, a3d_pf** pf);

/* The draws a session's next `count` steps would use (for tests): z4 for
 * predict (4*N per step) and u01 per step, starting at step `first`. */
int a3d_session_draws(uint64_t seed, uint64_t step, int num_particles, float* z4, float* u01);

#ifdef __cplusplus
}
#endif
#endif /* AMCL3D_GPU_H */
