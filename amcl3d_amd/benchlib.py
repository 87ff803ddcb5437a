"""Workloads of the batched and served forms, for bench.py.

batched: 256 independent filters on the room map (200 x 200 x 60 cells at
0.1 m); per filter N log-uniform in 500..50,000 particles, M uniform in
1,000..30,000 points, 1..8 beacons, all seeded.  One persistent launch per
update versus the plain single-filter kernel launched once per filter.

served: sessions on resident maps, each sending Poisson update requests at
10 Hz; per session N log-uniform in 1,000..50,000, M uniform in
2,000..20,000.  p50/p99 latency is request submit -> mean on the host.
"""
import heapq
import time

import numpy as np

from . import autotune, fleet as fleet_mod, serving


def large_filter_workload(amcl3d_amd, W, steps_total, seed, n, m, nb, log=print):
    """The single-large-filter workload: a seeded building, 100 x 100 x 30 m at
    0.05 m (2000 x 2000 x 600 cells) built on the device; a smooth trajectory;
    per step an m-point cloud (noise 0.03 m), nb beacons, n-particle draws."""
    t0 = time.time()
    bm = W.building_map(seed)
    pts = bm.points(dedup=False)
    log("map: %d boxes, %d occupied points (%.1fs)" % (bm.boxes.shape[0], pts.shape[0], time.time() - t0))
    grid = amcl3d_amd.Grid3d()
    amcl3d_amd.synchronize()
    t1 = time.time()
    grid.open(pts, bm.min, bm.max, bm.res, 0.05)
    amcl3d_amd.synchronize()
    build_s = time.time() - t1
    log("computeGrid on the device: %d cells in %.2fs" % (grid.cells, build_s))
    del pts
    traj = W.trajectory(bm, steps_total, seed, z=1.5)
    view = W.storey_view(bm, 1.5, 3.0)
    rng = np.random.default_rng(seed)
    anchors = rng.uniform(bm.min + 1, [bm.max[0] - 1, bm.max[1] - 1, 3.0], (nb, 3))
    clouds, ranges = [], []
    t2 = time.time()
    for k in range(steps_total):
        pose = traj[k + 1]
        clouds.append(W.sensor_cloud(view, pose, m, 0.03, rng))
        r = np.linalg.norm(anchors - pose[:3], axis=1) + rng.normal(0, 0.1, nb)
        ranges.append(np.concatenate([r[:, None], anchors], 1).astype(np.float32))
    log("clouds: %d x %d points (%.1fs)" % (steps_total, m, time.time() - t2))
    parts = W.particles_around(traj[0], n, rng)
    deltas = [W.odom_delta(traj[k], traj[k + 1]) for k in range(steps_total)]
    z4 = [rng.standard_normal((n, 4)).astype(np.float32) for _ in range(steps_total)]
    u01 = [float(rng.uniform()) for _ in range(steps_total)]
    return dict(bm=bm, grid=grid, build_s=build_s, traj=traj, clouds=clouds, ranges=ranges, parts=parts,
                deltas=deltas, z4=z4, u01=u01, view=view)


def room_grid(amcl3d_amd, W, seed=1):
    bm = W.room_map(seed)
    g = amcl3d_amd.Grid3d()
    g.open(bm.points(), bm.min, bm.max, bm.res, 0.1)
    return bm, g


def fleet_case(W, bm, nfilters, seed):
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(np.log(500), np.log(50000), nfilters))).astype(np.uint32)
    ms = rng.integers(1000, 30001, nfilters)
    nbs = rng.integers(1, 9, nfilters)
    parts, clouds, ranges = [], [], []
    for f in range(nfilters):
        pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), rng.uniform(-np.pi, np.pi))
        clouds.append(W.sensor_cloud(bm, pose, int(ms[f]), 0.05, rng))
        ranges.append(W.beacons(bm, pose, int(nbs[f]), 0.1, rng))
        parts.append(W.particles_around(pose, int(counts[f]), rng))
    return counts, parts, clouds, ranges


def batched(amcl3d_amd, W, bm, grid, nfilters=256, seed=3, reps=3):
    counts, parts, clouds, ranges = fleet_case(W, bm, nfilters, seed)
    fl = fleet_mod.Fleet(counts)
    fl.set_particles(np.concatenate(parts))

    def prepare(_shape):
        fl.stage_update(clouds, ranges, 0.8, 0.1)

    def run_persistent():
        fl.update_staged(grid, fleet_mod.PERSISTENT)
        fl.synchronize()

    shape, timings = autotune.tune(run_persistent, prepare=prepare)
    fl.timing(reset=True)
    t_batched = min(_timed(run_persistent) for _ in range(reps))
    kernel_ms, _, _ = fl.timing(reset=True)

    def run_loop():
        fl.update_staged(grid, fleet_mod.PLAIN_LOOP)
        fl.synchronize()

    run_loop()
    t_loop = min(_timed(run_loop) for _ in range(reps))
    ev = fl.evals
    fl.close()
    return {"workload": "%d filters on the 200x200x60 room map, N log-uniform 500..50k, M 1k..30k, 1..8 beacons"
                        % nfilters,
            "evals_per_update": ev, "shape": autotune.shape_name(shape),
            "autotune_s": {autotune.shape_name(k): v for k, v in timings.items()},
            "batched_evals_per_s": ev / t_batched, "plain_loop_evals_per_s": ev / t_loop,
            "batched_over_plain_loop": t_loop / t_batched, "batched_weight_kernel_ms": kernel_ms,
            "batched_update_ms": t_batched * 1e3, "plain_loop_update_ms": t_loop * 1e3}


def _timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def served(amcl3d_amd, W, maps, nsessions=256, rate_hz=10.0, duration_s=8.0, seed=5, max_inflight=3):
    """maps: list of (BoxMap view for ray casting, Grid3d, z of the sensor)."""
    rng = np.random.default_rng(seed)
    srv = serving.Server(max_inflight=max_inflight, max_batch_sessions=512, max_batch_evals=6_000_000_000)
    mids = [srv.add_map(g) for _, g, _ in maps]
    sess = []
    t0 = time.time()
    for s in range(nsessions):
        k = s % len(maps)
        view, grid, z = maps[k]
        n = int(np.round(np.exp(rng.uniform(np.log(1000), np.log(50000)))))
        m = int(rng.integers(2000, 20001))
        pose = np.append(W.free_pose(view, rng, z_range=(z, z)), rng.uniform(-np.pi, np.pi)).astype(np.float32)
        cloud = W.sensor_cloud(view, pose, m, 0.05, rng)
        xyz = tuple(np.ascontiguousarray(cloud[:, j]) for j in range(3))
        rg = W.beacons(view, pose, 3, 0.1, rng)
        sid = srv.open_session(mids[k], n, pose, [0.5, 0.5, 0.1, 0.2], seed * 100003 + s)
        sess.append((sid, n, m, xyz, rg))
    setup_s = time.time() - t0
    mods = np.array([0.1, 0.1, 0.1, 0.3])
    deltas = np.array([0.05, 0.0, 0.0, 0.01])
    # Poisson arrivals per session
    heap = [(rng.exponential(1.0 / rate_hz), i) for i in range(nsessions)]
    heapq.heapify(heap)
    lat, offered = [], 0
    start = time.perf_counter()
    while heap:
        t_next, i = heap[0]
        if t_next > duration_s:
            break
        now = time.perf_counter() - start
        if now < t_next:
            for r in srv.poll():
                lat.append(r.latency_ms)
            time.sleep(min(t_next - now, 0.0005))
            continue
        heapq.heappop(heap)
        sid, n, m, xyz, rg = sess[i]
        srv.submit(sid, mods, deltas, xyz, rg, 0.8, 0.1)
        offered += n * m
        heapq.heappush(heap, (t_next + rng.exponential(1.0 / rate_hz), i))
    srv.drain()
    wall = time.perf_counter() - start
    for r in srv.poll(1 << 20):
        lat.append(r.latency_ms)
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    return {"workload": "%d sessions on %d resident maps, Poisson %.0f Hz each, N log-uniform 1k..50k, "
                        "M 2k..20k, %.0f s" % (nsessions, len(maps), rate_hz, duration_s),
            "requests": int(lat.size), "batches": st["batches"],
            "offered_evals_per_s": offered / duration_s, "served_evals_per_s": st["evals"] / wall,
            "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
            "max_ms": float(lat.max()) if lat.size else None, "session_setup_s": setup_s}


def extras(amcl3d_amd, W, plain_value=None, building=None):
    bm, grid = room_grid(amcl3d_amd, W)
    out = {"batched": batched(amcl3d_amd, W, bm, grid)}
    maps = [(bm, grid, 1.5)]
    if building is not None:
        maps.append(building)
    out["served"] = served(amcl3d_amd, W, maps)
    if plain_value:
        out["served_over_plain_large_filter"] = out["served"]["served_evals_per_s"] / plain_value
    out["served_over_plain_loop"] = out["served"]["served_evals_per_s"] / out["batched"]["plain_loop_evals_per_s"]
    return out
