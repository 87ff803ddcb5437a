"""Workloads of the batched and served forms, for bench.py.

batched: 256 independent filters on the room map (200 x 200 x 60 cells at
0.1 m); per filter N log-uniform in 500..50,000 particles, M uniform in
1,000..30,000 points, 1..8 beacons, all seeded.  One persistent launch per
update versus the plain single-filter kernel launched once per filter.  The
same fleet on the large filter's building map (2000 x 2000 x 600 cells, where
spread particle sets make the gathers HBM-bound) is reported beside it.

map_file: the building map stored as an OcTree .bt file (one occupied leaf
per occupied cell) and read back the way Grid3d::open reads it.

node_step: the node's point-cloud callback (odometry increment, VoxelGrid,
predict, update, resample) on the large filter's map, trajectory and clouds.

served: sessions on resident maps, each sending Poisson update requests at
10 Hz; per session N log-uniform in 1,000..50,000, M uniform in
2,000..20,000.  p50/p99 latency is request submit -> mean on the host.
"""
import heapq
import time

import numpy as np

from . import autotune, fleet as fleet_mod, serving


# Trajectories are generated this long (and cut to the steps asked for), so
# that step k of every workload of a seed is the same whatever the length.
TRAJECTORY_STEPS = 256


def large_filter_workload(amcl3d_amd, W, steps_total, seed, n, m, nb, log=print):
    """The single-large-filter workload: a seeded building, 100 x 100 x 30 m at
    0.05 m (2000 x 2000 x 600 cells) built on the device; a smooth trajectory;
    per step an m-point cloud (noise 0.03 m), nb beacons, n-particle draws.

    Every input of step k (pose, cloud, ranges, draws) comes from its own
    generator, seeded by (seed, k), and the particles from (seed,), so step k
    is the same in a workload of any length (a profile of 105 updates covers
    any bench run of warmup + steps <= 105 update for update)."""
    if steps_total > TRAJECTORY_STEPS:
        raise ValueError("at most %d steps" % TRAJECTORY_STEPS)
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
    traj = W.trajectory(bm, TRAJECTORY_STEPS, seed, z=1.5)[:steps_total + 1]
    view = W.storey_view(bm, 1.5, 3.0)
    rng = np.random.default_rng([seed, 0])
    anchors = rng.uniform(bm.min + 1, [bm.max[0] - 1, bm.max[1] - 1, 3.0], (nb, 3))
    parts = W.particles_around(traj[0], n, rng)
    clouds, ranges, z4, u01 = [], [], [], []
    t2 = time.time()
    for k in range(steps_total):
        rk = np.random.default_rng([seed, 1, k])
        pose = traj[k + 1]
        clouds.append(W.sensor_cloud(view, pose, m, 0.03, rk))
        r = np.linalg.norm(anchors - pose[:3], axis=1) + rk.normal(0, 0.1, nb)
        ranges.append(np.concatenate([r[:, None], anchors], 1).astype(np.float32))
        z4.append(rk.standard_normal((n, 4)).astype(np.float32))
        u01.append(float(rk.uniform()))
    log("clouds: %d x %d points (%.1fs)" % (steps_total, m, time.time() - t2))
    deltas = [W.odom_delta(traj[k], traj[k + 1]) for k in range(steps_total)]
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


def batched(amcl3d_amd, W, bm, grid, nfilters=256, seed=3, reps=3, map_name="the building map", converge=0):
    """The fleet's first update (every filter's particles spread around its
    pose) in one persistent launch and in the plain loop.  converge > 0: then
    that many update -> resample -> predict cycles (odometry 2 cm and 0.5 deg
    a cycle with the launch file's noise factors), after which the filters
    track their poses, and the same two timings again ("converged")."""
    entry_shape = autotune.get_shape()  # restored on return: later phases run at the shape they found
    counts, parts, clouds, ranges = fleet_case(W, bm, nfilters, seed)
    fl = fleet_mod.Fleet(counts)
    fl.set_particles(np.concatenate(parts))

    def prepare(_shape):
        fl.stage_update(clouds, ranges, 0.8, 0.1)

    def run_persistent():
        fl.update_staged(grid, fleet_mod.PERSISTENT)
        fl.synchronize()

    shape, timings = autotune.tune(run_persistent, candidates=autotune.FLEET_CANDIDATES, prepare=prepare)
    fl.timing(reset=True)
    t_batched = min(_timed(run_persistent) for _ in range(reps))
    kernel_ms, _, _ = fl.timing(reset=True)

    def run_loop():
        fl.update_staged(grid, fleet_mod.PLAIN_LOOP)
        fl.synchronize()

    # the plain loop at its own best shape, not the fleet's
    loop_shape, loop_timings = autotune.tune(run_loop, candidates=autotune.LOOP_CANDIDATES, prepare=prepare)
    t_loop = min(_timed(run_loop) for _ in range(reps))
    autotune.set_shape(shape)
    prepare(shape)
    ev = fl.evals
    conv = None
    if converge > 0:
        rng = np.random.default_rng(seed + 1)
        F = len(counts)
        mods = np.tile([0.1, 0.1, 0.1, 0.3], (F, 1))
        deltas = np.tile([0.02, 0.0, 0.0, 0.0087], (F, 1))
        for _ in range(converge):
            prepare(shape)  # predict replaced the staged descriptors
            run_persistent()
            fl.resample(rng.uniform(0, 1, F).astype(np.float32))
            fl.predict(mods, deltas, rng.standard_normal((fl.total, 4)).astype(np.float32))
        prepare(shape)
        fl.synchronize()
        tb = min(_timed(run_persistent) for _ in range(reps))
        autotune.set_shape(loop_shape)
        prepare(loop_shape)
        tl = min(_timed(run_loop) for _ in range(reps))
        autotune.set_shape(shape)
        conv = {"cycles": converge, "batched_evals_per_s": ev / tb, "plain_loop_evals_per_s": ev / tl,
                "batched_over_plain_loop": tl / tb, "batched_update_ms": tb * 1e3, "plain_loop_update_ms": tl * 1e3}
    fl.close()
    autotune.set_shape(entry_shape)
    out = {"workload": "%d filters on %s, N log-uniform 500..50k, M 1k..30k, 1..8 beacons" % (nfilters, map_name),
            "evals_per_update": ev, "shape": autotune.shape_name(shape),
            "autotune_s": {autotune.shape_name(k): v for k, v in timings.items()},
            "plain_loop_shape": autotune.shape_name(loop_shape),
            "plain_loop_autotune_s": {autotune.shape_name(k): v for k, v in loop_timings.items()},
            "batched_evals_per_s": ev / t_batched, "plain_loop_evals_per_s": ev / t_loop,
            "batched_over_plain_loop": t_loop / t_batched, "batched_weight_kernel_ms": kernel_ms,
            "batched_update_ms": t_batched * 1e3, "plain_loop_update_ms": t_loop * 1e3}
    if conv:
        out["converged"] = conv
    return out


def _timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def served(amcl3d_amd, W, maps, nsessions=256, rate_hz=10.0, duration_s=8.0, seed=5, max_inflight=3,
           max_batch_evals=800_000_000, latest_only=True, poses_per_map=12, n_range=(1000, 50000),
           m_range=(2000, 20000), slo_ms=None, ladder=(1.3, 1.2, 1.1, 1.0, 0.9, 0.8, 0.7, 0.6, 0.45, 0.3, 0.15),
           slo_duration_s=4.0):
    """Sessions sending Poisson update requests at rate_hz each.

    maps: list of (BoxMap view for ray casting, Grid3d, sensor height z).
    Session s uses map s mod len(maps); N is log-uniform in n_range, M uniform
    in m_range.  Each map has poses_per_map seeded sensor poses with their
    clouds (M points ray-cast); a session takes one of them.  Latency is
    submit -> mean on the host, for the requests that ran; with latest_only a
    request still queued when its session's next one arrives is superseded.

    With slo_ms, the same sessions then run at rates around the one the
    device kept up with at rate_hz (its served / offered load), highest
    first, until one keeps p99 under slo_ms with at most 1% of its requests
    superseded; the result gains "slo" (that rate's figures) and "ladder"
    (every rate tried)."""
    rng = np.random.default_rng(seed)
    srv = serving.Server(max_inflight=max_inflight, max_batch_sessions=4096, max_batch_evals=max_batch_evals,
                         latest_only=latest_only)
    mids = [srv.add_map(g) for _, g, _ in maps]
    t0 = time.time()
    views = []
    for view, grid, z in maps:
        poses = []
        for _ in range(poses_per_map):
            pose = np.append(W.free_pose(view, rng, z_range=(z, z)), rng.uniform(-np.pi, np.pi)).astype(np.float32)
            cloud = W.sensor_cloud(view, pose, int(m_range[1]), 0.05, rng)
            poses.append((pose, cloud, W.beacons(view, pose, 3, 0.1, rng)))
        views.append(poses)
    sess = []
    for s in range(nsessions):
        k = s % len(maps)
        n = int(np.round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        m = int(rng.integers(m_range[0], m_range[1] + 1))
        pose, cloud, rg = views[k][int(rng.integers(0, poses_per_map))]
        xyz = tuple(np.ascontiguousarray(cloud[:m, j]) for j in range(3))
        sid = srv.open_session(mids[k], n, pose, [0.5, 0.5, 0.1, 0.2], seed * 100003 + s)
        sess.append((sid, n, m, xyz, rg))
    setup_s = time.time() - t0
    try:
        out = _serve_phase(srv, sess, rate_hz, duration_s, rng)
        out["workload"] = ("%d sessions on %d resident maps, Poisson %.1f Hz each, N log-uniform %d..%d, "
                           "M uniform %d..%d, %.0f s, latest_only=%s, max_inflight=%d, max_batch_evals=%.2g"
                           % (nsessions, len(maps), rate_hz, n_range[0], n_range[1], m_range[0], m_range[1],
                              duration_s, latest_only, max_inflight, max_batch_evals))
        out["session_setup_s"] = setup_s
        out["shape"] = autotune.shape_name(autotune.get_shape())
        if slo_ms:
            keep = min(1.0, out["served_evals_per_s"] / max(out["offered_evals_per_s"], 1.0))
            out["ladder"] = []
            for f in ladder:
                r = rate_hz * keep * f
                ph = _serve_phase(srv, sess, r, slo_duration_s, rng)
                ph["rate_hz"] = r
                out["ladder"].append(ph)
                if ph["p99_ms"] is not None and ph["p99_ms"] < slo_ms and ph["superseded"] <= 0.01 * ph["requests"]:
                    out["slo"] = dict(ph, slo_ms=slo_ms)
                    break
    finally:
        srv.close()
    return out


def _serve_phase(srv, sess, rate_hz, duration_s, rng):
    """One run of Poisson arrivals at rate_hz per session on an open server."""
    mods = np.array([0.1, 0.1, 0.1, 0.3])
    deltas = np.array([0.05, 0.0, 0.0, 0.01])
    ev_t, ev_s = [], []
    for i in range(len(sess)):
        t = rng.exponential(1.0 / rate_hz)
        while t < duration_s:
            ev_t.append(t)
            ev_s.append(i)
            t += rng.exponential(1.0 / rate_hz)
    order = np.argsort(ev_t, kind="stable")
    ev_t = np.asarray(ev_t)[order]
    ev_s = np.asarray(ev_s)[order]
    lat, nsup, offered = [], 0, 0

    def collect(results):
        nonlocal nsup
        for r in results:
            if r.superseded:
                nsup += 1
            else:
                lat.append(r.latency_ms)

    st0 = srv.stats()
    start = time.perf_counter()
    k = 0
    while k < len(ev_t):
        now = time.perf_counter() - start
        if now < ev_t[k]:
            collect(srv.poll())
            time.sleep(min(ev_t[k] - now, 0.0005))
            continue
        while k < len(ev_t) and ev_t[k] <= now:
            sid, n, m, xyz, rg = sess[ev_s[k]]
            srv.submit(sid, mods, deltas, xyz, rg, 0.8, 0.1)
            offered += n * m
            k += 1
    srv.drain()
    wall = time.perf_counter() - start
    collect(srv.poll(1 << 22))
    st = srv.stats()
    lat = np.array(lat)
    return {"rate_hz": rate_hz, "duration_s": duration_s, "requests": int(len(ev_t)), "ran": int(lat.size),
            "superseded": int(nsup), "batches": st["batches"] - st0["batches"],
            "offered_evals_per_s": offered / duration_s, "served_evals_per_s": (st["evals"] - st0["evals"]) / wall,
            "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
            "max_ms": float(lat.max()) if lat.size else None}


def serving_maps(amcl3d_amd, W, building=None, log=print):
    """The four resident maps of the serving workload: the 20x20x6 m room at
    0.1 m (200x200x60 cells), buildings of 50x50x10 m at 0.05 m
    (1000x1000x200), 100x100x30 m at 0.05 m (2000x2000x600; the caller's, if
    given) and 200x200x20 m at 0.1 m (2000x2000x200), all built on the device."""
    bm, grid = room_grid(amcl3d_amd, W)
    maps = [(bm, grid, 1.5)]
    specs = [((50.0, 50.0, 10.0), 0.05, 150, 11), ((100.0, 100.0, 30.0), 0.05, 500, 7),
             ((200.0, 200.0, 20.0), 0.1, 800, 13)]
    for ext, res, nbox, seed in specs:
        if ext[0] == 100.0 and building is not None:
            maps.append(building)
            continue
        t0 = time.time()
        b = W.building_map(seed, extent=ext, res=res, n_boxes=nbox)
        pts = b.points(dedup=False)
        g = amcl3d_amd.Grid3d()
        g.open(pts, b.min, b.max, b.res, 0.05)
        del pts
        log("serving map %s at %.2f m: %d cells (%.1fs)" % (ext, res, g.cells, time.time() - t0))
        maps.append((W.storey_view(b, 1.5, 3.0), g, 1.5))
    return maps


def map_file(amcl3d_amd, bm, sensor_dev=0.05, log=print):
    """The map as an OcTree file (a .bt of its occupied cells), read back as
    Grid3d::open reads it: openOcTree + computePointCloud on the host, then
    computeGrid on the device from the points read.  Writing the file is set
    up, not timed; the .grid cache Grid3d::open writes (8 bytes per cell) is
    left out."""
    import os
    import shutil
    import tempfile

    from . import octree

    d = tempfile.mkdtemp(prefix="a3d_map_")
    path = os.path.join(d, "building.bt")
    try:
        t0 = time.time()
        pts = bm.points(dedup=False)
        nodes = octree.write_points(path, pts, bm.res)
        del pts
        write_s = time.time() - t0
        nbytes = os.path.getsize(path)
        t1 = time.perf_counter()
        tree = octree.openOcTree(path)
        info = octree.computePointCloud(tree)
        read_s = time.perf_counter() - t1
        tree.close()
        g = amcl3d_amd.Grid3d()
        amcl3d_amd.synchronize()
        t2 = time.perf_counter()
        g.open(info.cloud, info.min, info.max, info.octo_resol, sensor_dev)
        amcl3d_amd.synchronize()
        build_s = time.perf_counter() - t2
        out = {"workload": "the 100x100x30 m building at 0.05 m as a .bt (%d nodes, %.1f MB): openOcTree + "
                           "computePointCloud on the host, computeGrid on the device" % (nodes, nbytes / 1e6),
               "occupied_points": int(info.cloud.shape[0]), "file_bytes": nbytes, "read_s": read_s,
               "read_MB_per_s": nbytes / read_s / 1e6, "read_points_per_s": info.cloud.shape[0] / read_s,
               "grid_cells": g.cells, "device_build_s": build_s, "open_without_cache_write_s": read_s + build_s,
               "file_write_s_untimed": write_s}
        g.close()
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def room_step(amcl3d_amd, W, grid, reps=50, warmup=5):
    """The reference's CPU-runnable configuration on the device: the room map
    (200 x 200 x 60 at 0.1 m), 1,000 particles, 2,000 points, 3 beacons; one
    synchronous predict -> update -> resample with host inputs, as the
    reference's node calls it.  Median and p99 of reps steps."""
    c = W.room_case(1)
    s = c["step"]
    pf = amcl3d_amd.ParticleFilter()
    times = []
    for r in range(warmup + reps):
        pf.set_particles(c["particles"])
        amcl3d_amd.synchronize()
        t0 = time.perf_counter()
        pf.predict(*s.mods, *s.deltas, s.z4)
        pf.update(grid, s.cloud, s.ranges, s.alpha, s.sigma)
        pf.resample(s.u01)
        pf.synchronize()
        if r >= warmup:
            times.append(time.perf_counter() - t0)
    pf.close()
    t = np.array(times) * 1e3
    return {"workload": "room map 200x200x60 at 0.1 m, 1,000 particles, 2,000 points, 3 beacons: one synchronous "
                        "predict -> update -> resample with host inputs", "reps": reps,
            "p50_ms": float(np.percentile(t, 50)), "p99_ms": float(np.percentile(t, 99))}


def node_step(amcl3d_amd, wl, n=100000, steps=20, warmup=3):
    """Latency of Node::pointcloudCallback (Node.cpp:112-191) through the
    library's node: the large filter's building map, trajectory, 20,000-point
    sensor clouds (voxelised at 0.1 m by the node) and 8 beacons, with the
    reference launch file's thresholds (amcl3d/launch/amcl3d.launch:
    voxel_size 0.1, update_rate 100, d_th 0.05, a_th 0.03, resample every
    update) and n particles.  Odometry is the trajectory relative to its
    first pose, which is also the initial pose."""
    import math

    from .node import Node

    traj = wl["traj"]
    x0, y0, z0, a0 = (float(v) for v in traj[0])
    params = dict(num_particles=n, resample_interval=0, set_initial_pose=1, init_x=x0, init_y=y0, init_z=z0,
                  init_a=a0, init_x_dev=0.05, init_y_dev=0.05, init_z_dev=0.05, init_a_dev=0.1, sensor_range=0.53,
                  voxel_size=0.1, odom_x_mod=0.1, odom_y_mod=0.1, odom_z_mod=0.01, odom_a_mod=0.3,
                  update_rate=100.0, d_th=0.05, a_th=0.03, take_off_height=-1.0, alpha=0.8)
    nd = Node(params, wl["grid"], seed=11)
    c, s_ = math.cos(a0), math.sin(a0)
    lat = []
    now = 0
    for k in range(min(steps + warmup, len(wl["clouds"]))):
        x, y, z, a = (float(v) for v in traj[k + 1])
        dx, dy = x - x0, y - y0
        rel = (c * dx + s_ * dy, -s_ * dx + c * dy, z - z0, a - a0)
        q = (0.0, 0.0, math.sin(rel[3] / 2), math.cos(rel[3] / 2))
        if k == 0:
            nd.odomCallback([0.0, 0.0, 0.0], (0.0, 0.0, 0.0, 1.0))  # initialises the filter
        nd.odomCallback(rel[:3], q)
        for r in wl["ranges"][k]:
            nd.rangeCallback(float(r[0]), float(r[1]), float(r[2]), float(r[3]))
        now += 100_000_000
        t0 = time.perf_counter()
        ran = nd.pointcloudCallback(now, wl["clouds"][k])
        dt = time.perf_counter() - t0
        if k >= warmup and ran:
            lat.append(dt * 1e3)
    st = nd.state()
    nd.close()
    lat = np.sort(np.asarray(lat))
    return {"workload": "node point-cloud callback: %d particles, 20,000-point clouds voxelised at 0.1 m, 8 beacons, "
                        "building map; host clouds in, mean on the host" % n,
            "updates_timed": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_xyz_final": [float(v) for v in st["mean"][:3]]}


def extras(amcl3d_amd, W, plain_value=None, building=None, building_map=None, wl=None, log=print,
           valu_per_64_evals=None):
    bm, grid = room_grid(amcl3d_amd, W)
    out = {"room_step": room_step(amcl3d_amd, W, grid)}
    log("room step:", out["room_step"])
    out["batched"] = batched(amcl3d_amd, W, bm, grid, map_name="the 200x200x60 room map")
    log("batched:", out["batched"]["batched_evals_per_s"], "evals/s")
    if building is not None:
        out["batched_large_map"] = batched(amcl3d_amd, W, building[0], building[1],
                                           map_name="the 100x100x30 m building at 0.05 m (2000x2000x600 cells)",
                                           converge=8)
        log("batched on the building map:", out["batched_large_map"]["batched_evals_per_s"], "evals/s")
    maps = serving_maps(amcl3d_amd, W, building, log=log)
    # the serving workload as specified: 1,024 sessions at 10 Hz each; then
    # the same sessions at the highest rate whose p99 stays under 10 ms
    sv = served(amcl3d_amd, W, maps, nsessions=1024, rate_hz=10.0, duration_s=6.0, slo_ms=10.0)
    out["served_slo"] = sv.pop("slo", None)
    out["served_ladder"] = [{k: ph[k] for k in ("rate_hz", "offered_evals_per_s", "served_evals_per_s", "p50_ms",
                                                "p99_ms", "superseded")} for ph in sv.pop("ladder", [])]
    out["served"] = sv
    log("served:", out["served"])
    log("served within a 10 ms p99:", out["served_slo"])
    if plain_value:
        out["served_over_plain_large_filter"] = out["served"]["served_evals_per_s"] / plain_value
    out["served_over_plain_loop"] = out["served"]["served_evals_per_s"] / out["batched"]["plain_loop_evals_per_s"]
    if building_map is not None:
        out["map_file"] = map_file(amcl3d_amd, building_map, log=log)
        log("map file:", out["map_file"])
    if wl is not None:
        out["node_step"] = node_step(amcl3d_amd, wl)
        log("node step:", out["node_step"])
    out["targets"] = targets(out, valu_per_64_evals)
    return out


def targets(out, valu_per_64_evals):
    """The served/batched targets beside what was measured, and what caps
    them.  Batched: at the cloud weight's vector instructions per evaluation,
    the chip's full issue rate (1,024 SIMDs at 2.4 GHz, a wave64 instruction
    per 4 cycles) is the most any schedule of the kernel runs, so the plain
    loop's rate bounds the ratio.  Served: at 10 Hz the server is saturated
    (requests are superseded), so what it completes is the GPU's rate on the
    mix of maps and sessions."""
    issue_peak = 1024 * 2.4e9 / 4
    ceiling = issue_peak * 64 / valu_per_64_evals if valu_per_64_evals else None
    sv = out["served"]
    t = {"batched_over_plain_loop": {"target": 10.0, "room_map": out["batched"]["batched_over_plain_loop"]},
         "served_p99_ms_at_10hz": {"target": 10.0, "measured": sv["p99_ms"], "p50": sv["p50_ms"]},
         "served_evals_per_s_at_10hz": {"offered": sv["offered_evals_per_s"], "served": sv["served_evals_per_s"],
                                        "superseded_fraction": sv["superseded"] / max(sv["requests"], 1)},
         "served_within_p99_10ms": None if not out.get("served_slo") else
         {k: out["served_slo"][k] for k in ("rate_hz", "served_evals_per_s", "p50_ms", "p99_ms")},
         "issue_bound_evals_per_s": ceiling,
         "valu_instr_per_64_evals": valu_per_64_evals}
    if "batched_large_map" in out:
        t["batched_over_plain_loop"]["building_map"] = out["batched_large_map"]["batched_over_plain_loop"]
        if "converged" in out["batched_large_map"]:
            t["batched_over_plain_loop"]["building_map_converged"] = \
                out["batched_large_map"]["converged"]["batched_over_plain_loop"]
    if ceiling:
        # the plain loop already runs the large filters, most of the work, near
        # the issue rate: no batching can beat it by more than this
        t["batched_over_plain_loop"]["cap_room_map"] = ceiling / out["batched"]["plain_loop_evals_per_s"]
    t["note"] = ("the serving configuration at 10 Hz offers %.3g evals/s; the server completed %.3g evals/s with "
                 "%.0f%% of requests superseded by their session's next scan: the offered load is above what one GPU "
                 "completes on this mix of maps and sessions, so queues grow (p99 %.0f ms) until requests are "
                 "superseded" % (sv["offered_evals_per_s"], sv["served_evals_per_s"],
                                 100.0 * sv["superseded"] / max(sv["requests"], 1), sv["p99_ms"] or 0.0))
    if ceiling:
        t["note"] += ("; batched: at %.1f vector instructions per 64 evaluations the chip issues at most %.3g "
                      "evals/s, %.1fx the plain loop's rate on the room map" % (
                          valu_per_64_evals, ceiling, t["batched_over_plain_loop"]["cap_room_map"]))
    return t
