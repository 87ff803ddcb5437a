"""Full-scale serving with a device-side map build (one MI355X).

    python -m amcl3d_amd.full_scale [--duration 5]

Map: a seeded city block of 200 x 200 x 40 m at 0.05 m (4000 x 4000 x 800 =
1.28e10 cells, 51.2 GB of FP32 prob, 64-bit cell indices) whose ~49M
occupied points (ground slab and 85 buildings) go through computeGrid on the
device (PointCloudTools.cpp:84-149).  tests/test_gpu_full_scale.py checks the
same build on sampled cells.

Sessions: 64 global-localisation sessions of 1,000,000 particles drawn
uniformly over the free space and 2,048 tracking sessions of U(5,000, 20,000)
particles around their pose, all with 10,000-point clouds, each sending
Poisson update requests at 10 Hz (the arrival process of the serving
workload) with latest-scan-wins queueing.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def free_particles(bm, n, rng):
    """n particles uniform over the free space of a box map (above the ground,
    outside every building), w = 1/n."""
    out = np.zeros((0, 3))
    lo = bm.min + np.array([0.0, 0.0, 2 * bm.res])
    hi = bm.max - bm.res
    while out.shape[0] < n:
        p = rng.uniform(lo, hi, (2 * n, 3))
        keep = np.ones(p.shape[0], bool)
        for b in bm.boxes[1:]:
            keep &= ~np.all((p >= b[:3]) & (p <= b[3:]), 1)
        out = np.concatenate([out, p[keep]])
    parts = np.zeros((n, 7), np.float32)
    parts[:, :3] = out[:n]
    parts[:, 3] = rng.uniform(-np.pi, np.pi, n)
    parts[:, 4] = 1.0 / n
    return parts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--duration", type=float, default=5.0)
    ap.add_argument("--rate", type=float, default=10.0)
    ap.add_argument("--global-sessions", type=int, default=64)
    ap.add_argument("--tracking-sessions", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    sys.path.insert(0, ROOT)
    import amcl3d_amd
    from amcl3d_amd import workloads as W

    amcl3d_amd.load()
    rng = np.random.default_rng(args.seed)
    out = {"workload": "full-scale serving: city block 200x200x40 m at 0.05 m, computeGrid on the device, "
                       "%d global sessions x 1e6 particles + %d tracking sessions, M=10,000, Poisson %.0f Hz"
                       % (args.global_sessions, args.tracking_sessions, args.rate)}
    t0 = time.time()
    bm = W.city_map(args.seed)
    pts = bm.points(dedup=False)
    log("city map: %s cells, %d occupied points (%.1fs)" % (bm.size, pts.shape[0], time.time() - t0))
    grid = amcl3d_amd.Grid3d()
    amcl3d_amd.synchronize()
    t1 = time.time()
    grid.open(pts, bm.min, bm.max, bm.res, 0.05)
    amcl3d_amd.synchronize()
    build_s = time.time() - t1
    out["map"] = {"cells": int(grid.cells), "points": int(pts.shape[0]), "prob_bytes": int(grid.cells) * 4,
                  "build_s": build_s, "cells_per_s": grid.cells / build_s}
    log("computeGrid on the device: %d cells in %.2fs" % (grid.cells, build_s))
    del pts

    # sessions
    from amcl3d_amd import serving

    srv = serving.Server(max_inflight=2, max_batch_sessions=4096, max_batch_evals=3_000_000_000, latest_only=True,
                         hbm_budget_bytes=160 << 30)
    mid = srv.add_map(grid)
    t2 = time.time()
    poses = []
    for _ in range(16):
        pose = np.append(W.free_pose(bm, rng, z_range=(1.5, 1.5)), rng.uniform(-np.pi, np.pi)).astype(np.float32)
        cloud = W.sensor_cloud(bm, pose, 10000, 0.05, rng)
        poses.append((pose, tuple(np.ascontiguousarray(cloud[:, j]) for j in range(3)),
                      W.beacons(bm, pose, 3, 0.1, rng)))
    sets = [free_particles(bm, 1_000_000, rng) for _ in range(2)]
    sess = []  # (sid, n, xyz, ranges, is_global)
    for s in range(args.global_sessions):
        _, xyz, rg = poses[s % len(poses)]
        sid = srv.open_session_particles(mid, sets[s % len(sets)], 1000 + s)
        sess.append((sid, 1_000_000, xyz, rg, True))
    for s in range(args.tracking_sessions):
        pose, xyz, rg = poses[int(rng.integers(0, len(poses)))]
        n = int(rng.integers(5000, 20001))
        sid = srv.open_session(mid, n, pose, [0.3, 0.3, 0.1, 0.2], 100000 + s)
        sess.append((sid, n, xyz, rg, False))
    out["session_setup_s"] = time.time() - t2
    log("sessions open (%.1fs)" % out["session_setup_s"])
    ev_t, ev_s = [], []
    for i in range(len(sess)):
        t = rng.exponential(1.0 / args.rate)
        while t < args.duration:
            ev_t.append(t)
            ev_s.append(i)
            t += rng.exponential(1.0 / args.rate)
    order = np.argsort(ev_t, kind="stable")
    ev_t = np.asarray(ev_t)[order]
    ev_s = np.asarray(ev_s)[order]
    is_global = {s[0]: s[4] for s in sess}
    lat = {True: [], False: []}
    nsup = {True: 0, False: 0}
    offered = {True: 0, False: 0}

    def collect(results):
        for r in results:
            g = is_global[r.session]
            if r.superseded:
                nsup[g] += 1
            else:
                lat[g].append(r.latency_ms)

    mods = [0.1, 0.1, 0.1, 0.3]
    deltas = [0.05, 0.0, 0.0, 0.01]
    start = time.perf_counter()
    k = 0
    while k < len(ev_t):
        now = time.perf_counter() - start
        if now < ev_t[k]:
            collect(srv.poll())
            time.sleep(min(ev_t[k] - now, 0.0005))
            continue
        while k < len(ev_t) and ev_t[k] <= now:
            sid, n, xyz, rg, g = sess[ev_s[k]]
            srv.submit(sid, mods, deltas, xyz, rg, 0.8, 0.1)
            offered[g] += n * xyz[0].shape[0]
            k += 1
    srv.drain()
    wall = time.perf_counter() - start
    collect(srv.poll(1 << 22))
    st = srv.stats()
    srv.close()
    for g, name in ((True, "global"), (False, "tracking")):
        a = np.asarray(lat[g])
        out[name] = {"offered_evals_per_s": offered[g] / args.duration, "ran": int(a.size),
                     "superseded": int(nsup[g]),
                     "p50_ms": float(np.percentile(a, 50)) if a.size else None,
                     "p99_ms": float(np.percentile(a, 99)) if a.size else None}
    out["served_evals_per_s"] = st["evals"] / wall
    out["batches"] = st["batches"]
    out["requests"] = int(len(ev_t))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
