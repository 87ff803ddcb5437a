#!/usr/bin/env python
"""Particle.point evaluations per second of the amcl3d measurement update on one MI355X.

Workload (the single-large-filter configuration): a seeded procedural building,
100 x 100 x 30 m at 0.05 m (2000 x 2000 x 600 cells, 9.6 GB of FP32 prob),
built on the device by computeGrid; one filter of 100,000 particles following
a seeded smooth trajectory; per step a 20,000-point cloud ray-cast from the
true pose with N(0, 0.03 m) noise and 8 range beacons.  A step is
predict -> update (cloud weight, range weight, normalise, mean) -> resample.
The timed steps run as one captured hipGraph with every input resident in
HBM.  The unit of work is one particle.point evaluation (the inner body of
Grid3d::computeCloudWeight, amcl3d/src/Grid3d.cpp:168-196), counted for
in-map particles only.

Prints one JSON line.  Besides the timed graph it runs, outside the timed
region: the same steps eagerly with HIP events around every cloud-weight
launch (roofline), a bit-exact comparison of the two runs, the oracle on a
bounded sample of the same step on one host core (cpu_baseline), and the
batched / served forms on their own workloads (extras).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle·point evals/s; p50/p99 update latency, served vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0
BYTES_PER_EVAL = 4  # one FP32 prob gather


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(amcl3d_amd, W, steps_total, seed, n, m, nb):
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
                deltas=deltas, z4=z4, u01=u01)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--particles", type=int, default=100000)
    ap.add_argument("--points", type=int, default=20000)
    ap.add_argument("--beacons", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU time of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the filter path has no multi-GPU form "
              "(sharding is described in DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import amcl3d_amd
    from amcl3d_amd import workloads as W

    amcl3d_amd.load()
    info = amcl3d_amd.device_info()
    log("device:", info)
    K, Wm = args.steps, args.warmup
    wl = build_workload(amcl3d_amd, W, K + Wm, args.seed, args.particles, args.points, args.beacons)
    grid = wl["grid"]
    mods = np.array([0.1, 0.1, 0.1, 0.3])
    alpha, sigma = 0.8, 0.1

    # inputs resident in HBM
    dbuf = []
    for k in range(K + Wm):
        c = wl["clouds"][k]
        x, y, z = (amcl3d_amd.DeviceBuffer.from_array(np.ascontiguousarray(c[:, j])) for j in range(3))
        r = amcl3d_amd.DeviceBuffer.from_array(wl["ranges"][k])
        dz = amcl3d_amd.DeviceBuffer.from_array(wl["z4"][k])
        dbuf.append((x, y, z, r, dz))

    pf = amcl3d_amd.ParticleFilter()
    pf.set_particles(wl["parts"])

    def enqueue(k):
        x, y, z, r, dz = dbuf[k]
        pf.predict_dev(mods, wl["deltas"][k], dz)
        pf.update_dev(grid, x.ptr, y.ptr, z.ptr, args.points, r.ptr, args.beacons, alpha, sigma)
        pf.resample_dev(wl["u01"][k])

    # warm-up (eager), then capture the K timed steps
    for k in range(Wm):
        enqueue(k)
    pf.synchronize()
    start_state = pf.particles()
    pf.capture_begin()
    for k in range(Wm, Wm + K):
        enqueue(k)
    graph = pf.capture_end()

    amcl3d_amd.synchronize()
    t0 = time.perf_counter()
    graph.launch()
    pf.synchronize()
    amcl3d_amd.synchronize()
    elapsed = time.perf_counter() - t0
    graph_state = pf.particles()
    graph_mean = pf.getMean()

    # the same steps eagerly, HIP events around each cloud-weight launch;
    # particles are read back after each step to count in-map particles
    pf.set_particles(start_state)
    pf.timing(reset=True)
    inmap_evals = 0
    bm = wl["bm"]
    for k in range(Wm, Wm + K):
        x, y, z, r, dz = dbuf[k]
        pf.predict_dev(mods, wl["deltas"][k], dz)
        pf.synchronize()
        p = pf.particles()
        inside = np.all((p[:, :3].astype(np.float64) >= bm.min) & (p[:, :3].astype(np.float64) < bm.max), 1)
        inmap_evals += int(inside.sum()) * args.points
        pf.update_dev(grid, x.ptr, y.ptr, z.ptr, args.points, r.ptr, args.beacons, alpha, sigma)
        pf.resample_dev(wl["u01"][k])
    weight_ms, launches, _ = pf.timing(reset=True)
    eager_state = pf.particles()
    identical = bool(np.array_equal(eager_state.view(np.int32), graph_state.view(np.int32)) and
                     np.array_equal(pf.getMean(), graph_mean))
    track_err = float(np.linalg.norm(graph_mean[:2] - wl["traj"][Wm + K][:2]))

    value = inmap_evals / elapsed
    evals_per_launch = inmap_evals / K
    achieved = evals_per_launch * BYTES_PER_EVAL / (weight_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 offsets, compares and cell division, as the reference)",
        "data": "synthetic",
        "config": {
            "workload": "single large filter, plain FP32 drop-in, hipGraph predict->update->resample chain",
            "map": "seeded building 100x100x30 m at 0.05 m (2000x2000x600 cells), computeGrid on the device",
            "particles": args.particles,
            "points": args.points,
            "beacons": args.beacons,
            "cells": int(grid.cells),
            "seed": args.seed,
            "tile_particles_per_wave": amcl3d_amd.get_tile(),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "cloud weight (weight_single_kernel)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_ms_avg": weight_ms,
            "kernel_launches": launches,
            "algorithmic_bytes_per_launch": evals_per_launch * BYTES_PER_EVAL,
        },
        "checks": {"graph_equals_eager_bit_exact": identical, "mean_xy_error_m": track_err,
                   "map_build_s": wl["build_s"], "map_build_cells_per_s": grid.cells / wl["build_s"]},
        "device": info["name"],
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(grid, wl, start_state, args, mods, alpha, sigma, Wm)
    if not args.no_extras:
        try:
            from amcl3d_amd import benchlib

            out["extras"] = benchlib.extras(amcl3d_amd, W, plain_value=value,
                                            building=(W.storey_view(wl["bm"], 1.5, 3.0), grid, 1.5))
        except Exception as e:  # extras never hide the main line
            out["extras"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


def cpu_baseline(grid, wl, start_state, args, mods, alpha, sigma, Wm):
    """The oracle (the CPU restatement of the reference) on one host core, on a
    bounded sample of the first timed step: its predict and update for the
    first S particles with the full cloud; S is sized to about args.cpu_seconds."""
    import oracle

    from amcl3d_amd import workloads as W

    log("cpu baseline: reading the grid back (%.1f GB)" % (grid.cells * 4 / 1e9))
    g = oracle.GridMeta(grid.min, grid.max, grid.res, grid.read_cells())
    k = Wm
    cloud, ranges = wl["clouds"][k], wl["ranges"][k]

    def run(s):
        p = oracle.pf_predict(start_state[:s], mods, wl["deltas"][k], wl["z4"][k][:s])
        t = time.perf_counter()
        oracle.pf_update(p, g, cloud, ranges, alpha, sigma, nthreads=1)
        return time.perf_counter() - t

    t = run(100)
    s = int(min(args.particles, max(100, 100 * args.cpu_seconds / max(t, 1e-6))))
    t = run(s)
    del g
    return {"value": s * args.points / t, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": "oracle ParticleFilter::update of the first %d of %d particles x %d points, one step, "
                      "single thread, %.1f s" % (s, args.particles, args.points, t),
            "host_cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
