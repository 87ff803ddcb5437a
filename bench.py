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
region: the same steps eagerly, back to back, with HIP events around every cloud-weight
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
# Per-update counters of the cloud-weight kernel on this workload (every step
# k of a seed's workload is the same whatever the run's length): HBM bytes
# read and written (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, separate runs,
# each scaled by its reading for a 1 GiB device copy) and vector instructions
# issued (SQ_INSTS_VALU).  Counters need profiler runs of their own, so the
# table is measured apart from this run; the bench averages it over exactly
# the updates it times.
PROFILE_TABLE = os.path.join(ROOT, "profiles", "cloud_weight_large_filter_per_update.md")
# The chip's vector issue rate: 1024 SIMDs at 2.4 GHz, 4 cycles per wave64
# instruction.
ISSUE_PEAK = 1024 * 2.4e9 / 4
# The shader clock the same kernel ran at, per update (rocprofv3 --pmc
# GRBM_GUI_ACTIVE over the kernel trace's duration, one run of the same
# workload and shape).
CLOCK_TABLE = os.path.join(ROOT, "profiles", "cloud_weight_large_filter_clock.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(amcl3d_amd, W, steps_total, seed, n, m, nb):
    from amcl3d_amd import benchlib

    return benchlib.large_filter_workload(amcl3d_amd, W, steps_total, seed, n, m, nb, log=log)


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
    ap.add_argument("--shape", default="profiled",
                    help="cloud-weight kernel shape 'MAPPING,PPW,ORDER[,TILE]' (e.g. 0,16,1,64), 'auto' to time "
                         "every shape and keep the fastest, or 'profiled' (default): the shape of the profile table "
                         "(the autotuner's choice when it was profiled), else 'auto'")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU. Filters and sessions are independent, so several GPUs would each "
              "run their own share with a replica of the maps; there is no multi-GPU form to time here. "
              "Run with --gpus 1.", file=sys.stderr)
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
    tuning = choose_shape(amcl3d_amd, pf, grid, dbuf[Wm], args, alpha, sigma)
    pf.set_particles(start_state)
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

    # the same steps eagerly, particles read back after each predict to count
    # the in-map particles; then again back to back (no host wait between
    # steps, as in the graph) with HIP events around each cloud-weight launch
    pf.set_particles(start_state)
    inmap_evals = 0
    evals_per_update = []
    bm = wl["bm"]
    for k in range(Wm, Wm + K):
        x, y, z, r, dz = dbuf[k]
        pf.predict_dev(mods, wl["deltas"][k], dz)
        pf.synchronize()
        p = pf.particles()
        inside = np.all((p[:, :3].astype(np.float64) >= bm.min) & (p[:, :3].astype(np.float64) < bm.max), 1)
        inmap_evals += int(inside.sum()) * args.points
        evals_per_update.append(int(inside.sum()) * args.points)
        pf.update_dev(grid, x.ptr, y.ptr, z.ptr, args.points, r.ptr, args.beacons, alpha, sigma)
        pf.resample_dev(wl["u01"][k])
    pf.synchronize()
    pf.set_particles(start_state)
    pf.timing(reset=True)
    for k in range(Wm, Wm + K):
        enqueue(k)
    pf.synchronize()
    weight_ms, launches, _ = pf.timing(reset=True)
    eager_state = pf.particles()
    identical = bool(np.array_equal(eager_state.view(np.int32), graph_state.view(np.int32)) and
                     np.array_equal(pf.getMean(), graph_mean))
    track_err = float(np.linalg.norm(graph_mean[:2] - wl["traj"][Wm + K][:2]))

    value = inmap_evals / elapsed
    evals_per_launch = inmap_evals / K
    counters = profile_counters(PROFILE_TABLE, Wm, K, evals_per_update, tuning["chosen"], seed=args.seed,
                                n=args.particles, m=args.points)
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
            "kernel_shape": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "cloud weight (%s)" % ("lanes_single_kernel" if amcl3d_amd.get_mapping() == 0
                                             else "weight_single_kernel"),
            "kernel_note": "the particle visiting order (4 small kernels, outside the HIP events) is in ms_per_step",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": counters.get("traffic"),
            "traffic_note": counters["note"],
            "issue": issue_block(counters, evals_per_launch, weight_ms, profile_clock(CLOCK_TABLE, Wm, K)),
            "kernel_ms_avg": weight_ms,
            "kernel_ms_profiled": counters.get("kernel_ms"),
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
                                            building=(W.storey_view(wl["bm"], 1.5, 3.0), grid, 1.5),
                                            building_map=wl["bm"], wl=wl, log=log,
                                            valu_per_64_evals=counters.get("valu_per_64_evals"))
        except Exception as e:  # extras never hide the main line
            out["extras"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


def profile_counters(path, warmup, steps, evals_per_update, shape, seed, n, m):
    """Means over updates warmup+1 .. warmup+steps (1-based, as profiled) of the
    profiled per-update counters, when the table is for this workload and
    kernel shape."""
    out = {"note": "no profile table"}
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return out
    meta, rows = {}, {}
    for line in lines:
        if line.startswith("- ") and ": " in line:
            k, v = line[2:].split(": ", 1)
            meta[k.strip()] = v.strip().strip("`")
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0].isdigit():
            rows[int(cells[0])] = [float(c) for c in cells[1:]]
    want = {"shape": shape, "seed": str(seed), "particles": str(n), "points": str(m)}
    if any(meta.get(k) != v for k, v in want.items()):
        out["note"] = "the profile table (%s) is for %s, this run is %s" % (
            os.path.relpath(path, ROOT), {k: meta.get(k) for k in want}, want)
        return out
    upd = list(range(warmup + 1, warmup + steps + 1))
    if not all(u in rows for u in upd):
        out["note"] = "the profile table covers updates 1..%d, this run times %d..%d" % (max(rows), upd[0], upd[-1])
        return out
    fetch = [rows[u][0] for u in upd]
    write = [rows[u][1] for u in upd]
    valu = [rows[u][2] for u in upd]
    out["traffic"] = float(sum(fetch) + sum(write)) / len(upd)
    out["read"] = float(sum(fetch)) / len(upd)
    out["written"] = float(sum(write)) / len(upd)
    out["valu_per_64_evals"] = float(sum(valu)) / (sum(evals_per_update) / 64.0)
    if all(len(rows[u]) > 3 for u in upd):
        out["kernel_ms"] = float(sum(rows[u][3] for u in upd)) / len(upd) / 1e3
    out["note"] = ("HBM bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch (separate runs of the same "
                   "workload and shape, each scaled by a 1 GiB copy's reading), mean over updates %d..%d, the ones "
                   "timed here (%s)" % (upd[0], upd[-1], os.path.relpath(path, ROOT)))
    return out


def profiled_shape(path):
    """The kernel shape ('MAPPING,PPW,ORDER,TILE') the profile table names."""
    try:
        for line in open(path):
            if line.startswith("- shape_args: "):
                return line.split(": ", 1)[1].strip().strip("`")
    except OSError:
        pass
    return None


def profile_clock(path, warmup, steps):
    """Mean shader clock (GHz) over updates warmup+1 .. warmup+steps of the
    clock table, or None when it does not cover them."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    clk = {}
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0].isdigit():
            clk[int(cells[0])] = float(cells[3])
    upd = range(warmup + 1, warmup + steps + 1)
    if not all(u in clk for u in upd):
        return None
    return sum(clk[u] for u in upd) / len(upd)


def issue_block(counters, evals_per_launch, weight_ms, clock_ghz=None):
    if "valu_per_64_evals" not in counters:
        return None
    v = counters["valu_per_64_evals"]
    rate = evals_per_launch / 64 * v / (weight_ms * 1e-3)
    out = {"valu_instr_per_64_evals": v, "achieved_wave_instr_per_s": rate, "peak_wave_instr_per_s": ISSUE_PEAK,
           "frac": rate / ISSUE_PEAK,
           "note": "SQ_INSTS_VALU per dispatch from the same profile table, over the same updates; the kernel is "
                   "bound by vector instruction issue, not HBM: per evaluation the reference's FP64 offset add, "
                   "float rounding, bounds and cell arithmetic per axis"}
    if clock_ghz:
        out["clock_ghz_profiled"] = clock_ghz
        out["frac_at_profiled_clock"] = rate / (1024 * clock_ghz * 1e9 / 4)
        out["clock_note"] = ("the shader clock the kernel ran at over the same updates (GRBM_GUI_ACTIVE per XCD over "
                             "the traced duration, %s); frac_at_profiled_clock is the issue rate over what the SIMDs "
                             "issue at that clock" % os.path.relpath(CLOCK_TABLE, ROOT))
    return out


def choose_shape(amcl3d_amd, pf, grid, bufs, args, alpha, sigma):
    """Kernel shape for the timed run: the one given, or the fastest of every
    shape on the update of the first timed step (all shapes give identical bits)."""
    from amcl3d_amd import autotune

    spec = args.shape
    if spec == "profiled":
        spec = profiled_shape(PROFILE_TABLE) or "auto"
    if spec != "auto":
        sh = tuple(int(v) for v in spec.split(","))
        autotune.set_shape(sh)
        return {"chosen": autotune.shape_name(sh), "from": args.shape}
    x, y, z, r, _ = bufs

    def run():
        pf.update_dev(grid, x.ptr, y.ptr, z.ptr, args.points, r.ptr, args.beacons, alpha, sigma)
        pf.synchronize()

    best, timings = autotune.tune(run)
    log("shape autotune:", {autotune.shape_name(k): round(v * 1e3, 3) for k, v in timings.items()}, "ms")
    return {"chosen": autotune.shape_name(best),
            "update_ms": {autotune.shape_name(k): v * 1e3 for k, v in timings.items()}}


def cpu_baseline(grid, wl, start_state, args, mods, alpha, sigma, Wm):
    """The oracle (the CPU restatement of the reference) on one host core.

    value: a bounded sample of the first timed step of this workload, its
    predict and update for the first S particles with the full cloud; S is
    sized to about args.cpu_seconds.  legs: the room configuration's full
    predict -> update -> resample step (1,000 particles, 2,000 points, 3
    beacons; median of 7), the host computeGrid (exact nearest neighbour) of
    that 200 x 200 x 60 map and a seeded sample of filters of the batched
    workload's distribution on it, single-threaded like the reference; and the large-filter sample on every
    host core (not the reference, which is serial)."""
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
    legs = {}
    # the same sample on every host core this process may use (the oracle's
    # particles split over threads): not the reference, which is serial
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    p = oracle.pf_predict(start_state[:s], mods, wl["deltas"][k], wl["z4"][k][:s])
    t0 = time.perf_counter()
    oracle.pf_update(p, g, cloud, ranges, alpha, sigma, nthreads=threads)
    tm = time.perf_counter() - t0
    legs["large_filter_all_cores"] = {"kind": "not the reference (the oracle's update split over threads)",
                                      "threads": threads, "particles": s, "seconds": tm,
                                      "evals_per_s": s * args.points / tm}
    del g
    # the room configuration: map build, then one full step
    c = W.room_case(1)
    bm = c["map"]
    pts = bm.points()
    t0 = time.perf_counter()
    _, prob = oracle.compute_grid(pts, bm.min, bm.max, bm.res, c["sensor_dev"], nthreads=1)
    tg = time.perf_counter() - t0
    cells = int(np.prod(bm.size))
    legs["room_compute_grid"] = {"cells": cells, "occupied_points": int(pts.shape[0]), "seconds": tg,
                                 "cells_per_s": cells / tg, "threads": 1}
    og = oracle.GridMeta(bm.min, bm.max, bm.res, prob)
    st, parts = c["step"], c["particles"]
    full, upd = [], []
    for _ in range(7):
        t0 = time.perf_counter()
        op = oracle.pf_predict(parts, st.mods, st.deltas, st.z4)
        t1 = time.perf_counter()
        op, _ = oracle.pf_update(op, og, st.cloud, st.ranges, st.alpha, st.sigma, nthreads=1)
        t2 = time.perf_counter()
        oracle.pf_resample(op, st.u01)
        full.append(time.perf_counter() - t0)
        upd.append(t2 - t1)
    inside = np.all((parts[:, :3].astype(np.float64) >= bm.min) & (parts[:, :3].astype(np.float64) < bm.max), 1)
    ev = int(inside.sum()) * st.cloud.shape[0]
    legs["room_full_step"] = {"particles": int(parts.shape[0]), "points": int(st.cloud.shape[0]),
                              "beacons": int(st.ranges.shape[0]), "full_step_ms_median": 1e3 * float(np.median(full)),
                              "update_ms_median": 1e3 * float(np.median(upd)),
                              "update_evals_per_s": ev / float(np.median(upd)), "runs": len(full), "threads": 1}
    # filters of the batched workload's distribution (N log-uniform 500..50k,
    # M 1k..30k, 1..8 beacons) on the same room map: a seeded sample of 16,
    # one full step each, single-threaded, until about 5 s of CPU time
    from amcl3d_amd import benchlib

    counts, fparts, fclouds, franges = benchlib.fleet_case(W, bm, 16, 29)
    rng = np.random.default_rng(12)
    evals, secs, done = 0, 0.0, 0
    for f in range(16):
        pf_, cl, rg = fparts[f], fclouds[f], franges[f]
        z4 = rng.standard_normal((pf_.shape[0], 4)).astype(np.float32)
        t0 = time.perf_counter()
        op = oracle.pf_predict(pf_, st.mods, st.deltas, z4)
        op, _ = oracle.pf_update(op, og, cl, rg, 0.8, 0.1, nthreads=1)
        oracle.pf_resample(op, 0.5)
        secs += time.perf_counter() - t0
        inside = np.all((op[:, :3].astype(np.float64) >= bm.min) & (op[:, :3].astype(np.float64) < bm.max), 1)
        evals += int(inside.sum()) * cl.shape[0]
        done += 1
        if secs > 5.0:
            break
    legs["batched_filters_sample"] = {"filters": done, "seed": 29, "evals": evals, "seconds": secs,
                                      "evals_per_s": evals / secs, "threads": 1}
    return {"value": s * args.points / t, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": "oracle ParticleFilter::update of the first %d of %d particles x %d points, one step, "
                      "single thread, %.1f s" % (s, args.particles, args.points, t),
            "host": host_info(), "legs": legs}


def host_info():
    """The host CPU as lscpu names it, and how many cores this process may use."""
    import subprocess

    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            key = key.strip()
            if key in ("Model name", "BIOS Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                       "CPU(s)", "CPU max MHz"):
                info["lscpu " + key] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpuinfo model name"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return info


if __name__ == "__main__":
    main()
