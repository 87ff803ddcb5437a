"""Profiling workload: the cloud-weight kernel of the large-filter workload (or,
with ``--fleet``, the persistent batched launch of a 256-filter fleet), then a
device copy of known size.

Under ``rocprofv3 --pmc ...`` the copy kernel's FETCH_SIZE / WRITE_SIZE
(it reads and writes ``--copy-mib`` MiB) calibrate the counters that the
cloud-weight dispatches report.

    python -m amcl3d_amd.profile_run --updates 3 --copy-mib 1024 --shape 0,16,1
    python -m amcl3d_amd.profile_run --fleet building --updates 3 --shape 0,16,1
"""
import argparse
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--updates", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--particles", type=int, default=100000)
    ap.add_argument("--points", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--shape", default="", help="cloud-weight kernel shape 'MAPPING,PPW,ORDER' (default: the library's)")
    ap.add_argument("--fleet", choices=("room", "building"), default="",
                    help="profile the batched launch of 256 filters on this map instead")
    args = ap.parse_args(argv)

    import amcl3d_amd
    from amcl3d_amd import _lib, benchlib, workloads as W

    amcl3d_amd.load()
    if args.shape:
        from amcl3d_amd import autotune

        autotune.set_shape(tuple(int(v) for v in args.shape.split(",")))
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.fleet:
        run_fleet(amcl3d_amd, benchlib, W, args, log)
    else:
        run_large_filter(amcl3d_amd, benchlib, W, args, log)
    nbytes = args.copy_mib << 20
    a = amcl3d_amd.DeviceBuffer(nbytes)
    b = amcl3d_amd.DeviceBuffer(nbytes)
    for _ in range(3):
        t0 = time.perf_counter()
        _lib.call("a3d_buffer_copy", b.ptr, a.ptr, nbytes)
        log("copy %d MiB: %.3f ms" % (args.copy_mib, (time.perf_counter() - t0) * 1e3))
    print("profile_run done", flush=True)


def run_fleet(amcl3d_amd, benchlib, W, args, log):
    """The batched workload: 256 filters (N log-uniform 500..50k, M 1k..30k)
    on the room map or on the building map, one persistent launch per update."""
    from amcl3d_amd import fleet as fleet_mod

    if args.fleet == "room":
        bm, grid = benchlib.room_grid(amcl3d_amd, W)
    else:
        bm = W.building_map(args.seed)
        grid = amcl3d_amd.Grid3d()
        grid.open(bm.points(dedup=False), bm.min, bm.max, bm.res, 0.05)
    counts, parts, clouds, ranges = benchlib.fleet_case(W, bm, 256, 3)
    fl = fleet_mod.Fleet(counts)
    fl.set_particles(np.concatenate(parts))
    fl.stage_update(clouds, ranges, 0.8, 0.1)
    for _ in range(args.updates):
        fl.update_staged(grid, fleet_mod.PERSISTENT)
        fl.synchronize()
    ms, n, ev = fl.timing()
    log("fleet cloud weight on the %s map: %d launches, %.3f ms average, %.3e evals/s"
        % (args.fleet, n, ms, ev / n / (ms * 1e-3)))
    fl.close()


def run_large_filter(amcl3d_amd, benchlib, W, args, log):
    wl = benchlib.large_filter_workload(amcl3d_amd, W, args.updates, args.seed, args.particles, args.points, 8,
                                        log=log)
    pf = amcl3d_amd.ParticleFilter()
    pf.set_particles(wl["parts"])
    mods = np.array([0.1, 0.1, 0.1, 0.3])
    for k in range(args.updates):
        c = wl["clouds"][k]
        x, y, z = (amcl3d_amd.DeviceBuffer.from_array(np.ascontiguousarray(c[:, j])) for j in range(3))
        r = amcl3d_amd.DeviceBuffer.from_array(wl["ranges"][k])
        dz = amcl3d_amd.DeviceBuffer.from_array(wl["z4"][k])
        pf.predict_dev(mods, wl["deltas"][k], dz)
        pf.update_dev(wl["grid"], x.ptr, y.ptr, z.ptr, args.points, r.ptr, 8, 0.8, 0.1)
        pf.resample_dev(wl["u01"][k])
        pf.synchronize()
    ms, n, ev = pf.timing()
    log("cloud weight: %d launches, %.3f ms average, %.3e evals/s" % (n, ms, ev / n / (ms * 1e-3)))


if __name__ == "__main__":
    main()
