"""Shape autotuner for the cloud-weight kernel.

A shape is (mapping, particles per wavefront, visiting order, points per
LDS cloud tile):
  * mapping 0 (MAP_PARTICLE_LANES): each lane carries one particle and
    64 / ppw consecutive cloud points per step (ppw 64, 32 or 16); the
    workgroup stages the cloud in LDS tiles of 32, 64, 128 or 256 points,
    one barrier per tile;
  * mapping 1 (MAP_POINT_LANES): the wavefront's 64 lanes are 64 consecutive
    points, transformed for ppw particles through an LDS tile (tiles are
    always 64 points, one per lane);
  * order 0 visits particles in index order, order 1 in a spatial order
    computed on the device before each weight launch (per filter, also in
    fleets and served batches);
  * persistent launches (fleets, served batches) add a fifth element, how the
    work list is dealt to the XCDs: 0 contiguous ranges, k runs of k blocks
    round-robin.
Every shape gives bit-identical results, so the choice is purely a timing
one: run the caller's workload once per candidate (after one untimed warm-up)
and keep the fastest.
"""
import time

from . import _lib

CLOUD_TILES = (32, 64, 128, 256)
CANDIDATES = tuple([(0, p, o, t) for o in (1, 0) for p in (64, 32, 16) for t in CLOUD_TILES] +
                   [(1, p, o, 64) for o in (1, 0) for p in (16, 32, 64)])
# fleets and served batches always visit particles in spatial order (index
# order loses on every measured workload, and a fleet update is long to time)
LOOP_CANDIDATES = tuple(c for c in CANDIDATES if c[2] == 1)
DEALS = (0, 4)
FLEET_CANDIDATES = tuple(c + (d,) for d in DEALS for c in LOOP_CANDIDATES)


def set_shape(shape):
    _lib.set_mapping(shape[0])
    _lib.set_tile(shape[1])
    if len(shape) > 2:
        _lib.set_order(shape[2])
    if len(shape) > 3:
        _lib.set_cloud_tile(shape[3])
    if len(shape) > 4:
        _lib.set_deal(shape[4])


def get_shape():
    return (_lib.get_mapping(), _lib.get_tile(), _lib.get_order(), _lib.get_cloud_tile(), _lib.get_deal())


def tune(run, candidates=CANDIDATES, reps=2, prepare=None):
    """run(): executes and synchronizes the workload once.  prepare(shape):
    optional re-staging after the shape changes.  Returns (best, {shape: s})."""
    timings = {}
    for sh in candidates:
        set_shape(sh)
        if prepare:
            prepare(sh)
        run()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            best = min(best, time.perf_counter() - t0)
        timings[sh] = best
    choice = min(timings, key=timings.get)
    set_shape(choice)
    if prepare:
        prepare(choice)
    return choice, timings


def shape_name(shape):
    name = "map%d-ppw%d" % tuple(shape[:2])
    if len(shape) > 2:
        name += "-" + ("spatial" if shape[2] else "index")
    if len(shape) > 3:
        name += "-tile%d" % shape[3]
    if len(shape) > 4:
        name += "-deal%d" % shape[4]
    return name
