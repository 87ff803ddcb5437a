"""Shape autotuner for the cloud-weight kernel.

A shape is (mapping, particles per wavefront, visiting order):
  * mapping 0 (MAP_PARTICLE_LANES): each lane carries one particle and
    64 / ppw consecutive cloud points per step (ppw 64, 32 or 16);
  * mapping 1 (MAP_POINT_LANES): the wavefront's 64 lanes are 64 consecutive
    points, transformed for ppw particles through an LDS tile;
  * order 0 visits particles in index order, order 1 in a spatial order
    computed on the device before each weight launch (per filter, also in
    fleets and served batches).
Every shape gives bit-identical results, so the choice is purely a timing
one: run the caller's workload once per candidate (after one untimed warm-up)
and keep the fastest.
"""
import time

from . import _lib

CANDIDATES = tuple((m, p, o) for o in (1, 0) for m, p in ((0, 64), (0, 32), (0, 16), (1, 16), (1, 32), (1, 64)))
FLEET_CANDIDATES = CANDIDATES


def set_shape(shape):
    _lib.set_mapping(shape[0])
    _lib.set_tile(shape[1])
    if len(shape) > 2:
        _lib.set_order(shape[2])


def get_shape():
    return (_lib.get_mapping(), _lib.get_tile(), _lib.get_order())


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
    return name
