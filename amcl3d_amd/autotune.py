"""Shape autotuner for the cloud-weight kernel.

A shape is (mapping, particles per wavefront):
  * mapping 0 (MAP_PARTICLE_LANES): each lane carries one particle and
    64 / ppw consecutive cloud points per step (ppw 64, 32 or 16);
  * mapping 1 (MAP_POINT_LANES): the wavefront's 64 lanes are 64 consecutive
    points, transformed for ppw particles through an LDS tile.
Every shape gives bit-identical results, so the choice is purely a timing
one: run the caller's workload once per candidate (after one untimed warm-up)
and keep the fastest.
"""
import time

from . import _lib

CANDIDATES = ((0, 64), (0, 32), (0, 16), (1, 16), (1, 32), (1, 64))


def set_shape(shape):
    _lib.set_mapping(shape[0])
    _lib.set_tile(shape[1])


def get_shape():
    return (_lib.get_mapping(), _lib.get_tile())


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
    return "map%d-ppw%d" % tuple(shape)
