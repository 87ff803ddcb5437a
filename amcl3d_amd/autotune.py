"""Tile autotuner for the cloud-weight kernel.

The kernel's tile is the number of particles one wavefront carries through
the cloud (16, 32 or 64; the point tile is the wavefront's 64 lanes).  Every
tile gives bit-identical results, so the choice is purely a timing one: run
the caller's workload once per candidate (after one untimed warm-up) and
keep the fastest.
"""
import time

from . import _lib

CANDIDATES = (16, 32, 64)


def tune(run, candidates=CANDIDATES, reps=2, prepare=None):
    """run(): executes and synchronizes the workload once.  prepare(tile):
    optional re-staging after the tile changes.  Returns (best, {tile: s})."""
    timings = {}
    for t in candidates:
        _lib.set_tile(t)
        if prepare:
            prepare(t)
        run()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            best = min(best, time.perf_counter() - t0)
        timings[t] = best
    choice = min(timings, key=timings.get)
    _lib.set_tile(choice)
    if prepare:
        prepare(choice)
    return choice, timings
