"""GPU: served sessions against the oracle, at the serving configuration's
scale.

The four resident maps of the serving configuration, built on the device:
the 20 x 20 x 6 m room at 0.1 m (200 x 200 x 60 cells), buildings of
50 x 50 x 10 m at 0.05 m (1000 x 1000 x 200), 100 x 100 x 30 m at 0.05 m
(2000 x 2000 x 600, 9.6 GB) and 200 x 200 x 20 m at 0.1 m (2000 x 2000 x 200):
13.6 GB of prob.

* Served sessions, on the room and on the upper storeys of the
  2000 x 2000 x 600 map, replayed step by step through the oracle's init,
  predict, update and resample (ParticleFilter.cpp:46-222) with the
  session's draws (a3d_session_draws): bit-exact.
* 1,024 sessions on the four maps (N log-uniform in 1e3..5e4, M uniform in
  2e3..2e4, two requests each) served together by continuous batching; a
  seeded sample of 16 sessions served again alone gives the same bits, and
  the sample's small sessions on the two smaller maps also match the oracle.
"""
import sys
import time

import numpy as np
import pytest

import oracle
from amcl3d_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32
DEVS = [0.5, 0.5, 0.1, 0.2]
MODS = [0.1, 0.1, 0.1, 0.3]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c4_maps(gpu):
    """[(BoxMap, Grid3d, sensor heights)] of the four serving maps."""
    specs = [("room", None), ((50.0, 50.0, 10.0), 0.05, 150, 11), ((100.0, 100.0, 30.0), 0.05, 500, 7),
             ((200.0, 200.0, 20.0), 0.1, 800, 13)]
    maps = []
    for spec in specs:
        t0 = time.time()
        if spec[0] == "room":
            b = W.room_map(1)
            g = gpu.Grid3d()
            g.open(b.points(), b.min, b.max, b.res, 0.1)
            zs = (1.5,)
        else:
            ext, res, nbox, seed = spec
            b = W.building_map(seed, extent=ext, res=res, n_boxes=nbox)
            g = gpu.Grid3d()
            g.open(b.points(dedup=False), b.min, b.max, b.res, 0.05)
            nfl = int(round(ext[2] / 3.0))
            zs = tuple(3.0 * k + 1.5 for k in range(nfl))
        _say("serving map %s: %d cells (%.1fs)" % (spec[0], g.cells, time.time() - t0))
        maps.append((b, g, zs))
    return maps


def _oracle_grid(g):
    return oracle.GridMeta(g.min, g.max, g.res, g.read_cells())


def _pose(b, z, rng):
    view = W.storey_view(b, z, 3.0) if b.max[2] > 6.5 else b
    p = W.free_pose(view, rng, z_range=(z, z), clearance=1.0)
    return np.append(p, rng.uniform(-np.pi, np.pi)).astype(F32), view


def _replay_oracle(og, n, pose, seed, steps):
    """The session's steps through the oracle chain with its draws."""
    from amcl3d_amd.serving import session_draws

    z0, _ = session_draws(seed, 0, n)
    p, _ = oracle.pf_init(n, pose, DEVS, z0)
    for k, (cloud, rg, d, rs) in enumerate(steps):
        z4, u01 = session_draws(seed, k + 1, n)
        p = oracle.pf_predict(p, MODS, d, z4)
        p, _ = oracle.pf_update(p, og, cloud, rg, 0.8, 0.1)
        if rs:
            p, _ = oracle.pf_resample(p, u01, 0)
    return p


def _serve(srv, mids, sessions):
    """Opens the sessions, submits every step of every session in step order
    and waits; returns the session ids."""
    ids = [srv.open_session(mids[k], n, pose, DEVS, seed) for (k, n, pose, seed, _) in sessions]
    nsteps = max(len(s[4]) for s in sessions)
    for j in range(nsteps):
        for sid, (_, _, _, _, steps) in zip(ids, sessions):
            if j < len(steps):
                cloud, rg, d, rs = steps[j]
                srv.submit(sid, MODS, d, tuple(np.ascontiguousarray(cloud[:, a]) for a in range(3)), rg, 0.8, 0.1,
                           resample=rs)
    srv.drain()
    return ids


def test_served_sessions_replay_through_oracle(gpu, c4_maps):
    """Three sessions on the room map and two on the top storeys of the
    2000 x 2000 x 600 map (cells past index 2^31), served together, each equal
    bit for bit to the oracle chain driven with the session's draws."""
    from amcl3d_amd.serving import Server

    rng = np.random.default_rng(41)
    plan = [(0, 700, 1.5, 1500), (0, 2000, 1.5, 400), (0, 3500, 1.5, 2600), (2, 1500, 28.5, 5000),
            (2, 2500, 25.5, 3000)]
    sessions = []
    for i, (k, n, z, m) in enumerate(plan):
        b = c4_maps[k][0]
        pose, view = _pose(b, z, rng)
        steps = []
        for j in range(3):
            cloud = W.sensor_cloud(view, pose, m, 0.05, rng)
            steps.append((cloud, W.beacons(view, pose, 3, 0.1, rng), rng.uniform(-0.1, 0.1, 4), j != 1))
        sessions.append((k, n, pose, 5000 + i, steps))
    srv = Server(max_inflight=3, max_batch_sessions=64)
    try:
        mids = [srv.add_map(g) for _, g, _ in c4_maps]
        ids = _serve(srv, mids, sessions)
        got = [srv.particles(sid) for sid in ids]
    finally:
        srv.close()
    ogs = {}
    for (k, n, pose, seed, steps), g in zip(sessions, got):
        if k not in ogs:
            ogs[k] = _oracle_grid(c4_maps[k][1])
        _say("oracle replay: map %d, %d particles at z=%.1f" % (k, n, pose[2]))
        op = _replay_oracle(ogs[k], n, pose, seed, steps)
        assert np.array_equal(g.view(np.int32), op.view(np.int32)), (k, n)


def test_serving_scale_same_alone_and_in_the_mix(gpu, c4_maps):
    """1,024 sessions on the four maps, two requests each, served together;
    16 sampled sessions served again alone (one session per launch) give the
    same bits, and the sampled sessions of at most 6,000 particles on the two
    smaller maps equal the oracle chain."""
    from amcl3d_amd.serving import Server

    rng = np.random.default_rng(43)
    poses = []
    for b, g, zs in c4_maps:
        ps = []
        for _ in range(6):
            pose, view = _pose(b, zs[int(rng.integers(0, len(zs)))], rng)
            ps.append((pose, W.sensor_cloud(view, pose, 20000, 0.05, rng), W.beacons(view, pose, 3, 0.1, rng)))
        poses.append(ps)
    sessions = []
    for s in range(1024):
        k = s % 4
        n = int(np.round(np.exp(rng.uniform(np.log(1000), np.log(50000)))))
        m = int(rng.integers(2000, 20001))
        pose, cloud, rg = poses[k][int(rng.integers(0, 6))]
        steps = [(cloud[:m], rg, rng.uniform(-0.05, 0.05, 4), j == 1) for j in range(2)]
        sessions.append((k, n, pose, 9000 + s, steps))
    evals = sum(n * st[0][0].shape[0] * 2 for _, n, _, _, st in sessions)
    t0 = time.time()
    srv = Server(max_inflight=3, max_batch_sessions=4096, max_batch_evals=800_000_000)
    try:
        mids = [srv.add_map(g) for _, g, _ in c4_maps]
        ids = _serve(srv, mids, sessions)
        stats = srv.stats()
        sample = sorted(np.random.default_rng(44).choice(1024, 16, replace=False).tolist())
        mixed = {i: srv.particles(ids[i]) for i in sample}
    finally:
        srv.close()
    _say("served 1,024 sessions x 2 requests (%.3g evals) in %d batches, %.1fs" % (evals, stats["batches"],
                                                                                time.time() - t0))
    assert stats["requests"] == 2048 and stats["batches"] < 2048
    alone = Server(max_inflight=1, max_batch_sessions=1)
    try:
        mids = [alone.add_map(g) for _, g, _ in c4_maps]
        for i in sample:
            (sid,) = _serve(alone, mids, [sessions[i]])
            got = alone.particles(sid)
            assert np.array_equal(got.view(np.int32), mixed[i].view(np.int32)), i
            alone.close_session(sid)
    finally:
        alone.close()
    ogs = {}
    checked = 0
    for i in sample:
        k, n, pose, seed, steps = sessions[i]
        if k > 1 or n > 6000:
            continue
        if k not in ogs:
            ogs[k] = _oracle_grid(c4_maps[k][1])
        op = _replay_oracle(ogs[k], n, pose, seed, steps)
        assert np.array_equal(mixed[i].view(np.int32), op.view(np.int32)), i
        checked += 1
    _say("oracle replays of sampled sessions: %d" % checked)
