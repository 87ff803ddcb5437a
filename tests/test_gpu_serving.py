"""GPU: served sessions give the same result alone and in the mix, and the
same as the plain ParticleFilter driven with the session's draws."""
import numpy as np
import pytest

from amcl3d_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def room_grid(gpu):
    bm = W.room_map(1)
    g = gpu.Grid3d()
    g.open(bm.points(), bm.min, bm.max, bm.res, 0.1)
    return bm, g


def _sessions(bm, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(k):
        n = int(rng.integers(50, 3000))
        m = int(rng.integers(20, 2500))
        pose = np.append(W.free_pose(bm, rng, z_range=(1.0, 2.5)), rng.uniform(-3, 3)).astype(F32)
        steps = []
        for _ in range(3):
            c = W.sensor_cloud(bm, pose, m, 0.05, rng)
            steps.append((tuple(np.ascontiguousarray(c[:, j]) for j in range(3)), W.beacons(bm, pose, 3, 0.1, rng),
                          rng.uniform(-0.1, 0.1, 4), bool(rng.integers(0, 2))))
        out.append((n, pose, 1000 + s, steps))
    return out


def _run(gpu, grid, sessions, alone, maps_of=None, max_batch_evals=None):
    """grid: one grid, or a list of grids with maps_of[i] the map of session i."""
    from amcl3d_amd.serving import Server

    kw = {} if max_batch_evals is None else {"max_batch_evals": max_batch_evals}
    srv = Server(max_inflight=1 if alone else 3, max_batch_sessions=1 if alone else 64, **kw)
    grids = grid if isinstance(grid, list) else [grid]
    mids = [srv.add_map(g) for g in grids]
    maps_of = maps_of or [0] * len(sessions)
    ids = [srv.open_session(mids[maps_of[i]], n, pose, [0.3, 0.3, 0.1, 0.2], seed)
           for i, (n, pose, seed, _) in enumerate(sessions)]
    mods = [0.1, 0.1, 0.1, 0.3]
    for k in range(3):
        for sid, (_, _, _, steps) in zip(ids, sessions):
            xyz, rg, d, rs = steps[k]
            srv.submit(sid, mods, d, xyz, rg, 0.8, 0.1, resample=rs)
            if alone:
                srv.drain()
    srv.drain()
    res = srv.poll(100000)
    out = [srv.particles(sid) for sid in ids]
    st = srv.stats()
    srv.close()
    return out, res, st


def test_same_alone_and_in_the_mix(gpu, room_grid):
    bm, grid = room_grid
    sessions = _sessions(bm, 10, 4)
    alone, res_a, _ = _run(gpu, grid, sessions, alone=True)
    mixed, res_m, st = _run(gpu, grid, sessions, alone=False)
    assert len(res_a) == len(res_m) == 30
    assert st["batches"] < 30  # requests were batched together
    for a, b in zip(alone, mixed):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    ma = {(r.session, r.ticket % 1000000): np.array(r.mean) for r in res_a}
    assert all(np.isfinite(r.latency_ms) and r.latency_ms >= 0 for r in res_m)
    assert len(ma) == 30


def test_backlog_grows_batches_same_results(gpu, room_grid):
    """A bound of about one session's evaluations per batch: with every
    session's request queued at once the backlog lets batches grow past it
    (fewer batches than requests), and every session still gives its alone
    result."""
    bm, grid = room_grid
    sessions = _sessions(bm, 12, 31)
    bound = int(np.median([n * len(st[0][0][0]) for n, _, _, st in sessions]))
    alone, _, _ = _run(gpu, grid, sessions, alone=True)
    grown, res, st = _run(gpu, grid, sessions, alone=False, max_batch_evals=bound)
    assert len(res) == 36
    assert st["batches"] < 36
    for a, b in zip(alone, grown):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))


def test_served_equals_plain_filter(gpu, room_grid):
    from amcl3d_amd.serving import session_draws

    bm, grid = room_grid
    sessions = _sessions(bm, 3, 9)
    served, _, _ = _run(gpu, grid, sessions, alone=False)
    mods = [0.1, 0.1, 0.1, 0.3]
    for (n, pose, seed, steps), got in zip(sessions, served):
        pf = gpu.ParticleFilter()
        z0, _ = session_draws(seed, 0, n)
        pf.init(n, *pose, 0.3, 0.3, 0.1, 0.2, z0)
        for k, (xyz, rg, d, rs) in enumerate(steps):
            z4, u01 = session_draws(seed, k + 1, n)
            pf.predict(*mods, *d, z4)
            pf.update(grid, np.stack(xyz, 1), rg, 0.8, 0.1)
            if rs:
                pf.resample(u01)
        assert np.array_equal(pf.particles().view(np.int32), got.view(np.int32))


def test_mixed_maps_in_one_batch(gpu, room_grid):
    """Sessions on two different maps share batches (one launch, each filter
    on its own grid) and still give their alone results and the plain
    filter's."""
    from amcl3d_amd.serving import session_draws

    bm, grid = room_grid
    bm2 = W.room_map(2, extent=(12.0, 16.0, 5.0), res=0.08)
    g2 = gpu.Grid3d()
    g2.open(bm2.points(), bm2.min, bm2.max, bm2.res, 0.15)
    sessions = _sessions(bm, 4, 21) + _sessions(bm2, 4, 22)
    maps_of = [0] * 4 + [1] * 4
    order = [0, 4, 1, 5, 2, 6, 3, 7]  # interleave the maps
    sessions = [sessions[i] for i in order]
    maps_of = [maps_of[i] for i in order]
    alone, _, _ = _run(gpu, [grid, g2], sessions, alone=True, maps_of=maps_of)
    mixed, _, st = _run(gpu, [grid, g2], sessions, alone=False, maps_of=maps_of)
    assert st["batches"] < 24
    for a, b in zip(alone, mixed):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    # a session on the second map against the plain filter
    n, pose, seed, steps = sessions[1]
    pf = gpu.ParticleFilter()
    z0, _ = session_draws(seed, 0, n)
    pf.init(n, *pose, 0.3, 0.3, 0.1, 0.2, z0)
    for k, (xyz, rg, d, rs) in enumerate(steps):
        z4, u01 = session_draws(seed, k + 1, n)
        pf.predict(0.1, 0.1, 0.1, 0.3, *d, z4)
        pf.update(g2, np.stack(xyz, 1), rg, 0.8, 0.1)
        if rs:
            pf.resample(u01)
    assert np.array_equal(pf.particles().view(np.int32), mixed[1].view(np.int32))


def test_latest_only_supersedes_and_composes(gpu, room_grid):
    """latest_only: a queued request is replaced by the session's newer one
    (the reference's scan subscriber keeps one message, Node.cpp:60) and the
    two odometry increments compose; the session equals the plain filter
    driven with the increments that ran."""
    from amcl3d_amd.serving import Server, session_draws

    bm, grid = room_grid
    sessions = _sessions(bm, 2, 31)
    # a large first session keeps the GPU busy so the small one's requests queue up
    n0, pose0, seed0, steps0 = sessions[0]
    n1, pose1, seed1, steps1 = sessions[1]
    srv = Server(max_inflight=1, max_batch_sessions=1, latest_only=True)
    mid = srv.add_map(grid)
    s0 = srv.open_session(mid, 60000, pose0, [0.3, 0.3, 0.1, 0.2], seed0)
    s1 = srv.open_session(mid, n1, pose1, [0.3, 0.3, 0.1, 0.2], seed1)
    big = W.sensor_cloud(bm, pose0, 20000, 0.05, np.random.default_rng(3))
    bigxyz = tuple(np.ascontiguousarray(big[:, j]) for j in range(3))
    mods = [0.1, 0.1, 0.1, 0.3]
    rng = np.random.default_rng(8)
    sent = {}
    for k in range(9):
        srv.submit(s0, mods, [0.01, 0.0, 0.0, 0.0], bigxyz, None, 0.8, 0.1)
        xyz, rg, _, _ = steps1[k % 3]
        d = rng.uniform(-0.1, 0.1, 4)
        rs = bool(k % 2)
        t = srv.submit(s1, mods, d, xyz, rg, 0.8, 0.1, resample=rs)
        sent[t] = (xyz, rg, rs)
    srv.drain()
    res = sorted((r for r in srv.poll(1000) if r.session == s1), key=lambda r: r.ticket)
    got = srv.particles(s1)
    srv.close()
    assert len(res) == 9
    ran = [r for r in res if not r.superseded]
    print("latest_only: %d of 9 requests superseded" % (9 - len(ran)), flush=True)
    assert ran and ran[-1].ticket == max(sent)  # the newest request always runs
    pf = gpu.ParticleFilter()
    z0, _ = session_draws(seed1, 0, n1)
    pf.init(n1, *pose1, 0.3, 0.3, 0.1, 0.2, z0)
    for k, r in enumerate(ran):
        xyz, rg, rs = sent[r.ticket]
        z4, u01 = session_draws(seed1, k + 1, n1)
        pf.predict(*mods, *list(r.deltas), z4)
        pf.update(grid, np.stack(xyz, 1), rg, 0.8, 0.1)
        if rs:
            pf.resample(u01)
    assert np.array_equal(pf.particles().view(np.int32), got.view(np.int32))


def test_session_from_particles(gpu, room_grid):
    """A session opened from given particles (global localisation) equals the
    plain filter started from the same particles."""
    from amcl3d_amd.serving import Server, session_draws

    bm, grid = room_grid
    rng = np.random.default_rng(12)
    n = 4000
    parts = np.zeros((n, 7), F32)
    parts[:, :3] = rng.uniform(bm.min + 0.5, bm.max - 0.5, (n, 3))
    parts[:, 3] = rng.uniform(-np.pi, np.pi, n)
    parts[:, 4] = 1.0 / n
    (_, pose, _, steps), = _sessions(bm, 1, 13)
    srv = Server(max_inflight=2)
    mid = srv.add_map(grid)
    sid = srv.open_session_particles(mid, parts, 77)
    for xyz, rg, d, rs in steps:
        srv.submit(sid, [0.1, 0.1, 0.1, 0.3], d, xyz, rg, 0.8, 0.1, resample=rs)
    srv.drain()
    got = srv.particles(sid)
    srv.close()
    pf = gpu.ParticleFilter()
    pf.set_particles(parts)
    for k, (xyz, rg, d, rs) in enumerate(steps):
        z4, u01 = session_draws(77, k + 1, n)
        pf.predict(0.1, 0.1, 0.1, 0.3, *d, z4)
        pf.update(grid, np.stack(xyz, 1), rg, 0.8, 0.1)
        if rs:
            pf.resample(u01)
    assert np.array_equal(pf.particles().view(np.int32), got.view(np.int32))


def test_admission_by_hbm_budget(gpu, room_grid):
    from amcl3d_amd.serving import Server

    _, grid = room_grid
    # a session of 10,000 particles holds ~0.8 MB (state, scratch, flags)
    srv = Server(hbm_budget_bytes=1 << 20)
    mid = srv.add_map(grid)
    srv.open_session(mid, 10000, [10, 10, 1.5, 0], [0.1, 0.1, 0.1, 0.1], 1)
    with pytest.raises(gpu.Amcl3dError) as e:
        srv.open_session(mid, 10000, [10, 10, 1.5, 0], [0.1, 0.1, 0.1, 0.1], 2)
    assert e.value.code == 4  # A3D_ERR_NOMEM
    srv.close()
