"""Seeded synthetic workloads: maps, clouds, beacons, particle sets.

There is no network and no recorded data here, so every workload is
procedural.  A map is a set of axis-aligned boxes (walls, floors, obstacles)
in metres; its occupied points are the centres of the voxels on the boxes'
surfaces, ``min + (i + 0.5) * res``, the points octomap's leaf iterator gives
(amcl3d/src/PointCloudTools.cpp:69-79).  Clouds are ray-cast analytically
from a true pose onto the boxes, expressed in the sensor frame (the frame
``computeCloudWeight`` rotates by yaw and translates,
amcl3d/src/Grid3d.cpp:174-176) with Gaussian noise per axis.  Ranges are
{r, ax, ay, az} rows (amcl3d/src/ParticleFilter.h:53-63).
"""
import numpy as np


class BoxMap:
    def __init__(self, mn, mx, res, boxes):
        self.min = np.asarray(mn, np.float64)
        self.max = np.asarray(mx, np.float64)
        self.res = float(res)
        self.boxes = np.asarray(boxes, np.float64).reshape(-1, 6)  # lo xyz, hi xyz

    @property
    def size(self):
        return np.ceil((self.max - self.min) / self.res).astype(np.int64)

    def points(self, dedup=True):
        """Centres of the voxels on the surface of every box, float32 (n, 3)."""
        res, mn, size = self.res, self.min, self.size
        chunks = []
        for b in self.boxes:
            lo = np.maximum(np.ceil((b[:3] - mn) / res - 0.5), 0).astype(np.int64)
            hi = np.minimum(np.floor((b[3:] - mn) / res - 0.5), size - 1).astype(np.int64)
            if np.any(hi < lo):
                continue
            ext = hi - lo + 1
            faces = []
            for a in range(3):
                o = [k for k in range(3) if k != a]
                g = np.stack(np.meshgrid(np.arange(lo[o[0]], hi[o[0]] + 1), np.arange(lo[o[1]], hi[o[1]] + 1),
                                         indexing="ij"), -1).reshape(-1, 2)
                for v in ({lo[a], hi[a]} if ext[a] > 1 else {lo[a]}):
                    f = np.empty((g.shape[0], 3), np.int64)
                    f[:, a] = v
                    f[:, o[0]] = g[:, 0]
                    f[:, o[1]] = g[:, 1]
                    faces.append(f)
            chunks.append(np.concatenate(faces))
        idx = np.concatenate(chunks) if chunks else np.zeros((0, 3), np.int64)
        if dedup and idx.shape[0]:
            lin = idx[:, 0] + size[0] * (idx[:, 1] + size[1] * idx[:, 2])
            _, keep = np.unique(lin, return_index=True)
            idx = idx[np.sort(keep)]
        return (mn + (idx + 0.5) * res).astype(np.float32)

    def inside_any(self, p, margin=0.0):
        p = np.asarray(p, np.float64).reshape(-1, 3)
        lo = self.boxes[:, :3] - margin
        hi = self.boxes[:, 3:] + margin
        return np.any(np.all((p[:, None, :] >= lo[None]) & (p[:, None, :] <= hi[None]), -1), -1)

    def raycast(self, origin, dirs, max_range=60.0):
        """Distance to the first box hit along each unit direction (inf if none)."""
        o = np.asarray(origin, np.float64)
        d = np.asarray(dirs, np.float64)
        t_best = np.full(d.shape[0], np.inf)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / d
        # only boxes that a ray of length max_range can reach
        lo = self.boxes[:, :3]
        hi = self.boxes[:, 3:]
        gap = np.maximum(np.maximum(lo - o, o - hi), 0.0)
        near = self.boxes[np.sum(gap * gap, 1) <= max_range * max_range]
        for c0 in range(0, near.shape[0], 64):
            bx = near[c0:c0 + 64]
            t1 = (bx[None, :, :3] - o) * inv[:, None, :]
            t2 = (bx[None, :, 3:] - o) * inv[:, None, :]
            tmin = np.nanmax(np.minimum(t1, t2), -1)
            tmax = np.nanmin(np.maximum(t1, t2), -1)
            hit = (tmax >= tmin) & (tmax > 1e-6)
            t = np.where(hit, np.where(tmin > 1e-6, tmin, np.inf), np.inf)
            t_best = np.minimum(t_best, t.min(1))
        t_best[t_best > max_range] = np.inf
        return t_best


def room_map(seed, n_boxes=30, extent=(20.0, 20.0, 6.0), res=0.1):
    """A room: 6 walls one voxel thick plus n_boxes boxes with sides U(0.2, 2) m."""
    rng = np.random.default_rng(seed)
    ex = np.asarray(extent, np.float64)
    t = res
    boxes = [
        [0, 0, 0, ex[0], ex[1], t], [0, 0, ex[2] - t, ex[0], ex[1], ex[2]],
        [0, 0, 0, t, ex[1], ex[2]], [ex[0] - t, 0, 0, ex[0], ex[1], ex[2]],
        [0, 0, 0, ex[0], t, ex[2]], [0, ex[1] - t, 0, ex[0], ex[1], ex[2]],
    ]
    for _ in range(n_boxes):
        side = rng.uniform(0.2, 2.0, 3)
        c = rng.uniform(side / 2, ex - side / 2)
        c[2] = side[2] / 2 + t  # resting on the floor
        boxes.append(list(c - side / 2) + list(c + side / 2))
    return BoxMap(np.zeros(3), ex, res, boxes)


def building_map(seed, extent=(100.0, 100.0, 30.0), res=0.05, floor_height=3.0, corridor=10.0, n_boxes=500):
    """A building: floor slabs, corridor walls with door gaps, n_boxes boxes."""
    rng = np.random.default_rng(seed)
    ex = np.asarray(extent, np.float64)
    t = res
    boxes = []
    nfl = int(round(ex[2] / floor_height))
    for k in range(nfl + 1):
        z = min(k * floor_height, ex[2] - t)
        boxes.append([0, 0, z, ex[0], ex[1], z + t])
    boxes += [[0, 0, 0, t, ex[1], ex[2]], [ex[0] - t, 0, 0, ex[0], ex[1], ex[2]],
              [0, 0, 0, ex[0], t, ex[2]], [0, ex[1] - t, 0, ex[0], ex[1], ex[2]]]
    for k in range(nfl):
        z0, z1 = k * floor_height + t, min((k + 1) * floor_height, ex[2] - t)
        for xw in np.arange(corridor, ex[0] - 1e-9, corridor):
            y = 0.0
            while y < ex[1]:  # wall segments with 1 m door gaps
                seg = rng.uniform(4.0, 9.0)
                boxes.append([xw, y, z0, xw + t, min(y + seg, ex[1]), z1])
                y += seg + 1.0
    for _ in range(n_boxes):
        side = rng.uniform(0.3, 2.0, 3)
        k = rng.integers(0, nfl)
        c = rng.uniform(side / 2, ex - side / 2)
        c[2] = k * floor_height + t + side[2] / 2
        boxes.append(list(c - side / 2) + list(c + side / 2))
    return BoxMap(np.zeros(3), ex, res, boxes)


def city_map(seed, extent=(200.0, 200.0, 40.0), res=0.05, n_boxes=85, footprint=(4.0, 16.0), height=(3.0, 35.0)):
    """A city block: a ground slab and n_boxes buildings (boxes) with footprint
    sides U(footprint) m and heights U(height) m, standing on the ground."""
    rng = np.random.default_rng(seed)
    ex = np.asarray(extent, np.float64)
    boxes = [[0, 0, 0, ex[0], ex[1], res]]
    for _ in range(n_boxes):
        w = rng.uniform(*footprint, 2)
        h = min(rng.uniform(*height), ex[2] - 2 * res)
        c = rng.uniform(w / 2, ex[:2] - w / 2)
        boxes.append([c[0] - w[0] / 2, c[1] - w[1] / 2, res, c[0] + w[0] / 2, c[1] + w[1] / 2, res + h])
    return BoxMap(np.zeros(3), ex, res, boxes)


def storey_view(bmap, z, floor_height):
    """The boxes a sensor at height z sees: its storey's floor, ceiling, walls, obstacles."""
    k = np.floor(z / floor_height)
    z0, z1 = k * floor_height - 0.01, (k + 1) * floor_height + 0.01
    keep = (bmap.boxes[:, 5] >= z0) & (bmap.boxes[:, 2] <= z1)
    return BoxMap(bmap.min, bmap.max, bmap.res, bmap.boxes[keep])


def trajectory(bmap, steps, seed, z=1.5, speed=0.25, floor_height=3.0):
    """A smooth seeded path in free space: (steps+1, 4) poses x, y, z, yaw, moving
    about `speed` m and turning a few degrees per step."""
    rng = np.random.default_rng(seed)
    view = storey_view(bmap, z, floor_height)
    for _ in range(200):
        start = free_pose(view, rng, z_range=(z, z), clearance=1.0)
        poses = [np.append(start, rng.uniform(-np.pi, np.pi))]
        turn = rng.uniform(-0.05, 0.05)
        ok = True
        for _k in range(steps):
            x, y, zz, a = poses[-1]
            for _try in range(20):
                a2 = a + turn
                nxt = np.array([x + speed * np.cos(a2), y + speed * np.sin(a2), zz, a2])
                if not view.inside_any(nxt[:3], 0.8)[0] and np.all(nxt[:2] > bmap.min[:2] + 1) and \
                        np.all(nxt[:2] < bmap.max[:2] - 1):
                    break
                turn = rng.uniform(-0.6, 0.6)
            else:
                ok = False
                break
            poses.append(nxt)
            turn = 0.8 * turn + rng.uniform(-0.02, 0.02)
        if ok:
            return np.array(poses)
    raise RuntimeError("no free trajectory found")


def odom_delta(p0, p1):
    """Odometry increment from pose p0 to p1 in p0's frame (Node.cpp:147-151)."""
    c, s = np.cos(p0[3]), np.sin(p0[3])
    dx, dy = p1[0] - p0[0], p1[1] - p0[1]
    return np.array([c * dx + s * dy, -s * dx + c * dy, p1[2] - p0[2], p1[3] - p0[3]])


def free_pose(bmap, rng, z_range=None, clearance=0.6):
    for _ in range(10000):
        p = rng.uniform(bmap.min + 1.0, bmap.max - 1.0)
        if z_range is not None:
            p[2] = rng.uniform(*z_range)
        if not bmap.inside_any(p, clearance)[0]:
            return p
    raise RuntimeError("no free pose found")


def scan_dirs(m, elev=(-0.5, 0.5)):
    """m unit directions in scan order: rows of constant elevation, azimuth sweeping."""
    rows = max(1, int(round(np.sqrt(m / 8.0))))
    cols = int(np.ceil(m / rows))
    el = np.linspace(elev[0], elev[1], rows)
    az = np.linspace(-np.pi, np.pi, cols, endpoint=False)
    E, A = np.meshgrid(el, az, indexing="ij")
    d = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
    return d[:m]


def sensor_cloud(bmap, pose, m, noise, rng, max_range=30.0):
    """m points ray-cast from pose (x, y, z, yaw), in the sensor frame, plus N(0, noise) per axis."""
    p = np.asarray(pose[:3], np.float64)
    yaw = float(pose[3])
    dirs = scan_dirs(int(m * 1.5) + 16)
    t = bmap.raycast(p, dirs, max_range)
    ok = np.isfinite(t)
    hits = dirs[ok] * t[ok, None]
    if hits.shape[0] < m:  # repeat rays to reach m
        hits = hits[np.resize(np.arange(hits.shape[0]), m)]
    hits = hits[:m]
    c, s = np.cos(yaw), np.sin(yaw)
    local = np.stack([c * hits[:, 0] + s * hits[:, 1], -s * hits[:, 0] + c * hits[:, 1], hits[:, 2]], -1)
    local += rng.normal(0.0, noise, local.shape)
    return local.astype(np.float32)


def beacons(bmap, pose, nb, sigma, rng):
    """nb range rows {r, ax, ay, az}: r = |true - anchor| + N(0, sigma)."""
    a = rng.uniform(bmap.min, bmap.max, (nb, 3))
    r = np.linalg.norm(a - np.asarray(pose[:3]), axis=1) + rng.normal(0, sigma, nb)
    return np.concatenate([r[:, None], a], 1).astype(np.float32)


def particles_around(pose, n, rng, dxyz=1.0, dyaw=0.3):
    """(n, 7) particles uniform in a box of +-dxyz m and +-dyaw rad around pose, w = 1/n."""
    p = np.zeros((n, 7), np.float32)
    p[:, :3] = np.asarray(pose[:3]) + rng.uniform(-dxyz, dxyz, (n, 3))
    p[:, 3] = pose[3] + rng.uniform(-dyaw, dyaw, n)
    p[:, 4] = 1.0 / n
    return p


class Step:
    """One predict -> update -> resample step of one filter, with its draws."""

    def __init__(self, cloud, ranges, mods, deltas, z4, u01, alpha, sigma):
        self.cloud, self.ranges, self.mods, self.deltas = cloud, ranges, mods, deltas
        self.z4, self.u01, self.alpha, self.sigma = z4, u01, alpha, sigma


def room_case(seed, n=1000, m=2000, nb=3, cloud_noise=0.05, range_sigma=0.1, alpha=0.8, sensor_dev=0.1):
    """The room workload: 200x200x60 map at 0.1 m, one filter and one step."""
    rng = np.random.default_rng(1000 + seed)
    bmap = room_map(seed)
    pose = np.append(free_pose(bmap, rng, z_range=(1.0, 2.5)), rng.uniform(-np.pi, np.pi))
    cloud = sensor_cloud(bmap, pose, m, cloud_noise, rng)
    ranges = beacons(bmap, pose, nb, range_sigma, rng)
    parts = particles_around(pose, n, rng)
    step = Step(cloud, ranges, np.array([0.1, 0.1, 0.1, 0.3]), np.array([0.05, -0.03, 0.01, 0.02]),
                rng.standard_normal((n, 4)).astype(np.float32), float(rng.uniform()), alpha, range_sigma)
    return {"map": bmap, "pose": pose, "particles": parts, "step": step, "sensor_dev": sensor_dev}
