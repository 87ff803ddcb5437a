"""Decode the reference's test data files into compact fixtures.

Reads the ROS1-serialised messages that amcl3d's own tests load
(amcl3d/tests/data/*.bin, Apache-2.0, FADA-CATEC) and writes the data they hold
as numpy arrays.  Nothing else is taken from the reference: the fixtures are
inputs and expected outputs only.

  grid_info.bin          geometry_msgs/PoseArray, 25,809 poses; the positions
                         are the sensor cloud of Grid3dTest.computeCloudWeightTest
                         (amcl3d/tests/Grid3dTest.cpp:135-157)
  mappointcloud_msg.bin  sensor_msgs/PointCloud2, 179,551 occupied-leaf centres of
                         the test map, shifted by the map minimum
                         (Grid3dTest.cpp:94-104, PointCloudToolsTest.cpp:166-173)
  nav_msg.bin            nav_msgs/OccupancyGrid 522x384, one z-slice of the
                         likelihood grid scaled to int8 0..100
                         (Grid3dTest.cpp:56-66, Grid3d.cpp:100-118)
  particle_info.bin      PoseArray with zero poses (Grid3dTest.cpp:215-226)

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import struct

import numpy as np

SRC = "/root/reference/amcl3d/tests/data"
OUT = os.path.dirname(os.path.abspath(__file__))


def _header(b, off):
    _seq, _sec, _nsec, ln = struct.unpack_from("<IIII", b, off)
    off += 16
    return b[off:off + ln].decode(), off + ln


def pose_array(path):
    b = open(path, "rb").read()
    _frame, off = _header(b, 0)
    (n,) = struct.unpack_from("<I", b, off)
    off += 4
    a = np.frombuffer(b, dtype="<f8", count=n * 7, offset=off).reshape(n, 7)
    assert off + n * 56 == len(b)
    return a


def point_cloud2(path):
    b = open(path, "rb").read()
    frame, off = _header(b, 0)
    h, w = struct.unpack_from("<II", b, off)
    off += 8
    (nf,) = struct.unpack_from("<I", b, off)
    off += 4
    fields = {}
    for _ in range(nf):
        (ln,) = struct.unpack_from("<I", b, off)
        off += 4
        name = b[off:off + ln].decode()
        off += ln
        o, dt, c = struct.unpack_from("<IBI", b, off)
        off += 9
        fields[name] = (o, dt, c)
    big_endian = b[off]
    off += 1
    point_step, row_step = struct.unpack_from("<II", b, off)
    off += 8
    (dl,) = struct.unpack_from("<I", b, off)
    off += 4
    raw = np.frombuffer(b, dtype=np.uint8, count=dl, offset=off).reshape(h * w, point_step)
    off += dl
    assert off + 1 == len(b) and not big_endian
    assert all(fields[k][1] == 7 for k in "xyz")  # FLOAT32
    xyz = np.stack([raw[:, fields[k][0]:fields[k][0] + 4].copy().view("<f4")[:, 0] for k in "xyz"], 1)
    return frame, xyz.astype(np.float32)


def occupancy_grid(path):
    b = open(path, "rb").read()
    frame, off = _header(b, 0)
    off += 8  # map_load_time
    res, width, height = struct.unpack_from("<fII", b, off)
    off += 12
    origin = struct.unpack_from("<7d", b, off)
    off += 56
    (dl,) = struct.unpack_from("<I", b, off)
    off += 4
    data = np.frombuffer(b, dtype=np.int8, count=dl, offset=off).reshape(height, width)
    assert off + dl == len(b)
    return frame, res, origin, data.copy()


def main():
    sensor = pose_array(os.path.join(SRC, "grid_info.bin"))
    assert not sensor[:, 3:].any()
    particles = pose_array(os.path.join(SRC, "particle_info.bin"))
    _, map_cloud = point_cloud2(os.path.join(SRC, "mappointcloud_msg.bin"))
    _, res, origin, nav = occupancy_grid(os.path.join(SRC, "nav_msg.bin"))
    np.savez_compressed(os.path.join(OUT, "kat_map_cloud.npz"), xyz=map_cloud)
    np.savez_compressed(os.path.join(OUT, "kat_sensor_cloud.npz"), xyz=sensor[:, :3])
    np.savez_compressed(os.path.join(OUT, "kat_nav_slice.npz"), data=nav)
    meta = {
        "source": "amcl3d/tests/data (decoded by tests/golden/make_fixtures.py)",
        "map_min": [-17.35, -9.5, -1.4],
        "map_max": [8.75, 9.7, 6.25],
        "map_res": 0.05,
        "map_points": int(map_cloud.shape[0]),
        "map_bounds_cite": "amcl3d/tests/PointCloudToolsTest.cpp:166-173",
        "sensor_dev": 0.05,
        "sensor_dev_cite": "amcl3d/tests/Grid3dTest.cpp:7",
        "kat_pose": {"tx": 20.017967, "ty": 10.140815, "tz": 3.372801, "roll": 0.0, "pitch": 0.0,
                     "yaw": 0.166781},
        "kat_weight": 3.8109049797058105,
        "kat_tolerance": 1e-4,
        "kat_cite": "amcl3d/tests/Grid3dTest.cpp:128-132,160-168",
        "nav_res": res,
        "nav_origin": origin,
        "nav_shape": list(nav.shape),
        "particle_info_poses": int(particles.shape[0]),
        "isinto_in": [1.0, 1.0, 1.0],
        "isinto_out": [-100.0, -100.0, -100.0],
        "isinto_cite": "amcl3d/tests/Grid3dTest.cpp:247-265",
    }
    with open(os.path.join(OUT, "kat_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote fixtures:", map_cloud.shape, sensor.shape, nav.shape)


if __name__ == "__main__":
    main()
