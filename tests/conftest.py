import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def kat():
    """The reference's test data, decoded by tests/golden/make_fixtures.py."""
    meta = json.load(open(os.path.join(GOLDEN, "kat_meta.json")))
    meta["map_cloud"] = np.load(os.path.join(GOLDEN, "kat_map_cloud.npz"))["xyz"]
    meta["sensor_cloud"] = np.load(os.path.join(GOLDEN, "kat_sensor_cloud.npz"))["xyz"].astype(np.float32)
    meta["nav"] = np.load(os.path.join(GOLDEN, "kat_nav_slice.npz"))["data"]
    return meta


@pytest.fixture(scope="session")
def kat_oracle_grid(kat):
    """Oracle computeGrid of the test map: the cloud is shifted by the map minimum
    (mappointcloud_msg.bin), so cells are sampled at i*res from zero."""
    import oracle

    ext = np.asarray(kat["map_max"]) - np.asarray(kat["map_min"])
    dist, prob = oracle.compute_grid(kat["map_cloud"], np.zeros(3), ext, kat["map_res"], kat["sensor_dev"])
    return dist, prob


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a GPU; tests marked gpu use it."""
    import amcl3d_amd

    amcl3d_amd.load()
    return amcl3d_amd
