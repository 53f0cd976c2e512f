import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "barnes2019-landscape_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP device 0)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_dem_shas():
    with open(os.path.join(GOLDEN, "dem_sha256.json")) as f:
        return json.load(f)


STAGE_FIXTURES = sorted(f for f in os.listdir(GOLDEN) if f.startswith("stages_"))
FINAL_FIXTURES = sorted(f for f in os.listdir(GOLDEN) if f.startswith("final_"))


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


def params_for(o, preset):
    return o.preset(int(preset))


def cone_terrain(W, H, period, offset):
    """Conical basins `period` cells wide: each cell drains to its basin's
    centre, up to period/2 receiver hops away (48-cell basins: 25 levels)."""
    x = np.arange(W, dtype=np.float64)[None, :]
    y = np.arange(H, dtype=np.float64)[:, None]
    cx = (x + offset) % period - period / 2 + 0.5
    cy = (y + offset) % period - period / 2 + 0.5
    h = 1.0 + np.sqrt(cx * cx + cy * cy) + 1e-7 * ((x * 7 + y * 13) % 11)
    h[:2, :] = h[-2:, :] = 0
    h[:, :2] = h[:, -2:] = 0
    return np.ascontiguousarray(h).ravel()
