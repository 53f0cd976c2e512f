"""GPU depression filling (fs_fill_depressions, fs_set_depression_fill)
against the oracle's Priority-Flood+epsilon, bit for bit.

The oracle's fill is pinned by an independent fixed-point computation
(tests/test_fill.py).  The GPU relaxes the same fixed point tile by tile,
so the filled surfaces must be identical, on single- and multi-tile grids,
ragged sizes and the benchmark terrain; with the fill mode on, every time-
step is the reference's step applied to the filled surface, and no interior
cell is left without a receiver.
"""
import sys

import numpy as np
import pytest

import fastscape_amd as fa
from test_fill import pit_terrain

pytestmark = pytest.mark.gpu


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("W,H,kind", [(5, 5, "pits"), (64, 64, "pits"), (200, 150, "pits"),
                                      (130, 301, "random"), (257, 129, "pits"),
                                      (1000, 1000, "random")])
def test_fill_equals_priority_flood(oracle, W, H, kind):
    note("fill %dx%d %s" % (W, H, kind))
    h = pit_terrain(W, H, W + H) if kind == "pits" else oracle.init_terrain(W, H, 11)
    want = oracle.fill_depressions(h, W, H)
    with fa.FastScape(W, H, fa.preset(fa.PRESET_BENCHMARK), h=h) as m:
        m.fill_depressions()
        np.testing.assert_array_equal(m.getH().ravel(), want)
        m.fill_depressions()                       # idempotent
        np.testing.assert_array_equal(m.getH().ravel(), want)


@pytest.mark.parametrize("preset", [fa.PRESET_BENCHMARK, fa.PRESET_REFERENCE])
def test_fill_mode_steps(oracle, preset):
    """Fill, then the reference's time-step, three times: state-synchronised
    against the oracle (receivers, accumulation, levels, elevations)."""
    W, H = 160, 120
    op = oracle.preset(preset)
    h = pit_terrain(W, H, 5)
    note("fill mode, preset %d" % preset)
    with fa.FastScape(W, H, fa.preset(preset), h=h) as m:
        m.set_depression_fill(True)
        for s in range(3):
            F = oracle.fill_depressions(h, W, H)
            st = oracle.step(F, W, H, op, oracle.POW_EXACT)
            m.setH(h)
            m.step(1)
            rec = m.receivers()
            np.testing.assert_array_equal(rec, st["rec"], err_msg="rec step %d" % s)
            assert np.all(rec.reshape(H, W)[2:H - 2, 2:W - 2] >= 0)
            np.testing.assert_array_equal(m.accum(), st["accum"], err_msg="accum %d" % s)
            np.testing.assert_array_equal(m.levels(), st["levels"], err_msg="levels %d" % s)
            np.testing.assert_array_equal(m.getH().ravel(), st["h"], err_msg="h step %d" % s)
            h = st["h"]
        m.set_depression_fill(False)
        m.setH(h)
        m.step(1)
        np.testing.assert_array_equal(m.getH().ravel(),
                                      oracle.step(h, W, H, op, oracle.POW_EXACT)["h"])
