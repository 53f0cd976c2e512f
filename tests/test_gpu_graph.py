"""GPU parity of the tile-graph path (drainage of any depth, per 64x64 tile
with no halo, tiles joined by the graph of the cells where drainage leaves a
tile) against the oracle and the level path, through the C ABI.

Bar: receivers, donors, the stack, levels[] and accumulation bit for bit;
elevations bit for bit for n = 1 and against the x*x oracle for n = 2 (and
within 1e-12 of glibc's pow), each step state-synchronised.  The terrains
put drainage paths across many tile borders (ramp, serpentine channel),
thousands of hops inside one tile (a serpentine channel in a 60x60 grid),
into partial tiles at the grid's edge, and through the reference preset's
first hundred steps (150 levels)."""
import numpy as np
import pytest

from conftest import cone_terrain

import fastscape_amd as fa
from test_gpu_parity import (assert_h, compare_step, gpu_params, note, oracle_params,
                             ramp_terrain, snake_terrain)

pytestmark = pytest.mark.gpu


def synced_steps(oracle, W, H, h, preset, steps, path, what):
    """State-synchronised steps on `path` against the oracle; returns the
    tile-graph stats of the model."""
    op = oracle_params(oracle, preset)
    hh = h.copy()
    with fa.FastScape(W, H, gpu_params(preset), h=hh) as m:
        m.set_path(path)
        for s in range(steps):
            st = oracle.step(hh, W, H, op, oracle.POW_LIBM)
            m.setH(hh)
            m.step(1)
            compare_step(m, st, "%s preset %d step %d" % (what, preset, s))
            exact = oracle.step(hh, W, H, op, oracle.POW_EXACT)["h"]
            assert_h(m.getH().ravel(), exact, st["h"], op.neq, "%s h step %d" % (what, s))
            hh = st["h"]
        return m.graph_stats(), m.path_stats(), st["levels"].size - 3


@pytest.mark.parametrize("kind,W,H", [("ramp", 400, 64), ("ramp", 333, 70),
                                      ("snake", 101, 101), ("snake", 60, 60),
                                      ("snake", 131, 67)])
def test_graph_path_deep_terrains_vs_oracle(oracle, kind, W, H):
    h = ramp_terrain(W, H) if kind == "ramp" else snake_terrain(W, H)[0]
    note("graph path %s %dx%d" % (kind, W, H))
    for preset in (1, 0):
        g, p, D = synced_steps(oracle, W, H, h, preset, 2, fa.PATH_TILE_GRAPH, kind)
        assert D > 64
        assert g["steps"] == 2 and p["tile"] == 0 and p["levels"] == 0
        if kind == "snake" and W > 64:
            assert g["tiers_last_step"] > 10, g   # the channel crosses a tile border per row


@pytest.mark.parametrize("W,H", [(7, 9), (64, 64), (65, 65), (130, 70), (200, 129)])
def test_graph_path_random_and_cone_terrains_vs_oracle(oracle, W, H):
    note("graph path random/cone %dx%d" % (W, H))
    for h in (oracle.init_terrain(W, H, 3), cone_terrain(W, H, 48, 5)):
        for preset in (1, 0):
            synced_steps(oracle, W, H, h, preset, 2, fa.PATH_TILE_GRAPH, "%dx%d" % (W, H))


@pytest.mark.parametrize("W,H,preset,steps", [(300, 300, 0, 100), (257, 190, 0, 60),
                                              (500, 400, 1, 30)])
def test_graph_path_equals_level_path(oracle, W, H, preset, steps):
    """Multi-step runs: the tile-graph path and the level path stay
    bit-identical (the reference preset passes 64 levels within the run)."""
    note("graph vs levels %dx%d preset %d" % (W, H, preset))
    h0 = oracle.init_terrain(W, H, 0)
    with fa.FastScape(W, H, gpu_params(preset), h=h0) as a, \
            fa.FastScape(W, H, gpu_params(preset), h=h0) as b:
        a.set_path(fa.PATH_TILE_GRAPH)
        b.set_path(fa.PATH_LEVELS)
        for s in range(steps):
            a.step(1)
            b.step(1)
            if s % 10 == 9 or s == steps - 1:
                np.testing.assert_array_equal(a.getH(), b.getH(), err_msg="h step %d" % s)
                np.testing.assert_array_equal(a.accum(), b.accum(), err_msg="accum %d" % s)
        np.testing.assert_array_equal(a.levels(), b.levels())
        np.testing.assert_array_equal(a.stack(), b.stack())
        depth = b.levels().size - 3
        if preset == 0:
            assert depth > 64, depth
        assert a.path_stats() == {"tile": 0, "levels": 0, "graph": steps}


def test_auto_path_takes_the_graph_for_deep_drainage(oracle):
    """The reference preset deepens past 32 levels: the default path moves
    from the halo tiles to the tile graph, never to the level path, and
    stays bit-identical to the level path."""
    W, H, steps = 300, 300, 80
    note("auto path 300^2 reference preset")
    h0 = oracle.init_terrain(W, H, 0)
    with fa.FastScape(W, H, gpu_params(0), h=h0) as a, \
            fa.FastScape(W, H, gpu_params(0), h=h0) as b:
        b.set_path(fa.PATH_LEVELS)
        a.step(steps)
        b.step(steps)
        np.testing.assert_array_equal(a.getH(), b.getH())
        np.testing.assert_array_equal(a.accum(), b.accum())
        pa = a.path_stats()
        assert pa["levels"] == 0 and pa["graph"] > 0 and pa["tile"] > 0, pa
        assert pa["tile"] + pa["graph"] == steps
