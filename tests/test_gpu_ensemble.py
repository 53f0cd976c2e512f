"""GPU parity of the batched ensemble (fs_ensemble_*, one persistent launch per
step) against single-grid models and the CPU oracle.

Every realization of an ensemble must give, step for step, the same bits as
a single-grid model (fs_model) with its size, parameters and terrain seed:
receivers, upstream accumulation, levels[] and elevations are compared
exactly.  The single-grid model is itself pinned to the reference (the
golden-vector tests), and the small cases here are also checked directly
against the oracle.  Realizations deeper than 16 levels take the tile-graph
step (all of them in the same launches), the others the halo tiles.  The
work schedule (receiver rows, tiles per claim) and the autotuner must not
change a bit.
"""
import sys

import numpy as np
import pytest

import fastscape_amd as fa
from conftest import cone_terrain

pytestmark = pytest.mark.gpu


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def bench(**kw):
    return fa.preset(fa.PRESET_BENCHMARK, **kw)


STRONG = dict(keq=1e-4, meq=0.6, ueq=5e-4)     # deep drainage within ~100 steps
MID = dict(keq=3e-5, meq=0.5, ueq=1e-3)

MIXED = [
    (37, 29, bench(), 3),
    (64, 64, bench(**STRONG), 1),
    (130, 70, fa.preset(fa.PRESET_REFERENCE), 0),   # n = 2
    (200, 150, bench(**MID), 7),
    (96, 96, bench(**STRONG), 11),
    (5, 5, bench(), 0),
    (65, 129, bench(keq=2e-5, meq=0.45, ueq=2e-3), 2),
]


def make(specs):
    return fa.Ensemble([fa.realization(W, H, p, s) for (W, H, p, s) in specs])


def singles(specs):
    return [fa.FastScape(W, H, p, engine_seed=s) for (W, H, p, s) in specs]


def compare_all(ens, models, specs, what, stages=True):
    for r, (m, (W, H, _, s)) in enumerate(zip(models, specs)):
        tag = "%s r%d %dx%d seed %d" % (what, r, W, H, s)
        np.testing.assert_array_equal(ens.getH(r), m.getH(), err_msg=tag + " h")
        if stages:
            np.testing.assert_array_equal(ens.receivers(r), m.receivers(), err_msg=tag + " rec")
            np.testing.assert_array_equal(ens.accum(r), m.accum(), err_msg=tag + " accum")
            np.testing.assert_array_equal(ens.levels(r), m.levels(), err_msg=tag + " levels")


def test_mixed_ensemble_step_by_step():
    note("mixed ensemble %d realizations, 160 steps" % len(MIXED))
    ens = make(MIXED)
    models = singles(MIXED)
    try:
        compare_all(ens, models, MIXED, "initial", stages=False)
        for step in range(160):
            ens.step()
            for m in models:
                m.step(1)
            compare_all(ens, models, MIXED, "step %d" % step, stages=(step % 8 == 0 or step > 150))
        deep = [ens.path_stats(r)[1] for r in range(len(MIXED))]
        tiles = [ens.path_stats(r)[0] for r in range(len(MIXED))]
        note("tile-graph steps per realization %s, tile steps %s" % (deep, tiles))
        assert sum(deep) > 0, "no realization took the tile-graph step"
        assert sum(tiles) > 0, "no realization took the tile path"
        assert all(t + d == 160 for t, d in zip(tiles, deep))
        for r, (W, H, _, _) in enumerate(MIXED):
            lv = ens.levels(r)
            assert lv[-1] == lv[-2] == W * H
    finally:
        ens.close()
        for m in models:
            m.close()


def test_run_matches_oracle(oracle):
    specs = [(48, 40, 5, dict(), fa.PRESET_BENCHMARK), (33, 57, 9, STRONG, fa.PRESET_BENCHMARK),
             (41, 41, 0, dict(), fa.PRESET_REFERENCE)]
    nstep = 90
    note("ensemble run(%d) vs oracle" % nstep)
    with fa.Ensemble([fa.realization(W, H, fa.preset(pr, **kw), s)
                      for (W, H, s, kw, pr) in specs]) as ens:
        ens.run(nstep)
        for r, (W, H, s, kw, pr) in enumerate(specs):
            op = oracle.preset(pr, **kw)
            ref = oracle.run(W, H, nstep, op, engine_seed=s, powmode=oracle.POW_EXACT)
            np.testing.assert_array_equal(ens.getH(r).ravel(), np.asarray(ref).ravel(),
                                          err_msg="realization %d" % r)


SCHED_SPECS = [
    (300, 300, bench(**STRONG), 4),
    (520, 260, bench(**STRONG), 5),
    (257, 193, bench(**MID), 6),
    (150, 400, bench(), 8),
    (70, 70, bench(**STRONG), 9),
]


@pytest.mark.parametrize("sched", [dict(recv_rows=16, tiles_per_claim=1, group_cells=16384),
                                   dict(recv_rows=256, tiles_per_claim=4, group_cells=65536),
                                   dict(recv_rows=64, tiles_per_claim=2, group_cells=0),
                                   dict(recv_rows=32, tiles_per_claim=16, group_cells=8192)])
def test_schedules_give_identical_bits(sched):
    note("schedule %s" % sched)
    with make(SCHED_SPECS) as base, make(SCHED_SPECS) as other:
        other.set_schedule(fa.schedule(**sched))
        base.run(129)
        other.run(129)
        for r in range(len(SCHED_SPECS)):
            np.testing.assert_array_equal(other.getH(r), base.getH(r), err_msg="r%d" % r)
            np.testing.assert_array_equal(other.levels(r), base.levels(r), err_msg="r%d" % r)
        assert sum(base.path_stats(r)[1] for r in range(len(SCHED_SPECS))) > 0


def test_schedule_matches_single_models():
    note("schedule with group_cells (no effect) vs single-grid models")
    with make(SCHED_SPECS) as ens:
        ens.set_schedule(fa.schedule(recv_rows=16, tiles_per_claim=3, group_cells=4096))
        models = singles(SCHED_SPECS)
        try:
            ens.run(119)
            for m in models:
                m.run(119)
            compare_all(ens, models, SCHED_SPECS, "after run(119)")
        finally:
            for m in models:
                m.close()


def test_autotune_keeps_results():
    note("autotune")
    with make(SCHED_SPECS) as tuned, make(SCHED_SPECS) as plain:
        tuned.autotune(2, interval=40)
        tuned.run(99)
        plain.run(99)
        log = tuned.tune_log()
        note("tune log %s" % log)
        assert len(log) >= 3
        assert all(ms > 0 for _, ms in log)
        steps, ms = tuned.stats()
        assert steps == 100 and ms > 0
        for r in range(len(SCHED_SPECS)):
            np.testing.assert_array_equal(tuned.getH(r), plain.getH(r), err_msg="r%d" % r)


def test_set_h_resume_and_errors():
    specs = [(40, 30, bench(), 1), (64, 65, bench(**STRONG), 2)]
    with make(specs) as ens, fa.FastScape(64, 65, bench(**STRONG), engine_seed=2) as m:
        ens.run(10)
        m.run(10)
        h = ens.getH(1)
        np.testing.assert_array_equal(h, m.getH())
        ens.setH(1, h)
        ens.run(20)
        m.run(20)
        np.testing.assert_array_equal(ens.getH(1), m.getH())
        with pytest.raises(fa.FastScapeError):
            ens.getH(5)
        with pytest.raises(fa.FastScapeError):
            ens.setH(0, np.zeros(7))
        with pytest.raises(fa.FastScapeError):
            ens.set_schedule(fa.schedule(recv_rows=20))


def test_ensemble_tile_graph_matches_deep_tile_halo():
    """A realization 17-32 levels deep (48-cell conical basins, 25 levels):
    the ensemble runs it on the tile graph, the single-grid model on the tile
    path with the 32-cell halo; both give the same bits at every stage."""
    note("ensemble with a 25-level realization")
    cone = cone_terrain(200, 160, 48, 0)
    specs = [(200, 160, bench(), 0), (64, 64, bench(), 1)]
    with make(specs) as ens, fa.FastScape(200, 160, bench(), h=cone) as m, \
            fa.FastScape(64, 64, bench(), engine_seed=1) as m1:
        ens.setH(0, cone)
        for step in range(4):
            ens.step()
            m.step(1)
            m1.step(1)
            compare_all(ens, [m, m1], specs, "cone step %d" % step)
            assert 16 < ens.levels(0).size - 3 <= 32
        assert ens.path_stats(0) == (0, 4)
        assert m.tile_stats()["halo32"] >= 1


def test_scaled_config4_sample():
    """The ensemble workload's parameter distribution at 1/8 of its sides."""
    reals, meta = fa.config4_realizations(24, master_seed=3, sides=(64, 96, 128, 192, 256))
    note("scaled ensemble workload, %d realizations" % len(reals))
    with fa.Ensemble(reals) as ens:
        ens.autotune(1)
        ens.run(149)
        for r in (0, 5, 11, 17, 23):
            d = meta[r]
            p = bench(keq=d["K"], ueq=d["U"], meq=d["m"])
            with fa.FastScape(d["side"], d["side"], p, engine_seed=d["seed"]) as m:
                m.run(149)
                np.testing.assert_array_equal(ens.getH(r), m.getH(), err_msg=str(d))
                np.testing.assert_array_equal(ens.levels(r), m.levels(), err_msg=str(d))
