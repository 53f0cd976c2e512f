"""GPU parity at the full benchmark workloads.

The other GPU tests pin every stage on small and mid-size grids; these run
the sizes the benchmark lines are quoted on, so that bit-exactness there
does not rest on extrapolation:

  * 10000 x 10000 (the headline grid): one state-synchronised step against
    the oracle after three GPU steps -- receivers, donor lists, the stack,
    levels[], accumulation and elevations compared bit for bit -- and a
    300-step run of the tile path against the level path through the switch
    from the 16- to the 32-cell tile halo (about step 261); and, with the
    reference's own constants, 64 steps of the default path (the tile graph
    from step 10, 180+ levels) against the level path;
  * the ensemble workload (256 realizations, sides 512..2048, sampled K, U,
    m): every realization's elevations and levels[] equal those of a
    single-grid model with its size, parameters and seed at steps 10, 20
    and 40 (the first realizations pass 16 levels near step 10), and one
    512 x 512 realization is state-synchronised against the oracle;
  * the learned proposal at 256 x 256: the network's forward pass and every
    layer's gradient against the CPU reference (tests/unet_ref.py), and an
    inversion with 2000-step exact runs, 2048 training runs and 8^3
    candidates.
The bar is the reference's compare.sh bar (every variant gives the same
output, compare.sh:5-23): bit-identical integers and, for n = 1, elevations.
"""
import hashlib
import sys

import numpy as np
import pytest

import fastscape_amd as fa
import unet_ref as ur
from fastscape_amd import inversion as inv

pytestmark = pytest.mark.gpu


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def bench_params(**kw):
    return fa.preset(fa.PRESET_BENCHMARK, **kw)


def compare_step(m, st, what):
    np.testing.assert_array_equal(m.receivers(), st["rec"], err_msg=what + " rec")
    ndon, donor = m.donors()
    np.testing.assert_array_equal(ndon, st["ndon"], err_msg=what + " ndon")
    np.testing.assert_array_equal(donor, st["donor"], err_msg=what + " donor")
    del ndon, donor
    np.testing.assert_array_equal(m.stack(), st["stack"], err_msg=what + " stack")
    np.testing.assert_array_equal(m.levels(), st["levels"], err_msg=what + " levels")
    np.testing.assert_array_equal(m.accum(), st["accum"], err_msg=what + " accum")


# ---------------------------------------------------------------------------
# the headline grid, 10000 x 10000
# ---------------------------------------------------------------------------
def test_headline_grid_state_synchronised_step(oracle):
    W = H = 10000
    note("10000^2: 3 GPU steps, then one step against the oracle")
    op = oracle.preset(1)
    with fa.FastScape(W, H, bench_params(), engine_seed=0) as m:
        m.step(3)
        assert m.path_stats() == {"tile": 3, "levels": 0, "graph": 0}
        h = m.getH().ravel().copy()
        st = oracle.step(h, W, H, op)
        del h
        m.step(1)
        assert m.path_stats()["tile"] == 4
        compare_step(m, st, "10000^2 step 3")
        np.testing.assert_array_equal(m.getH().ravel(), st["h"], err_msg="10000^2 h")


def test_headline_grid_tile_path_equals_level_path_through_halo_switch():
    """300 steps from the seed-0 terrain: the tile path takes the 16-cell halo
    until the drainage passes 16 levels (about step 261), then the 32-cell
    halo; the level path runs the reference's stage order every step."""
    W = H = 10000
    note("10000^2: tile path vs level path, 300 steps")
    with fa.FastScape(W, H, bench_params(), engine_seed=0) as a, \
            fa.FastScape(W, H, bench_params(), engine_seed=0) as b:
        b.set_path(fa.PATH_LEVELS)
        done = 0
        for upto in (100, 262, 300):
            a.step(upto - done)
            b.step(upto - done)
            done = upto
            note("10000^2 compare at step %d: tile stats %s" % (upto, a.tile_stats()))
            np.testing.assert_array_equal(a.getH(), b.getH(), err_msg="h step %d" % upto)
            np.testing.assert_array_equal(a.accum(), b.accum(), err_msg="accum %d" % upto)
            np.testing.assert_array_equal(a.levels(), b.levels(), err_msg="levels %d" % upto)
            np.testing.assert_array_equal(a.stack(), b.stack(), err_msg="stack %d" % upto)
        ta = a.tile_stats()
        assert a.path_stats()["tile"] == 300 and b.path_stats() == {"tile": 0, "levels": 300, "graph": 0}
        assert ta["halo16"] > 200 and ta["halo32"] > 0, ta


def test_headline_grid_reference_preset_graph_path_equals_level_path():
    """The reference's own constants (fastscape_RB.cpp:58-65) at 10000^2:
    drainage passes 32 levels within ten steps, so the default path moves
    from the halo tiles to the tile graph (180+ levels by step 60); it stays
    bit-identical to the level path, including the readouts rebuilt from the
    graph steps' receivers."""
    W = H = 10000
    note("10000^2 reference preset: default path vs level path, 64 steps")
    ref = fa.preset(fa.PRESET_REFERENCE)
    with fa.FastScape(W, H, ref, engine_seed=0) as a, \
            fa.FastScape(W, H, ref, engine_seed=0) as b:
        b.set_path(fa.PATH_LEVELS)
        done = 0
        for upto in (8, 30, 64):
            a.step(upto - done)
            b.step(upto - done)
            done = upto
            note("10000^2 reference preset at step %d: paths %s, graph %s"
                 % (upto, a.path_stats(), a.graph_stats()))
            np.testing.assert_array_equal(a.getH(), b.getH(), err_msg="h step %d" % upto)
            np.testing.assert_array_equal(a.accum(), b.accum(), err_msg="accum %d" % upto)
            np.testing.assert_array_equal(a.levels(), b.levels(), err_msg="levels %d" % upto)
        pa = a.path_stats()
        assert pa["graph"] > 50 and pa["levels"] == 0, pa
        assert b.levels().size - 3 > 150


# ---------------------------------------------------------------------------
# the ensemble workload: 256 realizations, sides 512..2048
# ---------------------------------------------------------------------------
CHECKPOINTS = (10, 20, 40)


def _digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def ensemble_digests():
    """Digests of every realization's elevations and levels[] at each
    checkpoint of one run of the whole ensemble workload (autotuned, as in
    the benchmark)."""
    reals, meta = fa.config4_realizations(256, master_seed=0)
    note("ensemble workload: 256 realizations, %d steps" % CHECKPOINTS[-1])
    dig = {c: [] for c in CHECKPOINTS}
    paths = None
    with fa.Ensemble(reals) as ens:
        ens.autotune(1, 20)
        done = 0
        for c in CHECKPOINTS:
            for _ in range(c - done):
                ens.step()
            done = c
            for r in range(len(reals)):
                dig[c].append(_digest(ens.getH(r), ens.levels(r)))
            note("ensemble digests at step %d" % c)
        paths = [ens.path_stats(r) for r in range(len(reals))]
    return meta, dig, paths


@pytest.mark.parametrize("part", range(4))
def test_ensemble_workload_equals_single_models(ensemble_digests, part):
    meta, dig, paths = ensemble_digests
    if part == 0:
        # the run crosses from tiles to the level path in some realizations
        assert sum(l for _, l in paths) > 0 and sum(t for t, _ in paths) > 0, paths
    rs = range(part * 64, (part + 1) * 64)
    note("ensemble workload: single-grid models %d..%d" % (rs[0], rs[-1]))
    for r in rs:
        d = meta[r]
        p = bench_params(keq=d["K"], ueq=d["U"], meq=d["m"])
        with fa.FastScape(d["side"], d["side"], p, engine_seed=d["seed"]) as m:
            done = 0
            for c in CHECKPOINTS:
                m.step(c - done)
                done = c
                assert _digest(m.getH(), m.levels()) == dig[c][r], \
                    "realization %d %s differs at step %d" % (r, d, c)


def test_ensemble_workload_state_synchronised_vs_oracle(oracle):
    """A 512 x 512 realization of the workload, 40 steps in, stepped once more
    from its own elevations against the oracle (every other realization
    advances with it)."""
    reals, meta = fa.config4_realizations(256, master_seed=0)
    r = next(i for i, d in enumerate(meta) if d["side"] == 512)
    d = meta[r]
    note("ensemble workload: realization %d %s vs the oracle" % (r, d))
    op = oracle.preset(1, keq=d["K"], ueq=d["U"], meq=d["m"])
    W = H = 512
    with fa.Ensemble(reals) as ens:
        for _ in range(40):
            ens.step()
        h = ens.getH(r).ravel().copy()
        st = oracle.step(h, W, H, op)
        ens.step()
        np.testing.assert_array_equal(ens.receivers(r), st["rec"], err_msg="rec")
        np.testing.assert_array_equal(ens.accum(r), st["accum"], err_msg="accum")
        np.testing.assert_array_equal(ens.levels(r), st["levels"], err_msg="levels")
        np.testing.assert_array_equal(ens.getH(r).ravel(), st["h"], err_msg="h")


# ---------------------------------------------------------------------------
# the learned proposal at 256 x 256
# ---------------------------------------------------------------------------
def _unet_sample(seed, B, H, W):
    rng = np.random.default_rng(seed)
    terrain = rng.random((B, H, W)).astype(np.float32)
    params = rng.uniform(-1, 1, (B, 3)).astype(np.float32)
    target = rng.normal(0, 1, (B, H, W)).astype(np.float32)
    return terrain, params, target


def _perturbed(net, seed):
    p = net.get_params().astype(np.float64)
    p += np.random.default_rng(seed).normal(0, 0.02, p.size)
    return p.astype(np.float32)


def test_unet_forward_and_gradients_at_256():
    """Same bars as the small-size network tests: forward within 3% of the
    output range, every layer's gradient within 5% relative L2 error with
    cosine > 0.995 (BF16 storage emulated in the reference)."""
    H = W = 256
    note("unet 256x256 forward (batch 2) and gradients (batch 1)")
    terrain, params, target = _unet_sample(31, 2, H, W)
    with fa.UNet(H, W, max_batch=2, seed=41) as net:
        p = _perturbed(net, 42)
        net.set_params(p)
        y = net.forward(terrain, params)
        loss, g = net.loss_grad(terrain[:1], params[:1], target[:1])
    ref = ur.Net(p.astype(np.float64), bf16=True)
    yr = ref.forward(terrain.astype(np.float64), params.astype(np.float64))
    err, span = np.abs(y - yr).max(), np.abs(yr).max()
    note("forward max err %.3e of range %.3e" % (err, span))
    assert err <= 0.03 * span + 1e-3
    assert np.corrcoef(y.ravel(), yr.ravel())[0, 1] > 0.999
    rl, rg = ref.loss_grad(terrain[:1].astype(np.float64), params[:1].astype(np.float64),
                           target[:1].astype(np.float64))
    note("loss gpu %.6f ref %.6f" % (loss, rl))
    assert abs(loss - rl) <= 1e-2 * abs(rl)
    off = 0
    for i, (ci, co, lv) in enumerate(ur.layer_shapes()):
        for what, n in (("W", 9 * ci * co), ("b", co)):
            a, b = g[off:off + n].astype(np.float64), rg[off:off + n]
            off += n
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            assert rel < 0.05 and cos > 0.995, (i, what, rel, cos)
    a, b = g[off:].astype(np.float64), rg[off:]
    assert np.linalg.norm(a - b) <= 0.02 * np.linalg.norm(b)


def test_inversion_at_256_with_2000_step_runs():
    """The inversion workload's grid and run length with a reduced search:
    2048 exact training runs, 8^3 candidates, the top 32 rerun exactly.  The
    proposal's winner must fit the target within 5% of the exact-only
    search's best misfit; the exact solver judges both."""
    side, nstep, ntrain, topk = 256, 2000, 2048, 32
    terrain, pn, t, params, secs = inv.make_training_set(ntrain, side, nstep, seed0=1,
                                                         master_seed=0, log=note)
    note("training set: %d exact runs in %.1f s" % (ntrain, secs))
    prop = inv.Proposal(side, max_batch=32, seed=0)
    try:
        hist = prop.train(terrain, pn, t, epochs=12, batch=32, lr=1e-3, seed=1)
        note("training loss %s" % np.array2string(hist, precision=4))
        assert hist[-1] < 0.5 * hist[0]
        del terrain, t
        hidden = np.array([[2.3e-5, 1.7e-3, 0.44]])
        seed = 900001
        target = {}
        inv.exact_runs(hidden, seed, side, nstep, on_result=lambda i, h: target.update(h=h))
        tt = inv.transform(target["h"])
        cands = inv.candidate_grid(8)
        best_e, mis_e, t_exact = inv.exact_only_search(tt, seed, cands, side, nstep)
        best_p, mis_p, top, pred, t_prop = inv.proposal_search(prop, tt, seed, cands, side,
                                                               nstep, top_k=topk)
        note("exact-only best %d misfit %.5f (%.2f s); proposal best %d misfit %.5f (%.2f s)"
             % (best_e, mis_e[best_e], t_exact, best_p, mis_p.min(), t_prop))
        assert mis_p.min() <= 1.05 * mis_e[best_e]
        assert t_prop < t_exact
    finally:
        prop.close()
