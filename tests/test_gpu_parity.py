"""GPU parity: the HIP time-step against the reference's golden vectors and
the CPU oracle, through the C ABI (libfastscape_hip.so).

Bar (from the reference's behaviour):
  * receivers, donor counts/lists, the stack, levels[] and accumulation are
    compared bit for bit;
  * elevations are bit-identical for n = 1 (the benchmark preset), where
    every operation of the reference's Newton loop is replayed exactly;
  * for n = 2 (the reference preset) the device evaluates pow(x,2) as x*x
    while glibc's pow differs from x*x in about 0.08% of arguments, so h is
    compared bit for bit with the oracle's product mode (POW_EXACT) and to a
    relative tolerance of 1e-12 per step with the libm mode.
Per-step checks are state-synchronised: the GPU starts each step from the
reference's (or the oracle's) elevations, so a difference cannot hide behind
an earlier one.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import (FINAL_FIXTURES, ROOT, STAGE_FIXTURES, cone_terrain, golden_dem_shas,
                      load_golden)

import fastscape_amd as fa
import tabular

pytestmark = pytest.mark.gpu
REL_TOL_N2 = 1e-12
CLI = os.path.join(ROOT, "barnes2019-landscape_amd", "lib", "fastscape_mi355x")


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def gpu_params(preset, **kw):
    return fa.preset(preset, **kw)


def oracle_params(oracle, preset, **kw):
    return oracle.preset(preset, **kw)


def assert_h(got, ref_exact, ref_libm, neq, what):
    if neq == 1:
        np.testing.assert_array_equal(got, ref_libm, err_msg=what)
    else:
        np.testing.assert_array_equal(got, ref_exact, err_msg=what + " (vs x*x oracle)")
        scale = np.maximum(np.abs(ref_libm), 1.0)
        assert np.all(np.abs(got - ref_libm) <= REL_TOL_N2 * scale), what


def compare_step(m, st, what):
    np.testing.assert_array_equal(m.receivers(), st["rec"], err_msg=what + " rec")
    ndon, donor = m.donors()
    np.testing.assert_array_equal(ndon, st["ndon"], err_msg=what + " ndon")
    np.testing.assert_array_equal(donor, st["donor"], err_msg=what + " donor")
    np.testing.assert_array_equal(m.stack(), st["stack"], err_msg=what + " stack")
    np.testing.assert_array_equal(m.levels(), st["levels"], err_msg=what + " levels")
    np.testing.assert_array_equal(m.accum(), st["accum"], err_msg=what + " accum")


@pytest.mark.parametrize("fx", STAGE_FIXTURES)
def test_stages_vs_reference_golden(oracle, fx):
    g = load_golden(fx)
    W, H, preset = int(g["W"]), int(g["H"]), int(g["preset"])
    note("golden %s %dx%d preset %d" % (fx, W, H, preset))
    op = oracle_params(oracle, preset)
    with fa.FastScape(W, H, gpu_params(preset), h=g["h_init"]) as m:
        h_in = g["h_init"]
        for s in range(int(g["nsteps"])):
            m.setH(h_in)
            m.step(1)
            st = {k: g["%s_%d" % (k, s)] for k in ("rec", "ndon", "donor", "stack", "levels",
                                                  "accum")}
            compare_step(m, st, "%s step %d" % (fx, s))
            exact = oracle.step(h_in, W, H, op, oracle.POW_EXACT)["h"]
            assert_h(m.getH().ravel(), exact, g["h_%d" % s], op.neq, "%s h step %d" % (fx, s))
            h_in = g["h_%d" % s]


# 4096x4096 is the single-grid workload whose receivers and levels are
# checked bit for bit every step (about 25 s, mostly the oracle's steps).
@pytest.mark.parametrize("W,H,preset,steps", [
    (64, 64, 1, 4), (257, 193, 1, 4), (193, 257, 0, 4), (1000, 1000, 1, 3), (1000, 1000, 0, 2),
    (5, 5, 1, 2), (6, 300, 0, 2), (300, 6, 1, 2), (4096, 4096, 1, 2)])
def test_step_sync_vs_oracle(oracle, W, H, preset, steps):
    note("step-sync %dx%d preset %d" % (W, H, preset))
    op = oracle_params(oracle, preset)
    h = oracle.init_terrain(W, H, 7)
    # let the drainage network develop on the CPU first
    for _ in range(3):
        h = oracle.step(h, W, H, op, oracle.POW_EXACT)["h"]
    with fa.FastScape(W, H, gpu_params(preset), h=h) as m:
        for s in range(steps):
            st = oracle.step(h, W, H, op, oracle.POW_LIBM)
            m.setH(h)
            m.step(1)
            compare_step(m, st, "%dx%d step %d" % (W, H, s))
            exact = oracle.step(h, W, H, op, oracle.POW_EXACT)["h"]
            assert_h(m.getH().ravel(), exact, st["h"], op.neq, "%dx%d h step %d" % (W, H, s))
            h = st["h"]


@pytest.mark.parametrize("fx", FINAL_FIXTURES)
def test_full_run_vs_reference(oracle, fx):
    """Whole trajectories (run(nstep)) with no resynchronisation."""
    g = load_golden(fx)
    W, H, nstep, preset = int(g["W"]), int(g["H"]), int(g["nstep"]), int(g["preset"])
    note("full run %s" % fx)
    with fa.FastScape(W, H, gpu_params(preset), engine_seed=0) as m:
        m.run(nstep)
        got = m.getH().ravel()
    if preset == 1:
        np.testing.assert_array_equal(got, g["h"])
    else:
        exact = oracle.run(W, H, nstep, oracle_params(oracle, preset), 0, oracle.POW_EXACT)
        np.testing.assert_array_equal(got, exact)


@pytest.mark.parametrize("key", ["base_501x501_120", "ref_501x501_120"])
def test_dem_vs_reference(oracle, tmp_path, key):
    """compare.sh's configuration: PrintDEM output of run(120) at 501x501."""
    info = golden_dem_shas()[key]
    W, H, preset = info["W"], info["H"], info["preset"]
    note("dem %s" % key)
    with fa.FastScape(W, H, gpu_params(preset), engine_seed=0) as m:
        m.run(info["nstep"])
        h = m.getH().ravel()
    dem = tmp_path / "g.dem"
    fa.write_dem(str(dem), h, W, H)
    sha = hashlib.sha256(dem.read_bytes()).hexdigest()
    if preset != 1:
        # n = 2: bit-exact against the x*x oracle; the file still matches the
        # reference's bytes (differences stay below 6 significant digits)
        exact = oracle.run(W, H, info["nstep"], oracle_params(oracle, preset), 0,
                           oracle.POW_EXACT)
        np.testing.assert_array_equal(h, exact)
    assert sha == info["sha256"]


def ramp_terrain(W, H):
    """Every interior cell drains west: levels grow with x (deep, > 64)."""
    h = np.zeros((H, W))
    xs = np.arange(W, dtype=np.float64)
    h[2:H - 2, 2:W - 2] = 1.0 + xs[2:W - 2][None, :] * 0.5
    h[2:H - 2, 2:W - 2] += np.linspace(0, 0.1, H - 4)[:, None] * 1e-3
    return h.ravel()


def snake_terrain(W, H):
    """A serpentine channel between high walls: one drainage path thousands
    of cells long."""
    h = np.full((H, W), 1e6)
    h[:2, :] = h[-2:, :] = 0
    h[:, :2] = h[:, -2:] = 0
    path = [(3, 2)]          # outlet cell, next to ring 1
    rows = list(range(3, H - 3, 2))
    for i, y in enumerate(rows):
        xs = range(3, W - 3) if i % 2 == 0 else range(W - 4, 2, -1)
        for x in xs:
            path.append((x, y))
        if i + 1 < len(rows):
            gx = W - 4 if i % 2 == 0 else 3
            path.append((gx, y + 1))
    for k, (x, y) in enumerate(path):
        h[y, x] = 1.0 + k * 1e-3
    return h.ravel(), len(path)


@pytest.mark.parametrize("kind", ["ramp", "snake"])
def test_deep_drainage_vs_oracle(oracle, kind):
    W, H = (400, 64) if kind == "ramp" else (101, 101)
    h = ramp_terrain(W, H) if kind == "ramp" else snake_terrain(W, H)[0]
    note("deep %s %dx%d" % (kind, W, H))
    for preset in (1, 0):
        op = oracle_params(oracle, preset)
        hh = h.copy()
        with fa.FastScape(W, H, gpu_params(preset), h=hh) as m:
            for s in range(2):
                st = oracle.step(hh, W, H, op, oracle.POW_LIBM)
                m.setH(hh)
                m.step(1)
                assert st["levels"].size > 3 + 64
                compare_step(m, st, "%s preset %d step %d" % (kind, preset, s))
                exact = oracle.step(hh, W, H, op, oracle.POW_EXACT)["h"]
                assert_h(m.getH().ravel(), exact, st["h"], op.neq, "%s h" % kind)
                hh = st["h"]


# 48-cell basins: 25 levels and at most 460 halo cells on any tile's chains,
# within the 32-cell halo's plan; 60-cell basins offset by 3: 31 levels but
# up to 1642 halo cells on one tile's chains, beyond it (level path).
@pytest.mark.parametrize("period,offset,halo32", [(48, 0, True), (60, 3, False)])
def test_deep_tile_path_vs_oracle(oracle, period, offset, halo32):
    W, H = 200, 160
    h = cone_terrain(W, H, period, offset)
    note("cone basins %d offset %d" % (period, offset))
    for preset in (1, 0):
        op = oracle_params(oracle, preset)
        hh = h.copy()
        with fa.FastScape(W, H, gpu_params(preset), h=hh) as m:
            for s in range(2):
                st = oracle.step(hh, W, H, op, oracle.POW_LIBM)
                m.setH(hh)
                m.step(1)
                if s == 0:
                    assert 16 < st["levels"].size - 3 <= 32
                    expect = {"halo16": 0, "halo32": 1 if halo32 else 0}
                    assert m.tile_stats() == expect
                    ps = m.path_stats()
                    assert ps["levels"] == 0 and ps["graph"] == (0 if halo32 else 1)
                compare_step(m, st, "cone %d preset %d step %d" % (period, preset, s))
                exact = oracle.step(hh, W, H, op, oracle.POW_EXACT)["h"]
                assert_h(m.getH().ravel(), exact, st["h"], op.neq, "cone h step %d" % s)
                hh = st["h"]


def test_flat_and_nan_terrain(oracle):
    W, H = 50, 40
    op = oracle_params(oracle, 1)
    note("flat terrain")
    flat = np.zeros(W * H)
    with fa.FastScape(W, H, gpu_params(1), h=flat) as m:
        m.step(1)
        st = oracle.step(flat, W, H, op)
        compare_step(m, st, "flat")
        np.testing.assert_array_equal(m.getH().ravel(), st["h"])
        assert m.levels().size == 3            # every cell is an outlet
    note("nan terrain")
    h = oracle.init_terrain(W, H, 3)
    h[W * 10 + 10] = np.nan
    h[W * 20 + 30] = np.nan
    with fa.FastScape(W, H, gpu_params(1), h=h) as m:
        m.step(1)
        st = oracle.step(h, W, H, op)
        compare_step(m, st, "nan")
        np.testing.assert_array_equal(m.getH().ravel(), st["h"])


def test_resume_equals_single_run(oracle):
    """run() is resumable (RB:334-340): run(10); run(10) == run(21)."""
    note("resume")
    W, H = 128, 96
    with fa.FastScape(W, H, gpu_params(1), engine_seed=0) as a, \
            fa.FastScape(W, H, gpu_params(1), engine_seed=0) as b:
        a.run(10)
        a.run(10)
        b.run(21)
        np.testing.assert_array_equal(a.getH(), b.getH())
        t = a.timers()
        assert t["Overall"] > 0 and t["Step7: Erosion"] >= 0


def test_api_errors_and_roundtrip():
    note("api")
    W, H = 33, 21
    with fa.FastScape(W, H, gpu_params(1), engine_seed=5) as m:
        np.testing.assert_array_equal(m.getH().ravel(), fa.generate_terrain(W, H, 5))
        x = np.random.default_rng(0).random(W * H)
        m.setH(x)
        np.testing.assert_array_equal(m.getH().ravel(), x)
        with pytest.raises(ValueError):
            m.setH(np.zeros(W * H + 1))
        import ctypes
        lib = fa.library()
        buf = np.zeros(5)
        assert lib.fs_get_h(m._m, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 5) == 8
        assert lib.fs_step(m._m, -1) == 1
        assert m.levels().size == 0            # no step taken yet
    with pytest.raises(fa.FastScapeError):
        fa.FastScape(10, 10, device=99)


def test_cli_dropin(oracle, tmp_path):
    """The drop-in executable: reference arguments and stdout prefixes, and
    the same .dem as the oracle (benchmark preset is bit-exact)."""
    note("cli")
    out = tmp_path / "c.dem"
    r = subprocess.run([CLI, "--preset=benchmark", "101", "30", str(out), "123"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("A ") and lines[1].startswith("C ")
    assert lines[2].startswith("h git_hash") and lines[3] == "m Random seed = 123"
    assert "p Step = 0" in lines and "p Step = 20" in lines and "p Step = 30" not in lines
    assert sum(1 for l in lines if l.startswith("t ")) == 9
    # the results pipeline's columns (d.make_tabular:27-44)
    tabular.check_cli_row(tabular.row(lines), 31)
    h = oracle.run(101, 101, 30, oracle_params(oracle, 1), 0)
    oracle.write_dem(str(tmp_path / "o.dem"), h, 101, 101)
    assert out.read_bytes() == (tmp_path / "o.dem").read_bytes()


def test_large_grid_invariants(oracle):
    """4096x4096 (the per-stage configuration): one state-synchronised step
    against the oracle after a few GPU steps."""
    W = H = 4096
    note("4096^2")
    with fa.FastScape(W, H, gpu_params(1), engine_seed=0) as m:
        m.step(3)
        h = m.getH().ravel().copy()
        st = oracle.step(h, W, H, oracle_params(oracle, 1))
        m.step(1)
        compare_step(m, st, "4096 step")
        np.testing.assert_array_equal(m.getH().ravel(), st["h"])


@pytest.mark.parametrize("W,H,preset,steps", [(1000, 800, 1, 30), (130, 70, 1, 40),
                                              (300, 300, 0, 60), (1000, 1000, 1, 500)])
def test_tile_path_equals_level_path(oracle, W, H, preset, steps):
    """The default path and the level path are bit-identical step after step.
    The reference preset deepens past 16 levels within ten steps (32-cell
    halo) and past 32 soon after (tile-graph path); the benchmark preset at
    1000^2 passes 16 levels after about 350 steps and stays on the tile
    path."""
    note("tile vs levels %dx%d preset %d" % (W, H, preset))
    h0 = oracle.init_terrain(W, H, 0)
    with fa.FastScape(W, H, gpu_params(preset), h=h0) as a, \
            fa.FastScape(W, H, gpu_params(preset), h=h0) as b:
        b.set_path(fa.PATH_LEVELS)
        for s in range(steps):
            a.step(1)
            b.step(1)
            if s % 10 == 9 or s == steps - 1:
                np.testing.assert_array_equal(a.getH(), b.getH(), err_msg="h step %d" % s)
                np.testing.assert_array_equal(a.accum(), b.accum(), err_msg="accum %d" % s)
                np.testing.assert_array_equal(a.levels(), b.levels(), err_msg="levels %d" % s)
                np.testing.assert_array_equal(a.stack(), b.stack(), err_msg="stack %d" % s)
        pa, pb = a.path_stats(), b.path_stats()
        assert pb == {"tile": 0, "levels": steps, "graph": 0}
        assert pa["levels"] == 0 and pa["tile"] + pa["graph"] == steps
        ta = a.tile_stats()
        assert ta["halo16"] + ta["halo32"] == pa["tile"]
        if preset == 1:
            assert pa["tile"] == steps            # shallow drainage: always tiles
        else:
            assert pa["graph"] > 0                 # deep drainage: the tile graph
        if steps >= 60:
            assert ta["halo16"] > 0 and ta["halo32"] > 0, ta


def test_benchmark_grid_takes_tile_path(oracle):
    note("tile path at 2000^2")
    W = H = 2000
    with fa.FastScape(W, H, gpu_params(1), engine_seed=0) as m:
        m.step(5)
        assert m.path_stats() == {"tile": 5, "levels": 0, "graph": 0}
        h = m.getH().ravel().copy()
        st = oracle.step(h, W, H, oracle_params(oracle, 1))
        m.step(1)
        compare_step(m, st, "2000 tile step")
        np.testing.assert_array_equal(m.getH().ravel(), st["h"])


def test_kernel_and_stage_timing_on_both_paths(oracle):
    """Per-kernel HIP-event timing and the stage timers while the tile path
    runs (its erosion is still in flight when the host collects) and on the
    level path; timing does not change the results."""
    note("kernel timing 300x260")
    W, H = 300, 260
    h0 = oracle.init_terrain(W, H, 0)
    with fa.FastScape(W, H, gpu_params(1), h=h0) as a, \
            fa.FastScape(W, H, gpu_params(1), h=h0) as b:
        a.set_kernel_timing(True)
        a.step(6)
        b.step(6)
        st = a.kernel_stats()
        for k in ("k_receivers", "k_tile_acc", "k_tile_erode"):
            assert st[k][1] == 6 and st[k][0] > 0, (k, st[k])
        assert a.path_stats() == {"tile": 6, "levels": 0, "graph": 0}
        np.testing.assert_array_equal(a.getH(), b.getH())
        a.set_path(fa.PATH_LEVELS)
        a.set_kernel_timing(True)
        a.step(2)
        st = a.kernel_stats()
        assert st["k_receivers"][1] == 2 and st["k_erode"][1] + st["k_uplift_roots"][1] > 0
        assert st["k_tile_acc"][1] == 0
        t = a.timers()
        assert all(v >= 0 for v in t.values()) and max(t.values()) > 0, t


def test_stage_timers_cover_every_stage_on_both_paths():
    """The reference's seven stage timers (RB:380-387) are all non-zero on the
    tile path too: its two kernels' event time is split over donors, order,
    accumulation, uplift and erosion by the phase ticks of sampled
    workgroups.  The split adds up to the kernels' time and stays within the
    wall time of the run."""
    note("stage timers 1000^2")
    W = H = 1000
    stages = fa.TIMER_NAMES[1:7]
    with fa.FastScape(W, H, gpu_params(1), engine_seed=0) as m:
        m.reset_timers()
        m.step(20)
        assert m.path_stats() == {"tile": 20, "levels": 0, "graph": 0}
        t = m.timers()
        note("tile path timers %s" % t)
        assert all(t[k] > 0 for k in stages), t
        assert sum(t[k] for k in stages) <= t["Overall"]
        # accumulation and erosion dominate, as in the reference
        assert t["Step7: Erosion"] > t["Step6: Uplift"]
        m.reset_timers()
        m.set_path(fa.PATH_LEVELS)
        m.step(5)
        t = m.timers()
        note("level path timers %s" % t)
        assert all(t[k] > 0 for k in stages), t
        assert sum(t[k] for k in stages) <= t["Overall"]


@pytest.mark.parametrize("W,H", [(300, 260), (130, 70)])
def test_receivers_of_consecutive_tile_steps(oracle, W, H):
    """Consecutive tile-path steps without set_h (the tile path alternates two
    elevation buffers): each step's receivers are the reference's receivers
    of that step's input, also after the device pointer was handed out."""
    note("consecutive-step receivers %dx%d" % (W, H))
    op = oracle_params(oracle, 1)
    with fa.FastScape(W, H, gpu_params(1), engine_seed=3) as m:
        m.step(2)
        for k in range(4):
            h_in = m.getH().ravel().copy()
            m.step(1)
            np.testing.assert_array_equal(m.receivers(), oracle.receivers(h_in, W, H, op),
                                          err_msg="step %d" % k)
        assert m.path_stats()["tile"] >= 5
        m.device_h()
        h_in = m.getH().ravel().copy()
        m.step(1)
        np.testing.assert_array_equal(m.receivers(), oracle.receivers(h_in, W, H, op))


def _hip_dtoh(ptr, count):
    """hipMemcpy of `count` doubles from a device address (the HIP runtime's
    own entry point, outside the library)."""
    import ctypes
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    out = np.empty(count, dtype=np.float64)
    assert hip.hipMemcpy(out.ctypes.data, ptr, count * 8, 2) == 0   # device to host
    return out


def test_device_pointer_names_the_field_until_the_next_step():
    """fs_device_h is valid until the next step: the tile path exchanges its
    two elevation buffers, so the pointer is fetched again after each step
    and then always holds what getH() returns."""
    note("device pointer lifetime")
    W, H = 300, 260
    with fa.FastScape(W, H, gpu_params(1), engine_seed=3) as m:
        p0 = m.device_h()
        np.testing.assert_array_equal(_hip_dtoh(p0, W * H), m.getH().ravel())
        m.step(1)
        p1 = m.device_h()
        np.testing.assert_array_equal(_hip_dtoh(p1, W * H), m.getH().ravel())
        assert m.path_stats() == {"tile": 1, "levels": 0, "graph": 0} and p1 != p0
        m.step(1)
        p2 = m.device_h()
        assert p2 == p0
        np.testing.assert_array_equal(_hip_dtoh(p2, W * H), m.getH().ravel())
        m.set_path(fa.PATH_LEVELS)          # the level path updates in place
        m.step(1)
        assert m.device_h() == p2
        np.testing.assert_array_equal(_hip_dtoh(p2, W * H), m.getH().ravel())
