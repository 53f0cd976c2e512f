"""CPU oracle pinned against the compiled reference and against invariants.

The fixtures in tests/golden/ were written by the reference's own code
(fastscape_RB.cpp compiled from its sources, tests/golden/make_golden.py).
Every oracle stage is compared bit-for-bit with them; the invariant tests
below fix what the fixtures leave open (brute force on small grids, closed
forms), chosen so that a swapped direction, an off-by-one level or a wrong
exponent fails at least one of them.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import FINAL_FIXTURES, STAGE_FIXTURES, golden_dem_shas, load_golden, params_for

SQRT2 = 1.414213562373095048801688724209698078569671875376948
OFFS = lambda W: [-1, -W - 1, -W, -W + 1, 1, W + 1, W, W - 1]   # RB:137


# ---------------------------------------------------------------------------
# random initialisation (random.cpp:19-49, RB:111-130)
# ---------------------------------------------------------------------------
def test_mt19937_known_answer(oracle):
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # std::mt19937 (seed 5489) is 4123659995.
    e = oracle.MT19937(5489)
    for _ in range(9999):
        e.next()
    assert e.next() == 4123659995


def test_canonical_matches_libstdcxx(oracle, tmp_path):
    """fso_canonical equals std::uniform_real_distribution<>(0,1) on the same
    engine, as compiled by this image's libstdc++ (random.cpp:45-49)."""
    src = tmp_path / "canon.cpp"
    src.write_text(
        "#include <random>\n#include <cstdio>\n#include <cstdlib>\n"
        "int main(int c, char**v){ std::mt19937 e(std::strtoul(v[1],0,10));"
        " std::uniform_real_distribution<> d; using P=std::uniform_real_distribution<>::param_type;"
        " for(int i=0;i<20000;i++){ double x=d(e,P{0,1}); std::printf(\"%a\\n\", x);} }\n")
    exe = tmp_path / "canon"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", str(src), "-o", str(exe)], check=True)
    for seed in (0, 1, 123, 4294967295):
        out = subprocess.run([str(exe), str(seed)], check=True, capture_output=True,
                             text=True).stdout.split()
        e = oracle.MT19937(seed)
        mine = [e.canonical() for _ in range(len(out))]
        assert [float.fromhex(x) for x in out] == mine


@pytest.mark.parametrize("fx", STAGE_FIXTURES)
def test_init_terrain_matches_reference(oracle, fx):
    g = load_golden(fx)
    W, H = int(g["W"]), int(g["H"])
    # the CLI seed was 123, which the reference maps to engine seed 0
    np.testing.assert_array_equal(oracle.init_terrain(W, H, 0), g["h_init"])


# ---------------------------------------------------------------------------
# per-stage parity with the reference
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("fx", STAGE_FIXTURES)
def test_stages_match_reference(oracle, fx):
    g = load_golden(fx)
    W, H = int(g["W"]), int(g["H"])
    p = params_for(oracle, g["preset"])
    h = g["h_init"]
    for s in range(int(g["nsteps"])):
        st = oracle.step(h, W, H, p, oracle.POW_LIBM)
        np.testing.assert_array_equal(st["rec"], g[f"rec_{s}"], err_msg=f"rec step {s}")
        np.testing.assert_array_equal(st["ndon"], g[f"ndon_{s}"], err_msg=f"ndon step {s}")
        np.testing.assert_array_equal(st["donor"], g[f"donor_{s}"], err_msg=f"donor step {s}")
        np.testing.assert_array_equal(st["stack"], g[f"stack_{s}"], err_msg=f"stack step {s}")
        np.testing.assert_array_equal(st["levels"], g[f"levels_{s}"], err_msg=f"levels {s}")
        np.testing.assert_array_equal(st["accum"], g[f"accum_{s}"], err_msg=f"accum step {s}")
        np.testing.assert_array_equal(st["hup"], g[f"hup_{s}"], err_msg=f"uplift step {s}")
        np.testing.assert_array_equal(st["h"], g[f"h_{s}"], err_msg=f"h step {s}")
        h = g[f"h_{s}"]


@pytest.mark.parametrize("fx", FINAL_FIXTURES)
def test_run_matches_reference(oracle, fx):
    g = load_golden(fx)
    W, H, nstep = int(g["W"]), int(g["H"]), int(g["nstep"])
    p = params_for(oracle, g["preset"])
    h = oracle.run(W, H, nstep, p, 0, oracle.POW_LIBM)
    np.testing.assert_array_equal(h, g["h"])


@pytest.mark.slow
@pytest.mark.parametrize("key", ["base_501x501_120", "ref_501x501_120"])
def test_dem_matches_reference(oracle, tmp_path, key):
    """compare.sh's configuration (501x501, run(120)): byte-identical .dem."""
    info = golden_dem_shas()[key]
    W, H = info["W"], info["H"]
    exe = oracle.CLI_PATH
    if not os.path.exists(exe):
        oracle.build()
    dem = tmp_path / "o.dem"
    subprocess.run([exe, str(W), str(H), str(info["nstep"]), "0", str(info["preset"]), "0", "0",
                    str(dem)], check=True, capture_output=True)
    data = dem.read_bytes()
    assert len(data) == info["bytes"]
    assert hashlib.sha256(data).hexdigest() == info["sha256"]


# ---------------------------------------------------------------------------
# invariants that do not depend on the fixtures
# ---------------------------------------------------------------------------
def _random_h(W, H, seed):
    rng = np.random.default_rng(seed)
    h = rng.random(W * H)
    h = h.reshape(H, W)
    h[:2, :] = h[-2:, :] = 0
    h[:, :2] = h[:, -2:] = 0
    return h.ravel()


def _brute_receivers(h, W, H, p):
    """Argmax of the eight slopes with the first maximum winning, evaluated
    with numpy's vectorised division (not the oracle's loop)."""
    hh = h.reshape(H, W)
    rec = np.full((H, W), -1, dtype=np.int32)
    c = hh[2:-2, 2:-2]
    d = [p.dx, p.dx * SQRT2] * 4
    nb = [hh[2:-2, 1:-3], hh[1:-3, 1:-3], hh[1:-3, 2:-2], hh[1:-3, 3:-1], hh[2:-2, 3:-1],
          hh[3:-1, 3:-1], hh[3:-1, 2:-2], hh[3:-1, 1:-3]]
    slopes = np.stack([(c - nb[n]) / d[n] for n in range(8)] + [np.zeros_like(c)])
    # index 8 (slope 0) represents NO_FLOW; argmax returns the first maximum,
    # so put the zero first and shift
    order = np.concatenate([slopes[8:9], slopes[:8]])
    k = np.argmax(order, axis=0) - 1
    # a direction wins only with a strictly positive slope
    k[np.max(slopes[:8], axis=0) <= 0] = -1
    rec[2:-2, 2:-2] = k
    return rec.ravel()


@pytest.mark.parametrize("W,H,seed", [(9, 7, 1), (40, 33, 2), (128, 96, 3)])
@pytest.mark.parametrize("preset", [0, 1])
def test_receivers_brute_force(oracle, W, H, seed, preset):
    p = oracle.preset(preset)
    h = _random_h(W, H, seed)
    # quantise to create exact ties between directions
    h = np.round(h * 8) / 8
    np.testing.assert_array_equal(oracle.receivers(h, W, H, p), _brute_receivers(h, W, H, p))


def _levels_by_walk(rec, W):
    off = OFFS(W)
    lvl = np.empty(rec.size, dtype=np.int64)
    for c in range(rec.size):
        d, x = 0, c
        while rec[x] != -1:
            x += off[rec[x]]
            d += 1
        lvl[c] = d
    return lvl


@pytest.mark.parametrize("W,H,seed,preset", [(31, 23, 5, 0), (64, 64, 6, 1), (20, 50, 7, 0)])
def test_order_and_accumulation_invariants(oracle, W, H, seed, preset):
    p = oracle.preset(preset)
    h = _random_h(W, H, seed)
    for _ in range(3):   # a few steps so the network develops
        h = oracle.step(h, W, H, p)["h"]
    st = oracle.step(h, W, H, p)
    rec, ndon, donor, stack, levels = st["rec"], st["ndon"], st["donor"], st["stack"], st["levels"]
    off = OFFS(W)
    # donors are exactly the inverse of rec, in ascending index
    inv = [[] for _ in range(W * H)]
    for c in range(W * H):
        if rec[c] != -1:
            inv[c + off[rec[c]]].append(c)
    for c in range(W * H):
        assert list(donor[c, :ndon[c]]) == inv[c]
    assert ndon.sum() == np.count_nonzero(rec != -1)
    # the stack holds every cell once; levels match the hop count to an outlet
    assert np.array_equal(np.sort(stack), np.arange(W * H))
    lvl = _levels_by_walk(rec, W)
    nl = levels.size
    assert levels[0] == 0 and levels[-1] == W * H and levels[-2] == W * H
    for L in range(nl - 1):
        assert np.all(lvl[stack[levels[L]:levels[L + 1]]] == L)
    assert nl == lvl.max() + 3
    # level 0 is the NO_FLOW cells in ascending index (ring 0 included)
    assert np.array_equal(stack[:levels[1]], np.flatnonzero(rec == -1))
    # accumulation = cell_area x number of cells whose path passes through c
    upstream = np.zeros(W * H, dtype=np.int64)
    for c in range(W * H):
        x = c
        upstream[x] += 1
        while rec[x] != -1:
            x += off[rec[x]]
            upstream[x] += 1
    np.testing.assert_array_equal(st["accum"], upstream * p.cell_area)


@pytest.mark.parametrize("preset", [0, 1])
def test_erosion_closed_form(oracle, preset):
    """For n = 1 the implicit update has the fixed point
    h = (h0 + F*hn)/(1 + F); the Newton replay must agree to 1e-12.  For
    n = 2 the converged h solves h - h0 + F*(h-hn)^2 = 0 to the tolerance."""
    W, H = 48, 40
    p = oracle.preset(preset)
    h = _random_h(W, H, 11) * 50
    st = oracle.step(h, W, H, p)
    rec, stack, accum, hup, hnew = st["rec"], st["stack"], st["accum"], st["hup"], st["h"]
    off = OFFS(W)
    d = [p.dx, p.dx * SQRT2] * 4
    checked = 0
    for c in stack:
        if rec[c] == -1:
            continue
        n = c + off[rec[c]]
        F = p.keq * p.dt * accum[c] ** p.meq / d[rec[c]] ** p.neq
        if p.neq == 1:
            expect = (hup[c] + F * hnew[n]) / (1 + F)
            assert abs(hnew[c] - expect) <= 1e-12 * max(1.0, abs(expect))
        else:
            resid = hnew[c] - hup[c] + F * (hnew[c] - hnew[n]) ** 2
            assert abs(resid) < 1e-2
            assert hnew[n] <= hnew[c] + 1e-9
        checked += 1
    assert checked > 100


def test_exact_pow_mode_equals_libm_for_n1(oracle):
    """pow(x,1) and pow(x,0) are exact in libm, so the product form used on
    the GPU (x, 1.0) changes nothing for n = 1."""
    W = H = 64
    p = oracle.preset(1)
    h = oracle.init_terrain(W, H, 0)
    for _ in range(5):
        a = oracle.step(h, W, H, p, oracle.POW_LIBM)["h"]
        b = oracle.step(h, W, H, p, oracle.POW_EXACT)["h"]
        np.testing.assert_array_equal(a, b)
        h = a


def test_parallel_erosion_equals_serial(oracle):
    """fastscape_RB+P's level-parallel erosion (P:313-351) gives RB's bits."""
    W, H = 200, 150
    p = oracle.preset(0)
    h = oracle.init_terrain(W, H, 0)
    for _ in range(15):
        st = oracle.step(h, W, H, p)
        par = oracle.erode_levels_omp(st["stack"], st["levels"], st["rec"], st["accum"],
                                      st["hup"], W, H, p)
        np.testing.assert_array_equal(par, st["h"])
        h = st["h"]


def test_planted_direction_swap_is_caught(oracle):
    """A plausible mistake (directions NE and SW swapped in the receiver
    offsets) must break the inverse-donor invariant used above."""
    W, H = 30, 30
    p = oracle.preset(1)
    h = _random_h(W, H, 3)
    rec = oracle.receivers(h, W, H, p)
    bad = rec.copy()
    bad[rec == 3], bad[rec == 7] = 7, 3
    ndon, donor = oracle.donors(bad, W, H)
    off = OFFS(W)
    ok = True
    for c in range(W * H):
        if rec[c] != -1 and c not in donor[c + off[rec[c]], :ndon[c + off[rec[c]]]]:
            ok = False
            break
    assert not ok
