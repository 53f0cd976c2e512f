"""CPU tests of the C-ABI library: it loads, exports every declared symbol,
and its host-side pieces (terrain generator, DEM writer, argument checks,
CLI usage, C++ drop-in header) behave like the reference's.  No GPU compute
is called here."""
import ctypes
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, STAGE_FIXTURES, golden_dem_shas, load_golden

import fastscape_amd as fa

HEADER = os.path.join(ROOT, "include", "fastscape.h")
CLI = os.path.join(ROOT, "barnes2019-landscape_amd", "lib", "fastscape_mi355x")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = fa.library()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(fa.EXPORTED) == names


def test_exported_symbols_in_dynamic_table():
    out = subprocess.run(["nm", "-D", "--defined-only", fa.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_symbols():
        assert name in syms, name


def test_presets():
    p = fa.preset(fa.PRESET_REFERENCE)
    assert (p.keq, p.neq, p.meq, p.ueq, p.dt, p.dx, p.cell_area, p.tol) == \
        (2e-6, 2.0, 0.8, 2e-3, 1000.0, 1.0, 40000.0, 1e-3)          # RB:58-65
    q = fa.preset(fa.PRESET_BENCHMARK)
    assert (q.keq, q.neq, q.meq, q.ueq, q.dt, q.dx, q.cell_area, q.tol) == \
        (1e-5, 1.0, 0.4, 2e-3, 1000.0, 100.0, 1e4, 1e-3)
    with pytest.raises(fa.FastScapeError):
        fa.preset(7)


@pytest.mark.parametrize("fx", STAGE_FIXTURES)
def test_terrain_generator_matches_reference(fx):
    g = load_golden(fx)
    W, H = int(g["W"]), int(g["H"])
    np.testing.assert_array_equal(fa.generate_terrain(W, H, 0), g["h_init"])


def test_terrain_generator_matches_oracle_large(oracle):
    np.testing.assert_array_equal(fa.generate_terrain(1000, 700, 0), oracle.init_terrain(1000, 700, 0))
    np.testing.assert_array_equal(fa.generate_terrain(333, 77, 99), oracle.init_terrain(333, 77, 99))
    # the seeds the ensemble and inversion workloads use
    for seed in (1, 255, 900001):
        np.testing.assert_array_equal(fa.generate_terrain(64, 96, seed),
                                      oracle.init_terrain(64, 96, seed))


def test_dem_writer_byte_identical(oracle, tmp_path):
    rng = np.random.default_rng(0)
    W, H = 37, 29
    h = rng.standard_normal(W * H) * 10.0 ** rng.integers(-8, 8, W * H)
    h[5] = -0.0
    h[6] = 0.0
    h[7] = 1e300
    fa.write_dem(str(tmp_path / "a.dem"), h, W, H)
    oracle.write_dem(str(tmp_path / "b.dem"), h, W, H)
    assert (tmp_path / "a.dem").read_bytes() == (tmp_path / "b.dem").read_bytes()


@pytest.mark.slow
def test_dem_writer_reproduces_reference_sha(oracle, tmp_path):
    """fs_write_dem on the oracle's run(120) at 501x501 (benchmark preset)
    gives the reference PrintDEM bytes (RB:20-40)."""
    info = golden_dem_shas()["base_501x501_120"]
    h = oracle.run(501, 501, 120, oracle.preset(1), 0)
    fa.write_dem(str(tmp_path / "x.dem"), h, 501, 501)
    assert hashlib.sha256((tmp_path / "x.dem").read_bytes()).hexdigest() == info["sha256"]


def test_create_rejects_bad_dimensions():
    lib = fa.library()
    m = ctypes.c_void_p()
    p = fa.preset(0)
    for W, H in [(4, 10), (10, 4), (0, 0), (-5, 7), (70000, 70000)]:
        assert lib.fs_create(W, H, ctypes.byref(p), 0, 0, ctypes.byref(m)) == 2
    assert lib.fs_create(10, 10, None, 0, 0, ctypes.byref(m)) == 1
    bad = fa.preset(0, dx=0.0)
    assert lib.fs_create(10, 10, ctypes.byref(bad), 0, 0, ctypes.byref(m)) == 1
    assert lib.fs_run(None, 1) == 1
    assert lib.fs_destroy(None) == 1
    assert lib.fs_write_dem(b"/nonexistent/x.dem", None, 10, 10) == 1
    h = np.zeros(100)
    assert lib.fs_write_dem(b"/nonexistent/dir/x.dem",
                            h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 10, 10) == 10
    assert lib.fs_strerror(5) == b"drainage deeper than FS_MAX_LEVELS"
    assert lib.fs_strerror(12345) == b"unknown status"


def test_cli_usage_matches_reference():
    """argc != 5 -> usage on stderr and exit -1 (RB:418-421)."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    r = subprocess.run([CLI, "10", "10"], capture_output=True, text=True)
    assert r.returncode == 255
    assert "Syntax:" in r.stderr and "<Dimension> <Steps> <Output Name> <Seed>" in r.stderr


def test_cpp_dropin_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "fastscape.hpp"\n'
        "int main(){ fastscape::seed_rand(123);"
        " if(false){ fastscape::FastScape tm(10,10); tm.run(1);"
        " fastscape::PrintDEM(\"x\", tm.getH(), 10, 10);} return 0; }\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                    str(src)], check=True)


def test_no_oracle_in_product_path():
    """The product sources never reference the oracle."""
    pkg = os.path.join(ROOT, "barnes2019-landscape_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".hpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "liboracle" not in text and "import oracle" not in text, f


def test_ensemble_argument_errors_without_device():
    lib = fa.library()
    e = ctypes.c_void_p()
    assert lib.fs_ensemble_create(None, 0, 0, ctypes.byref(e)) == 1
    bad = (fa.FsRealization * 2)(fa.realization(64, 64, fa.preset(fa.PRESET_BENCHMARK)),
                                 fa.realization(3, 64, fa.preset(fa.PRESET_BENCHMARK)))
    assert lib.fs_ensemble_create(bad, 2, 0, ctypes.byref(e)) == 2
    nan = fa.preset(fa.PRESET_BENCHMARK, keq=float("nan"))
    one = (fa.FsRealization * 1)(fa.realization(64, 64, nan))
    assert lib.fs_ensemble_create(one, 1, 0, ctypes.byref(e)) == 1
    assert lib.fs_ensemble_destroy(None) == 1
    assert lib.fs_ensemble_run(None, 1) == 1


def test_ensemble_workload_sampler():
    reals, meta = fa.config4_realizations(256)
    again, meta2 = fa.config4_realizations(256)
    assert meta == meta2
    assert [m["seed"] for m in meta] == list(range(256))
    assert {m["side"] for m in meta} <= {512, 768, 1024, 1536, 2048}
    K = np.array([m["K"] for m in meta])
    U = np.array([m["U"] for m in meta])
    mm = np.array([m["m"] for m in meta])
    assert K.min() >= 1e-6 and K.max() <= 1e-4
    assert U.min() >= 5e-4 * (1 - 1e-12) and U.max() <= 5e-3 * (1 + 1e-12)
    assert mm.min() >= 0.3 and mm.max() <= 0.6
    for r, m in zip(reals, meta):
        assert r.width == r.height == m["side"]
        assert r.params.neq == 1.0 and r.params.keq == m["K"] and r.params.dx == 100.0
