"""The .dem comparator (the role of RichDEM's rd_compare in compare.sh) and
the compare.sh-style harness, on the CPU."""
import os
import stat
import sys

import numpy as np
import pytest

import fastscape_amd as fa
from fastscape_amd import demtools

from conftest import ROOT


def write(path, h, W, H):
    fa.write_dem(str(path), np.ascontiguousarray(h, dtype=np.float64), W, H)


def test_compare_identical_and_tolerances(tmp_path):
    rng = np.random.default_rng(0)
    W, H = 23, 17
    h = rng.random(W * H) * 100
    write(tmp_path / "a.dem", h, W, H)
    write(tmp_path / "b.dem", h, W, H)
    r = demtools.compare(tmp_path / "a.dem", tmp_path / "b.dem")
    assert r["same_bytes"] and r["equal"] and r["ndiff"] == 0 and r["cells"] == (W - 2) * (H - 2)
    h2 = h.copy()
    h2[5 * W + 7] += 0.5
    write(tmp_path / "c.dem", h2, W, H)
    r = demtools.compare(tmp_path / "a.dem", tmp_path / "c.dem")
    assert not r["same_bytes"] and not r["equal"] and r["ndiff"] == 1
    assert abs(r["max_abs"] - 0.5) < 1e-3
    assert demtools.compare(tmp_path / "a.dem", tmp_path / "c.dem", atol=0.6)["equal"]
    assert demtools.main(["compare", str(tmp_path / "a.dem"), str(tmp_path / "c.dem")]) == 1
    assert demtools.main(["compare", str(tmp_path / "a.dem"), str(tmp_path / "b.dem")]) == 0


def test_read_dem_values_match_writer(tmp_path):
    W, H = 9, 6
    h = np.arange(W * H, dtype=np.float64) * 1.25
    write(tmp_path / "a.dem", h, W, H)
    hdr, v = demtools.read_dem(tmp_path / "a.dem")
    assert hdr["ncols"] == str(W - 2) and hdr["nrows"] == str(H - 2)
    np.testing.assert_allclose(v, h.reshape(H, W)[1:-1, 1:-1], rtol=1e-6)


def test_size_mismatch_raises(tmp_path):
    write(tmp_path / "a.dem", np.zeros(36), 6, 6)
    write(tmp_path / "b.dem", np.zeros(49), 7, 7)
    with pytest.raises(ValueError):
        demtools.compare(tmp_path / "a.dem", tmp_path / "b.dem")


def test_harness_runs_argv_contract(tmp_path):
    # two stand-in executables with the reference's argv contract: one writes
    # the terrain, the other the terrain plus a perturbation
    pkg = os.path.join(ROOT, "barnes2019-landscape_amd")
    for name, bump in (("a", 0.0), ("b", 1.0)):
        script = tmp_path / ("exe_" + name)
        script.write_text(
            "#!%s\nimport sys\nsys.path.insert(0, %r)\nimport fastscape_amd as f\n"
            "n = int(sys.argv[1])\nh = f.generate_terrain(n, n, 0)\nh[n * 3 + 3] += %r\n"
            "f.write_dem(sys.argv[3], h, n, n)\n" % (sys.executable, pkg, bump))
        script.chmod(script.stat().st_mode | stat.S_IEXEC)
    rows, outs = demtools.harness(11, 1, 123, [("A", str(tmp_path / "exe_a")),
                                               ("B", str(tmp_path / "exe_b"))],
                                  str(tmp_path), log=lambda *a: None)
    assert rows[0][2]["same_bytes"] and not rows[1][2]["equal"]
    assert rows[1][2]["ndiff"] == 1
