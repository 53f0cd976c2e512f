"""File-level drivers and compare.sh-style checks on the GPU: the ensemble
driver's .dem files, the drop-in CLI's outputs and metadata lines, compared
with the reference's golden results through the .dem comparator."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import fastscape_amd as fa
from fastscape_amd import demtools, drivers

from conftest import ROOT, golden_dem_shas, load_golden

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "barnes2019-landscape_amd", "lib", "fastscape_mi355x")


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def test_ensemble_driver_writes_reference_dems(tmp_path):
    note("ensemble driver")
    entries = [{"side": 101, "seed": 0, "preset": "benchmark"},
               {"width": 64, "height": 48, "seed": 0, "preset": "benchmark"},
               {"side": 96, "seed": 7, "K": 1e-4, "m": 0.6, "U": 5e-4}]
    summary = drivers.run_ensemble(entries, 121, str(tmp_path), log=note)
    assert len(summary["realizations"]) == 3
    # realization 0: the reference's run(120) at 101x101, benchmark preset
    g = load_golden("final_base_101x101_120.npz")
    fa.write_dem(str(tmp_path / "golden.dem"), g["h"].ravel(), 101, 101)
    r = demtools.compare(str(tmp_path / "golden.dem"), str(tmp_path / "r0.dem"))
    assert r["same_bytes"], r
    with open(tmp_path / "manifest.json") as f:
        man = json.load(f)
    assert man["steps"] == 121
    assert man["realizations"][2]["params"]["keq"] == 1e-4
    assert all(x["tile_path_steps"] + x["level_path_steps"] == 121 for x in man["realizations"])


def test_cli_harness_matches_reference_sha_and_ensemble(tmp_path):
    note("compare.sh-style harness")
    rows, outs = demtools.harness(501, 120, 123, [("MI355X_ref", CLI),
                                                  ("MI355X_ref2", CLI + " --device=0")],
                                  str(tmp_path), log=note)
    sha = hashlib.sha256(open(outs[0][1], "rb").read()).hexdigest()
    assert sha == golden_dem_shas()["ref_501x501_120"]["sha256"]
    assert all(c["same_bytes"] for _, _, c in rows)
    stdout = outs[0][2]
    for key in ("m cell_steps_per_s", "m nlevel_last_step", "m tile_path_steps",
                "m level_path_steps", "t Overall", "A FastScape"):
        assert key in stdout, key
    # the benchmark preset through the CLI equals the ensemble driver's output
    out = tmp_path / "cli_base.dem"
    subprocess.run([CLI, "--preset=benchmark", "64", "40", str(out), "5"], check=True,
                   capture_output=True)
    drivers.run_ensemble([{"side": 64, "seed": 0}], 41, str(tmp_path / "ens"), log=note)
    assert demtools.compare(str(out), str(tmp_path / "ens" / "r0.dem"))["same_bytes"]
