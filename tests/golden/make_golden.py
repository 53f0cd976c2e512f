#!/usr/bin/env python3
"""Generate the golden fixtures in tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_rb (the reference fastscape_RB.cpp compiled from its own
sources by `make -C oracle ref`, driven by oracle/ref_harness.cpp) and stores
its outputs as small compressed .npz / .json fixtures:

  stages_<name>.npz   per-stage arrays of the first time-steps: h_init, and
                      for each step s: rec_s, ndon_s, donor_s, stack_s,
                      levels_s, accum_s, hup_s (after uplift), h_s (after
                      erosion)
  final_<preset>_<W>x<H>_<nstep>.npz   elevations after run(nstep)
  dem_sha256.json     SHA-256 of the PrintDEM output of run(120) at 501x501
                      (the compare.sh configuration) for both presets

Only this script reads the reference; the fixtures are data.  Run from the
repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden")
REF_RB = os.path.join(ROOT, "oracle", "_ref", "ref_rb")

# (name, W, H, nsteps, cli_seed, preset)
STAGE_CASES = [
    ("ref_8x8", 8, 8, 3, 123, 0),
    ("base_17x13", 17, 13, 3, 123, 1),
    ("ref_13x17", 13, 17, 3, 123, 0),
    ("ref_64x64", 64, 64, 3, 123, 0),
    ("base_64x64", 64, 64, 3, 123, 1),
    ("base_5x5", 5, 5, 2, 123, 1),
]
FINAL_CASES = [(101, 101, 120, 0), (101, 101, 120, 1), (64, 48, 40, 1)]
DEM_CASES = [(501, 120, 0), (501, 120, 1)]


def run(args):
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def load(d, name, dtype, n=None):
    a = np.fromfile(os.path.join(d, name + ".bin"), dtype=dtype)
    return a if n is None else a[:n]


def stage_case(name, W, H, nsteps, seed, preset):
    with tempfile.TemporaryDirectory() as d:
        run([REF_RB, "dump", str(W), str(H), str(nsteps), str(seed), str(preset), d])
        out = dict(W=np.int32(W), H=np.int32(H), preset=np.int32(preset),
                   nsteps=np.int32(nsteps), h_init=load(d, "h_init", "<f8"))
        for s in range(nsteps):
            out[f"rec_{s}"] = load(d, f"rec_{s}", "<i4")
            out[f"ndon_{s}"] = load(d, f"ndon_{s}", "<i4")
            donor = load(d, f"donor_{s}", "<i4").reshape(-1, 8).copy()
            # slots beyond ndon are never written by the reference; zero them
            for k in range(8):
                donor[out[f"ndon_{s}"] <= k, k] = 0
            out[f"donor_{s}"] = donor
            out[f"stack_{s}"] = load(d, f"stack_{s}", "<i4")
            out[f"levels_{s}"] = load(d, f"levels_{s}", "<i4")
            out[f"accum_{s}"] = load(d, f"accum_{s}", "<f8")
            out[f"hup_{s}"] = load(d, f"hup_{s}", "<f8")
            out[f"h_{s}"] = load(d, f"h_{s}", "<f8")
    np.savez_compressed(os.path.join(OUT, f"stages_{name}.npz"), **out)


def final_case(W, H, nstep, preset):
    with tempfile.TemporaryDirectory() as d:
        run([REF_RB, "dump", str(W), str(H), str(nstep + 1), "123", str(preset), d])
        h = load(d, f"h_{nstep}", "<f8")
        levels = load(d, f"levels_{nstep}", "<i4")
    tag = "ref" if preset == 0 else "base"
    np.savez_compressed(os.path.join(OUT, f"final_{tag}_{W}x{H}_{nstep}.npz"),
                        W=np.int32(W), H=np.int32(H), nstep=np.int32(nstep),
                        preset=np.int32(preset), h=h, nlevel_last=np.int32(len(levels)))


def dem_case(n, nstep, preset):
    with tempfile.TemporaryDirectory() as d:
        dem = os.path.join(d, "out.dem")
        run([REF_RB, "run", str(n), str(n), str(nstep), "123", str(preset), dem])
        with open(dem, "rb") as f:
            data = f.read()
    return hashlib.sha256(data).hexdigest(), len(data)


def main():
    if not os.path.exists(REF_RB):
        sys.exit("build the reference harness first: make -C oracle ref")
    for case in STAGE_CASES:
        stage_case(*case)
        print("stages", case[0])
    for case in FINAL_CASES:
        final_case(*case)
        print("final", case)
    shas = {}
    for n, nstep, preset in DEM_CASES:
        tag = "ref" if preset == 0 else "base"
        sha, size = dem_case(n, nstep, preset)
        shas[f"{tag}_{n}x{n}_{nstep}"] = {"sha256": sha, "bytes": size,
                                          "W": n, "H": n, "nstep": nstep,
                                          "preset": preset, "cli_seed": 123}
        print("dem", tag, sha)
    with open(os.path.join(OUT, "dem_sha256.json"), "w") as f:
        json.dump(shas, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
