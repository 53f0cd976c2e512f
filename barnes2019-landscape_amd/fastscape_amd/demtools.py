"""`.dem` tools: a comparator in the role of RichDEM's rd_compare, which the
reference's compare.sh calls (compare.sh:14-23), and a compare.sh-style
harness over executables that follow the reference's argv contract
`<exe> <Dimension> <Steps> <Output Name> <Seed>` (fastscape_RB.cpp:418-427).

    python -m fastscape_amd.demtools compare A.dem B.dem [--atol X] [--rtol Y]
    python -m fastscape_amd.demtools harness --size 501 --steps 120 --seed 123 \\
        --exe RB=./fastscape_RB.exe --exe MI355X="lib/fastscape_mi355x"

Files are ArcGIS ASCII grids as PrintDEM writes them (fastscape_RB.cpp:20-40).
"""
import argparse
import os
import shlex
import subprocess
import sys
import tempfile

import numpy as np

HEADER_KEYS = ("ncols", "nrows", "xllcorner", "yllcorner", "cellsize", "NODATA_value")


def read_dem(path):
    """(header dict, values as a float64 array of shape (nrows, ncols))."""
    header = {}
    with open(path, "r") as f:
        for key in HEADER_KEYS:
            line = f.readline().split()
            if len(line) != 2 or line[0] != key:
                raise ValueError("%s: expected '%s <value>' header line" % (path, key))
            header[key] = line[1]
        data = np.loadtxt(f, dtype=np.float64, ndmin=2)
    ncols, nrows = int(header["ncols"]), int(header["nrows"])
    if data.shape != (nrows, ncols):
        raise ValueError("%s: %s values for a %dx%d grid" % (path, data.shape, nrows, ncols))
    return header, data


def compare(path_a, path_b, atol=0.0, rtol=0.0):
    """Cell-by-cell comparison.  Returns a dict with `same_bytes`, `equal`
    (every |a-b| <= atol + rtol*|b|), `max_abs`, `max_rel`, `ndiff` (cells
    that differ at all) and `nbad` (cells outside the tolerance)."""
    with open(path_a, "rb") as f:
        ba = f.read()
    with open(path_b, "rb") as f:
        bb = f.read()
    ha, a = read_dem(path_a)
    hb, b = read_dem(path_b)
    if (ha["ncols"], ha["nrows"]) != (hb["ncols"], hb["nrows"]):
        raise ValueError("grid sizes differ: %sx%s vs %sx%s" % (ha["ncols"], ha["nrows"],
                                                               hb["ncols"], hb["nrows"]))
    d = np.abs(a - b)
    rel = d / np.maximum(np.abs(b), np.finfo(float).tiny)
    bad = d > atol + rtol * np.abs(b)
    return {"same_bytes": ba == bb, "equal": not bool(bad.any()), "max_abs": float(d.max()),
            "max_rel": float(rel.max()), "ndiff": int(np.count_nonzero(d)),
            "nbad": int(np.count_nonzero(bad)), "cells": int(a.size)}


def harness(size, steps, seed, exes, outdir, atol=0.0, rtol=0.0, log=print):
    """Run each (name, command) with the reference's argv contract and compare
    every output with the first one's, as compare.sh does."""
    outs = []
    for name, cmd in exes:
        out = os.path.join(outdir, "out_%s.dem" % name)
        argv = shlex.split(cmd) + [str(size), str(steps), out, str(seed)]
        r = subprocess.run(argv, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s failed (%d): %s" % (name, r.returncode, r.stderr.strip()))
        outs.append((name, out, r.stdout))
    rows = []
    ref_name, ref_out, _ = outs[0]
    for name, out, _ in outs:
        c = compare(ref_out, out, atol, rtol)
        rows.append((ref_name, name, c))
        log("%-12s vs %-12s  bytes %s  equal %s  max|d| %.3g  cells differing %d of %d"
            % (ref_name, name, "same" if c["same_bytes"] else "diff", c["equal"], c["max_abs"],
               c["ndiff"], c["cells"]))
    return rows, outs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("compare")
    c.add_argument("a")
    c.add_argument("b")
    c.add_argument("--atol", type=float, default=0.0)
    c.add_argument("--rtol", type=float, default=0.0)
    h = sub.add_parser("harness")
    h.add_argument("--size", type=int, default=501)
    h.add_argument("--steps", type=int, default=120)
    h.add_argument("--seed", type=int, default=123)
    h.add_argument("--exe", action="append", required=True, metavar="NAME=COMMAND")
    h.add_argument("--outdir", default=None)
    h.add_argument("--atol", type=float, default=0.0)
    h.add_argument("--rtol", type=float, default=0.0)
    a = ap.parse_args(argv)
    if a.cmd == "compare":
        r = compare(a.a, a.b, a.atol, a.rtol)
        print("bytes %s  equal %s  max|d| %.6g  max rel %.3g  differing %d  outside tolerance %d "
              "of %d" % ("same" if r["same_bytes"] else "diff", r["equal"], r["max_abs"],
                         r["max_rel"], r["ndiff"], r["nbad"], r["cells"]))
        return 0 if r["equal"] else 1
    exes = []
    for e in a.exe:
        name, _, cmd = e.partition("=")
        if not cmd:
            ap.error("--exe wants NAME=COMMAND")
        exes.append((name, cmd))
    outdir = a.outdir or tempfile.mkdtemp(prefix="dem_compare_")
    rows, _ = harness(a.size, a.steps, a.seed, exes, outdir, a.atol, a.rtol)
    return 0 if all(c["equal"] for _, _, c in rows) else 1


if __name__ == "__main__":
    sys.exit(main())
