"""File-level drivers for ensembles and parameter inversion.

    python -m fastscape_amd.drivers ensemble --manifest in.json --steps 500 --outdir out/
    python -m fastscape_amd.drivers ensemble --sample 256 --steps 500 --outdir out/
    python -m fastscape_amd.drivers inversion --target target.dem --seed 900001 --out result.json

Ensemble manifest: a JSON list of realizations, each
{"width": W, "height": H (or "side": S), "seed": engine seed,
 "preset": "benchmark" | "reference" (default benchmark),
 and any of "K", "n", "m", "U", "dt", "dx", "cell_area", "tol" overriding it}.
Each realization's final elevations are written as out/r<index>.dem in the
reference's PrintDEM format (fastscape_RB.cpp:20-40), and out/manifest.json
records the parameters used, the deepest level of the last step and the
steps taken by each path.

Inversion: the target .dem is the exact solver's output for hidden (K, U, m)
from the terrain of `--seed`; the driver trains the proposal network on
exact runs, ranks the candidate grid, reruns the best `--top-k` exactly and
writes the winner with its exact misfit (inversion.py defines the misfit).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from ._lib import PRESET_BENCHMARK, PRESET_REFERENCE, preset, write_dem

_KEYS = {"K": "keq", "n": "neq", "m": "meq", "U": "ueq", "dt": "dt", "dx": "dx",
         "cell_area": "cell_area", "tol": "tol"}


def realization_from_entry(e):
    from .ensemble import realization
    which = {"benchmark": PRESET_BENCHMARK, "reference": PRESET_REFERENCE}[e.get("preset",
                                                                               "benchmark")]
    p = preset(which, **{_KEYS[k]: float(v) for k, v in e.items() if k in _KEYS})
    W = int(e.get("width", e.get("side", 0)))
    H = int(e.get("height", e.get("side", 0)))
    return realization(W, H, p, int(e.get("seed", 0)))


def run_ensemble(entries, steps, outdir, autotune=1, log=print):
    """run(steps - 1) of every realization; writes r<i>.dem and manifest.json."""
    from .ensemble import Ensemble
    os.makedirs(outdir, exist_ok=True)
    reals = [realization_from_entry(e) for e in entries]
    with Ensemble(reals) as ens:
        if autotune:
            ens.autotune(autotune, 100)
        t0 = time.perf_counter()
        ens.run(steps - 1)
        ens.synchronize()
        secs = time.perf_counter() - t0
        out = []
        for i, (e, r) in enumerate(zip(entries, reals)):
            path = os.path.join(outdir, "r%d.dem" % i)
            write_dem(path, ens.getH(i), r.width, r.height)
            tile, lev = ens.path_stats(i)
            out.append(dict(e, index=i, width=r.width, height=r.height,
                            params=r.params.as_dict(), dem=os.path.basename(path),
                            deepest_level=int(ens.levels(i).size) - 3, tile_path_steps=tile,
                            level_path_steps=lev))
        cells = ens.cells
    summary = {"steps": steps, "seconds": secs, "cell_steps_per_s": cells * steps / secs,
               "realizations": out}
    with open(os.path.join(outdir, "manifest.json"), "w") as f:
        json.dump(summary, f, indent=1)
    log("%d realizations, %d steps, %.2f s (%.3g cell-steps/s) -> %s"
        % (len(out), steps, secs, summary["cell_steps_per_s"], outdir))
    return summary


def sample_entries(count, master_seed=0, sides=(512, 768, 1024, 1536, 2048)):
    from .ensemble import config4_realizations
    _, meta = config4_realizations(count, master_seed, sides)
    return [{"side": d["side"], "K": d["K"], "U": d["U"], "m": d["m"], "n": 1.0,
             "seed": d["seed"]} for d in meta]


def run_inversion(target_dem, seed, steps, train_runs, epochs, axis, top_k, log=print):
    from . import inversion as inv
    from .demtools import read_dem
    _, inner = read_dem(target_dem)
    side = inner.shape[0] + 2
    if inner.shape[1] + 2 != side or side % 32:
        raise ValueError("the target must be a square grid whose side is a multiple of 32")
    h = np.zeros((side, side))
    h[1:-1, 1:-1] = inner
    target_t = inv.transform(h)
    terrain, pn, t, _, gen_s = inv.make_training_set(train_runs, side, steps, seed0=1, log=log)
    prop = inv.Proposal(side, max_batch=32, seed=0)
    try:
        hist = prop.train(terrain, pn, t, epochs=epochs, batch=min(32, train_runs))
        cands = inv.candidate_grid(axis)
        best, mis, top, pred, secs = inv.proposal_search(prop, target_t, seed, cands, side, steps,
                                                         top_k=top_k)
    finally:
        prop.close()
    K, U, m = cands[best]
    res = {"K": K, "U": U, "m": m, "exact_misfit": float(mis.min()), "search_seconds": secs,
           "training_runs": train_runs, "training_solver_seconds": gen_s,
           "final_training_loss": float(hist[-1]), "candidates": len(cands), "top_k": top_k}
    log("best K %.4g U %.4g m %.3f, exact misfit %.5f (%.2f s)" % (K, U, m, res["exact_misfit"],
                                                                 secs))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description="FastScape ensemble and inversion drivers")
    sub = ap.add_subparsers(dest="cmd", required=True)
    e = sub.add_parser("ensemble")
    g = e.add_mutually_exclusive_group(required=True)
    g.add_argument("--manifest")
    g.add_argument("--sample", type=int)
    e.add_argument("--master-seed", type=int, default=0)
    e.add_argument("--steps", type=int, default=500)
    e.add_argument("--outdir", required=True)
    e.add_argument("--autotune", type=int, default=1)
    i = sub.add_parser("inversion")
    i.add_argument("--target", required=True)
    i.add_argument("--seed", type=int, required=True)
    i.add_argument("--steps", type=int, default=2000)
    i.add_argument("--train-runs", type=int, default=2048)
    i.add_argument("--epochs", type=int, default=4)
    i.add_argument("--candidates-axis", type=int, default=16)
    i.add_argument("--top-k", type=int, default=32)
    i.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "ensemble":
        if a.manifest:
            with open(a.manifest) as f:
                entries = json.load(f)
        else:
            entries = sample_entries(a.sample, a.master_seed)
        run_ensemble(entries, a.steps, a.outdir, a.autotune)
        return 0
    res = run_inversion(a.target, a.seed, a.steps, a.train_runs, a.epochs, a.candidates_axis,
                        a.top_k)
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
