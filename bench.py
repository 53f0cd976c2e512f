#!/usr/bin/env python3
"""Benchmark: FastScape time-steps on one MI355X.

A step is one pass of the time-step (receivers, donors, order, flow
accumulation, uplift, erosion: fastscape_RB.cpp:366-376) over the whole grid.
Default workload: 10000 x 10000 cells, FP64, benchmark preset (K=1e-5, n=1,
m=0.4, U=2e-3, dt=1000, dx=dy=100), terrain from mt19937 seeded 0 (the
reference's terrain for any nonzero seed), state resident in HBM before the
timed region.

    python bench.py [--gpus 1] [--steps 121] [--warmup 5] [--size 10000]
    python bench.py --workload ensemble [--steps 500]     # 256 realizations
    python bench.py --workload inversion                  # learned proposal vs exact-only


Prints one JSON line.  `value` is cell-steps/s over the timed steps;
`roofline` is the dominant kernel's algorithmic bytes / its HIP-event time;
`cpu_baseline` times the reference's RB+P (OpenMP) on a bounded sample on
this host.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "barnes2019-landscape_amd"))

METRIC = "cell-steps/s at 10000² FP64 vs RB+P CPU; exact-vs-proposal inversion wall time"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per processed cell for each kernel (reads + writes
# of the minimal data the kernel must touch; neighbour re-reads excluded).
KERNEL_BYTES = {
    "k_receivers": 9,      # h 8 + rcv 1
    "k_topology": 4,       # rcv 1 + donor mask 1 + level 2
    "k_bfs": 9,            # stack entry 4 + donor mask 1 + donor index out 4
    "k_flowacc": 13,       # stack 4 + mask 1 + donor counts 4 (one donor/cell) + count 4
    "k_erode": 33,         # stack 4 + rcv 1 + count 4 + h 8 + receiver h 8 + h out 8
    # tile path, per cell, at the per-stage algorithmic figures of a whole
    # time-step (79 B/cell-step in total: receivers 9, donors 2, order 10,
    # accumulation 21, uplift + erosion 37): donors + order + accumulation 33;
    # uplift + erosion 37
    "k_tile_acc": 33,
    "k_tile_erode": 37,
    # tile-graph path (drainage of any depth), per cell: k_deep_local reads
    # the receiver byte and writes the local count and the tail/hops word;
    # k_deep_counts reads those, the receiver and the elevation, writes the
    # count and, for the root-tailed cells (most of a tile), the new elevation
    "k_deep_local": 9,
    "k_deep_counts": 29,
}


# HBM bytes per cell actually moved by one launch (FETCH_SIZE x 2.0 + WRITE_SIZE,
# calibrated against a 1 GiB copy) at 10000 x 10000 with the benchmark preset:
# profiles/tile_path_10000x10000_hbm_traffic.md.  Reported as `traffic` for
# that workload only.
MEASURED_TRAFFIC_PER_CELL = {
    "k_tile_erode": 23.89,
    "k_tile_acc": 8.61,
    "k_receivers": 10.42,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["grid", "ensemble", "inversion", "unet"],
                    default="grid",
                    help="grid: one 10000^2 grid (default); ensemble: 256 realizations of "
                         "sides 512..2048 with sampled K, U, m; inversion: learned proposal "
                         "vs exact-only search at 256^2, 2000 steps")
    ap.add_argument("--train-runs", type=int, default=20480)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--candidates-axis", type=int, default=16)
    ap.add_argument("--top-k", type=int, default=32)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (grid: 121; ensemble: 500)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--realizations", type=int, default=256)
    ap.add_argument("--tune-trial", type=int, default=1,
                    help="ensemble: steps per autotuner candidate (0 = no tuning)")
    ap.add_argument("--tune-interval", type=int, default=100)
    ap.add_argument("--size", type=int, default=10000)
    ap.add_argument("--preset", choices=["benchmark", "reference"], default="benchmark",
                    help="grid workload parameters: the benchmark preset (K=1e-5, n=1, m=0.4, "
                         "dx=100) or the reference's own (K=2e-6, n=2, m=0.8, dx=1), whose "
                         "drainage deepens past the tile halos within tens of steps")
    ap.add_argument("--advance", type=int, default=0,
                    help="grid workload: untimed steps before the warmup, so the timed window "
                         "starts from a deeper drainage network (at 10000^2, 300 gives 17-32 "
                         "levels: the 32-cell-halo tile path)")
    ap.add_argument("--profile-steps", type=int, default=10,
                    help="steps timed per kernel with HIP events for the roofline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-size", type=int, default=10000)
    ap.add_argument("--cpu-sample-run", type=int, default=2,
                    help="run(n) of the CPU sample, i.e. n+1 time-steps")
    ap.add_argument("--cpu-reps", type=int, default=3,
                    help="repetitions of the CPU sample (the median is reported)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a 1 GiB plain copy kernel (counter calibration)")
    return ap.parse_args()


def host_cpu():
    """CPU model, physical cores (distinct (package, core) pairs) and the CPUs
    this process may run on."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return model, (len(cores) or None), allowed


def _timed_runs(cmd, env, reps):
    secs = []
    for _ in range(reps):
        out = subprocess.run(cmd, check=True, capture_output=True, text=True, env=env).stdout
        secs.append(float(re.search(r"m run_seconds = ([0-9.eE+-]+)", out).group(1)))
    return secs


def cpu_baseline(size, nrun, reps=3, serial_nrun=0, serial_reps=3, serial=True):
    """The reference's RB+P (fastscape_RB+P.cpp, OpenMP) built from the
    reference sources under oracle/_ref, timed as tests.sh does
    (tests/tests.sh:45-48): `reps` repetitions of run(nrun) on the benchmark
    preset, median taken.  Threads: OMP_NUM_THREADS when set (the CPUs this
    host grants per GPU), else every CPU this process may use.  With
    `serial`, also the serial RB (fastscape_RB.cpp) on one thread, median of
    `serial_reps` runs of run(serial_nrun).  Falls back to the oracle's own
    restatements (kind "port") when the reference build is absent."""
    import statistics
    model, physical, allowed = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or allowed
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    kind = "reference" if os.path.exists(os.path.join(ref_dir, "ref_rbp")) else "port"

    def cmd(par, n):
        if kind == "reference":
            exe = os.path.join(ref_dir, "ref_rbp" if par else "ref_rb")
            return [exe, "run", str(size), str(size), str(n), "123", "1", "-"]
        return [os.path.join(ROOT, "oracle", "oracle_cli"), str(size), str(size), str(n), "0",
                "1", "0", "1" if par else "0", "-"]

    secs = _timed_runs(cmd(True, nrun), env, reps)
    med = statistics.median(secs)
    cells = size * size * (nrun + 1)
    line = {"value": cells / med, "unit": "cell-steps/s", "cores": threads, "kind": kind,
            "sample": "%dx%d benchmark preset, run(%d) = %d time-steps, median of %d: %s s, %s"
                      % (size, size, nrun, nrun + 1, reps,
                         "/".join("%.2f" % x for x in secs),
                         "fastscape_RB+P.cpp" if kind == "reference"
                         else "oracle RB+P restatement"),
            "cpu_model": model, "cores_physical": physical, "cpus_allowed": allowed,
            "repetitions": reps, "seconds": secs}
    if serial:
        one = dict(env, OMP_NUM_THREADS="1")
        ssecs = _timed_runs(cmd(False, serial_nrun), one, serial_reps)
        smed = statistics.median(ssecs)
        line["serial_rb_1_thread"] = {
            "value": size * size * (serial_nrun + 1) / smed, "unit": "cell-steps/s",
            "cores": 1, "seconds": ssecs,
            "sample": "%dx%d benchmark preset, run(%d), median of %d, %s"
                      % (size, size, serial_nrun, serial_reps,
                         "fastscape_RB.cpp" if kind == "reference" else "oracle RB restatement")}
    return line


def bench_ensemble(args, fa):
    """The ensemble workload: `--realizations` grids with square sides drawn
    from {512, 768, 1024, 1536, 2048}, K log-uniform [1e-6, 1e-4], U
    log-uniform [5e-4, 5e-3], m uniform [0.3, 0.6], n = 1, terrain seed =
    realization index; every step is one persistent launch over all of them,
    then one tile-graph step (nine launches) over every realization whose
    drainage is deeper than 16 levels.  The autotuner runs inside the warmup
    and every --tune-interval steps inside the timed region (its trial steps
    are real steps)."""
    steps = args.steps or 500
    reals, meta = fa.config4_realizations(args.realizations, master_seed=0)
    ens = fa.Ensemble(reals)
    cells = ens.cells
    ens.autotune(args.tune_trial, args.tune_interval)
    for _ in range(args.warmup):
        ens.step()
    ens.synchronize()
    s0, ms0 = ens.stats()
    t0 = time.perf_counter()
    ens.run(steps - 1)          # run(n) = n+1 time-steps
    ens.synchronize()
    t1 = time.perf_counter()
    s1, ms1 = ens.stats()
    secs = t1 - t0
    value = cells * steps / secs
    kernel_ms = ms1 - ms0
    per_step_bytes = 79 * cells   # 79 B per cell-step, every stage of the time-step
    achieved = per_step_bytes * steps / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_ens_step + the deep realizations' tile-graph kernels",
                "launches_per_step": None, "kernel_ms_per_step": kernel_ms / steps}
    paths = [ens.path_stats(r) for r in range(len(reals))]
    depth = [int(ens.levels(r).size) - 3 for r in range(len(reals))]
    sched = ens.get_schedule().as_dict()
    tune = ens.tune_log()
    ens.close()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(1024, 29, reps=3, serial_nrun=9)
    tiles = sum(t for t, _ in paths)
    lvls = sum(l for _, l in paths)
    line = {"metric": "ensemble cell-steps/s, FP64 (one persistent launch per step, deep "
                      "realizations batched on the tile graph)",
            "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": steps,
            "warmup": args.warmup, "ms_per_step": secs * 1e3 / steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (per-realization mt19937 terrains; parameters from PCG64 seed 0)",
            "config": {"workload": "ensemble of %d realizations, sides U{512..2048}, "
                                   "K logU[1e-6,1e-4], U logU[5e-4,5e-3], m U[0.3,0.6], n=1, "
                                   "dx=100, dt=1000, %d steps" % (len(reals), steps),
                       "cells": cells, "schedule_final": sched,
                       "realization_steps_tile_path": tiles,
                       "realization_steps_tile_graph": lvls,
                       "deepest_level_last_step": max(depth),
                       "median_level_last_step": sorted(depth)[len(depth) // 2],
                       "autotune_candidates_timed": len(tune)},
            "roofline": roofline, "cpu_baseline": cpu}
    if cpu:
        line["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(line))


def bench_inversion(args, fa):
    """Parameter inversion on 256 x 256 grids, 2000 time-steps per exact run:
    a training set of --train-runs exact runs (parameters from the ensemble
    workload's distribution, terrain seeds 1..n), the U-Net trained on it,
    then a hidden-parameter target (terrain seed 900001) searched over a
    16^3 candidate grid by (a) running every candidate exactly and (b)
    ranking with the network and running the top 32 exactly.  `value` is
    (b)'s wall time; the exact solver judges both results."""
    import numpy as np
    from fastscape_amd import inversion as inv
    side, nstep = 256, 2000

    def log(msg):
        print("[inversion] " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    terrain, pn, t, params, gen_s = inv.make_training_set(args.train_runs, side, nstep, seed0=1,
                                                          master_seed=0, log=log)
    gen_wall = time.perf_counter() - t0
    log("training set: %d exact runs, %.1f s in the solver, %.1f s wall"
        % (args.train_runs, gen_s, gen_wall))
    prop = inv.Proposal(side, max_batch=32, seed=0)
    hist = prop.train(terrain, pn, t, epochs=args.epochs, batch=32, lr=1e-3, seed=1)
    st = prop.net.stats()
    train_tflops = 3 * st["flops_per_sample"] * st["samples"] / (st["train_ms"] * 1e-3) / 1e12
    log("trained %d epochs in %.1f s (%.1f TFLOP/s), losses %s"
        % (args.epochs, st["train_ms"] / 1e3, train_tflops, np.array2string(hist, precision=4)))
    del terrain, t
    hidden = np.array([[2.3e-5, 1.7e-3, 0.44]])
    seed = 900001
    target = {}
    inv.exact_runs(hidden, seed, side, nstep, on_result=lambda i, h: target.update(h=h))
    tt = inv.transform(target["h"])
    cands = inv.candidate_grid(args.candidates_axis)
    best_e, mis_e, t_exact = inv.exact_only_search(tt, seed, cands, side, nstep)
    log("exact-only: %d candidates in %.2f s, best %d misfit %.5f"
        % (len(cands), t_exact, best_e, mis_e[best_e]))
    best_p, mis_p, top, pred, t_prop = inv.proposal_search(prop, tt, seed, cands, side, nstep,
                                                           top_k=args.top_k)
    log("proposal: top %d in %.2f s, best %d misfit %.5f"
        % (args.top_k, t_prop, best_p, mis_p.min()))
    # inference alone, for the MFMA roofline
    terrain1 = fa.generate_terrain(side, side, seed).reshape(side, side).astype(np.float32)
    pn_c = inv.normalise(cands[:, 0], cands[:, 1], cands[:, 2])
    prop.rank(terrain1, pn_c[:64], tt)
    ti = time.perf_counter()
    prop.rank(terrain1, pn_c, tt)
    t_inf = time.perf_counter() - ti
    inf_tflops = st["flops_per_sample"] * len(cands) / t_inf / 1e12
    rel = inv.relative_param_error(cands[best_p], cands[best_e])
    rank_corr = float(np.corrcoef(np.argsort(np.argsort(pred)),
                                  np.argsort(np.argsort(mis_e)))[0, 1])
    prop.close()
    cpu = None
    if not args.no_cpu_baseline:
        one = cpu_baseline(side, nstep - 1, reps=1, serial=False)
        per_run = side * side * nstep / one["value"]
        cpu = {"value": per_run * len(cands), "unit": "s (exact-only search on the CPU)",
               "cores": one["cores"], "kind": one["kind"],
               "sample": "one %dx%d run(%d) of the benchmark preset, %.2f s, times %d candidates"
                         % (side, side, nstep - 1, per_run, len(cands))}
    # one-off cost of the proposal: exact training runs and training; what it
    # buys per target: t_exact - t_prop
    train_s = st["train_ms"] / 1e3
    one_off = gen_wall + train_s
    gain = t_exact - t_prop
    break_even = one_off / gain if gain > 0 else None
    end_to_end = {
        "one_off_seconds": one_off,
        "training_set_wall_seconds": gen_wall,
        "training_seconds": train_s,
        "speedup_per_target_excluding_one_off": t_exact / t_prop,
        "speedup_single_target_including_one_off": t_exact / (one_off + t_prop),
        "break_even_targets": break_even,
        "speedup_amortised_over_100_targets": 100 * t_exact / (one_off + 100 * t_prop),
    }
    log("end to end: one-off %.1f s (training set %.1f s + training %.1f s); per target %.2f s vs "
        "%.2f s exact-only; single target %.3fx; break-even after %s targets"
        % (one_off, gen_wall, train_s, t_prop, t_exact, end_to_end[
            "speedup_single_target_including_one_off"],
           "%.1f" % break_even if break_even else "no"))
    line = {"metric": "exact-vs-proposal inversion wall time", "value": t_prop, "unit": "s",
            "n_gpus": 1, "steps": nstep, "warmup": 0, "ms_per_step": None,
            "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": "f64 solver, bf16 network (fp32 accumulate)",
            "data": "synthetic (mt19937 terrains; parameters from PCG64 seed 0)",
            "config": {"workload": "inversion: %d^3 candidates, top %d exact reruns, %dx%d, "
                                   "%d steps, network trained on %d exact runs x %d epochs"
                                   % (args.candidates_axis, args.top_k, side, side, nstep,
                                      args.train_runs, args.epochs),
                       "hidden_params": hidden[0].tolist(), "terrain_seed": seed},
            "exact_only_seconds": t_exact, "speedup_vs_exact_only": t_exact / t_prop,
            "end_to_end": end_to_end,
            "result": {"exact_only_best": cands[best_e].tolist(),
                       "proposal_best": cands[best_p].tolist(),
                       "same_candidate": bool(best_p == best_e),
                       "relative_param_error": rel.tolist(),
                       "misfit_exact_only": float(mis_e[best_e]),
                       "misfit_proposal": float(mis_p.min()),
                       "rank_correlation_predicted_vs_exact": rank_corr},
            "training": {"runs": args.train_runs, "solver_seconds": gen_s,
                         "train_seconds": st["train_ms"] / 1e3, "epochs": args.epochs,
                         "final_loss": float(hist[-1]), "tflops": train_tflops},
            "roofline": {"bound": "mfma", "achieved": round(inf_tflops, 2), "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": round(inf_tflops / 2500.0, 4),
                         "traffic": None, "kernel": "U-Net inference (k_conv3x3)",
                         "inference_seconds": t_inf},
            "cpu_baseline": cpu}
    print(json.dumps(line))


def bench_unet(args, fa):
    """Throughput of the learned proposal's network alone: BF16 MFMA forward
    passes (inference) and training steps (forward + backward + Adam) on
    synthetic 256 x 256 inputs, 32 samples per batch."""
    import numpy as np
    side, batch, n = 256, 32, 256
    rng = np.random.default_rng(0)
    terrain = rng.random((n, side, side), dtype=np.float32)
    params = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    target = rng.normal(0, 1, (n, side, side)).astype(np.float32)
    with fa.UNet(side, side, max_batch=batch, seed=0) as net:
        fl = net.stats()["flops_per_sample"]
        net.forward(terrain[:batch], params[:batch])
        t0 = time.perf_counter()
        net.forward(terrain, params)
        t_inf = time.perf_counter() - t0
        net.set_data(terrain, params, target)
        net.train(1, batch)
        st0 = net.stats()
        net.train(args.epochs, batch)
        st1 = net.stats()
    inf = fl * n / t_inf / 1e12
    tr_ms = st1["train_ms"] - st0["train_ms"]
    tr_samples = st1["samples"] - st0["samples"]
    tr = 3 * fl * tr_samples / (tr_ms * 1e-3) / 1e12
    line = {"metric": "U-Net inference TFLOP/s (BF16 MFMA, FP32 accumulate)", "value": inf,
            "unit": "TFLOP/s", "n_gpus": 1, "steps": n // batch, "warmup": 1,
            "ms_per_step": t_inf * 1e3 / (n // batch), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic (uniform terrain, uniform parameters, normal targets)",
            "config": {"workload": "4-level U-Net 32/64/128/160, %dx%d, batch %d, %d samples"
                                   % (side, side, batch, n),
                       "gflop_per_sample_forward": fl / 1e9,
                       "training_tflops": tr, "training_samples": tr_samples},
            "roofline": {"bound": "mfma", "achieved": round(inf, 2), "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": round(inf / 2500.0, 4), "traffic": None,
                         "kernel": "U-Net forward (k_conv3x3)"},
            "cpu_baseline": None}
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; --gpus %d is not supported "
              "(a multi-GPU run would shard ensembles by whole realizations and a single "
              "grid by row strips with a 16-row halo)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import fastscape_amd as fa

    if args.calibrate:
        ms = fa.calibrate_copy(1 << 30)
        print("calibration copy: 1 GiB in %.3f ms (%.1f GB/s read+write)"
              % (ms, 2 * (1 << 30) / ms / 1e6), file=sys.stderr)
    if args.workload == "ensemble":
        bench_ensemble(args, fa)
        return
    if args.workload == "inversion":
        bench_inversion(args, fa)
        return
    if args.workload == "unet":
        bench_unet(args, fa)
        return
    args.steps = args.steps or 121
    W = H = args.size
    N = W * H
    which = fa.PRESET_BENCHMARK if args.preset == "benchmark" else fa.PRESET_REFERENCE
    m = fa.FastScape(W, H, fa.preset(which), engine_seed=0, device=0)
    m.step(args.advance + args.warmup)
    m.synchronize()
    t0 = time.perf_counter()
    m.step(args.steps)
    m.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = N * args.steps / secs

    # per-kernel HIP-event timing over a further window of steps
    m.set_kernel_timing(True)
    m.step(args.profile_steps)
    stats = m.kernel_stats()
    m.set_kernel_timing(False)
    levels = m.levels()
    paths = m.path_stats()
    paths.update(m.tile_stats())
    n0 = int(levels[1]) if levels.size > 1 else N
    units = {"k_receivers": (W - 4) * (H - 4), "k_topology": N, "k_bfs": N - n0,
             "k_flowacc": N, "k_erode": N - n0, "k_tile_acc": N, "k_tile_erode": N,
             "k_deep_local": N, "k_deep_counts": N}
    dom = max(stats, key=lambda k: stats[k][0])
    ms, launches = stats[dom]
    per_step_bytes = KERNEL_BYTES.get(dom, 0) * units.get(dom, N)
    achieved = per_step_bytes * args.profile_steps / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = None
    if (args.size == 10000 and args.advance == 0 and args.preset == "benchmark"
            and dom in MEASURED_TRAFFIC_PER_CELL):
        traffic = MEASURED_TRAFFIC_PER_CELL[dom] * N
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom, "launches_per_step": launches / args.profile_steps,
                "kernel_ms_per_step": ms / args.profile_steps}
    m.close()

    cpu = None
    if not args.no_cpu_baseline and args.preset == "benchmark":   # timed on that preset
        cpu = cpu_baseline(args.cpu_sample_size, args.cpu_sample_run, reps=args.cpu_reps,
                           serial_nrun=0, serial_reps=args.cpu_reps)

    line = {"metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (mt19937 seed 0 uniform[0,1) terrain, rings 0/1 at 0)",
            "config": {"workload": ("FastScape %dx%d FP64 time-steps, K=1e-5 n=1 m=0.4 U=2e-3 "
                                    "dt=1000 dx=100" if args.preset == "benchmark" else
                                    "FastScape %dx%d FP64 time-steps, reference preset K=2e-6 "
                                    "n=2 m=0.8 U=2e-3 dt=1000 dx=1") % (W, H),
                       "grid": [W, H], "first_timed_step": args.advance + args.warmup,
                       "nlevel_last_step": int(levels.size),
                       "step_paths": paths},
            "roofline": roofline, "cpu_baseline": cpu,
            "stage_kernel_ms_per_step": {k: round(v[0] / args.profile_steps, 4)
                                         for k, v in stats.items()}}
    if cpu:
        line["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
