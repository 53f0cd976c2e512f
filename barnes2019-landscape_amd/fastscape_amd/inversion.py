"""Parameter inversion with a learned proposal.

Problem: a target topography was produced by the exact time-step (the
reference's, fastscape_RB.cpp:366-376) from a known initial terrain with
hidden (K, U, m); find the candidate set whose exact result fits it best.

* exact-only search: every candidate runs through the exact solver (batched
  in ensembles, one persistent launch per step) and the best misfit wins;
* proposal search: the U-Net predicts each candidate's topography, the
  candidates are ranked by predicted misfit, and only the top few run
  through the exact solver, which picks the winner -- the exact solver stays
  the judge of every reported result.

Misfit: root-mean-square difference of log(1 + h) over the interior cells
(rings 0 and 1 are held at 0 by the model).  Parameters enter the network
normalised to [-1, 1]: log10 K over [-6, -4], log10 U over
[log10 5e-4, log10 5e-3], m over [0.3, 0.6].
"""
import time

import numpy as np

from ._lib import PRESET_BENCHMARK, generate_terrain, preset
from .ensemble import Ensemble, realization
from .unet import UNet

LOGK = (-6.0, -4.0)
LOGU = (float(np.log10(5e-4)), float(np.log10(5e-3)))
MR = (0.3, 0.6)


def normalise(K, U, m):
    K, U, m = np.asarray(K, float), np.asarray(U, float), np.asarray(m, float)
    a = 2 * (np.log10(K) - LOGK[0]) / (LOGK[1] - LOGK[0]) - 1
    b = 2 * (np.log10(U) - LOGU[0]) / (LOGU[1] - LOGU[0]) - 1
    c = 2 * (m - MR[0]) / (MR[1] - MR[0]) - 1
    return np.stack([a, b, c], axis=-1).astype(np.float32)


def sample_params(n, rng):
    """(K, U, m) from the ensemble workload's distribution."""
    K = 10.0 ** rng.uniform(*LOGK, n)
    U = 10.0 ** rng.uniform(*LOGU, n)
    m = rng.uniform(*MR, n)
    return np.stack([K, U, m], axis=1)


def candidate_grid(per_axis=16):
    """per_axis^3 candidates on a regular grid in (log10 K, log10 U, m)."""
    k = np.linspace(*LOGK, per_axis)
    u = np.linspace(*LOGU, per_axis)
    mm = np.linspace(*MR, per_axis)
    g = np.stack(np.meshgrid(k, u, mm, indexing="ij"), axis=-1).reshape(-1, 3)
    return np.stack([10.0 ** g[:, 0], 10.0 ** g[:, 1], g[:, 2]], axis=1)


def transform(h):
    return np.log1p(np.maximum(h, 0.0))


def interior(a):
    return a[..., 2:-2, 2:-2]


def misfit(t_a, t_b):
    """RMS difference of transformed topographies over the interior."""
    d = interior(t_a) - interior(t_b)
    return np.sqrt(np.mean(d * d, axis=(-2, -1)))


def exact_runs(params, seeds, side, nstep, batch=1024, on_result=None, autotune=1, log=None):
    """Final elevations of one exact run per (K, U, m) row and terrain seed:
    run(nstep - 1), i.e. nstep time-steps, batched `batch` realizations per
    ensemble.  on_result(i, h) receives each result (float64 side x side);
    returns the device wall seconds spent in the runs."""
    params = np.asarray(params, float)
    seeds = np.broadcast_to(np.asarray(seeds, dtype=np.int64), (len(params),))
    secs = 0.0
    for b0 in range(0, len(params), batch):
        rows = range(b0, min(len(params), b0 + batch))
        reals = [realization(side, side, preset(PRESET_BENCHMARK, keq=params[i, 0],
                                                ueq=params[i, 1], meq=params[i, 2], neq=1.0),
                             int(seeds[i])) for i in rows]
        with Ensemble(reals) as ens:
            if autotune:
                ens.autotune(autotune, 500)
            t0 = time.perf_counter()
            ens.run(nstep - 1)
            ens.synchronize()
            secs += time.perf_counter() - t0
            for j, i in enumerate(rows):
                on_result(i, ens.getH(j))
        if log:
            log("exact runs %d/%d done (%.1f s in the solver)" % (rows[-1] + 1, len(params), secs))
    return secs


class Proposal:
    """The trained network and the target normalisation it was trained with."""

    def __init__(self, side, max_batch=16, seed=0):
        self.side = side
        self.net = UNet(side, side, max_batch=max_batch, seed=seed)
        self.mean, self.std = 0.0, 1.0
        self.history = None

    def close(self):
        self.net.close()

    def train(self, terrain, params_norm, t, epochs, batch, lr=1e-3, seed=0):
        """t: transformed exact topographies (n, side, side)."""
        self.mean = float(t.mean())
        self.std = float(t.std()) or 1.0
        self.net.set_data(terrain, params_norm, (t - self.mean) / self.std)
        self.history = self.net.train(epochs, batch, lr=lr, seed=seed)
        return self.history

    def rank(self, terrain1, params_norm, target_t):
        """Predicted misfit of each candidate (shared terrain) against target_t,
        computed on the device."""
        tn = ((np.asarray(target_t, np.float64) - self.mean) / self.std).astype(np.float32)
        return self.net.rank(terrain1, params_norm, tn).astype(np.float64) * self.std

    def predict(self, terrain, params_norm):
        """Predicted transformed topographies (n, side, side)."""
        y = self.net.forward(terrain, params_norm)
        return y.astype(np.float64) * self.std + self.mean


def make_training_set(count, side, nstep, seed0=1, master_seed=0, batch=1024, log=None):
    """`count` exact runs with parameters from the workload's distribution and
    terrain seeds seed0, seed0+1, ...; returns (terrain, params_norm, t,
    params, seconds)."""
    rng = np.random.default_rng(master_seed)
    params = sample_params(count, rng)
    seeds = np.arange(seed0, seed0 + count)
    t = np.empty((count, side, side), dtype=np.float32)

    def keep(i, h):
        t[i] = transform(h)

    secs = exact_runs(params, seeds, side, nstep, batch=batch, on_result=keep, log=log)
    terrain = np.empty((count, side, side), dtype=np.float32)
    for i, s in enumerate(seeds):
        terrain[i] = generate_terrain(side, side, int(s)).reshape(side, side)
    return terrain, normalise(params[:, 0], params[:, 1], params[:, 2]), t, params, secs


def exact_only_search(target_t, seed, candidates, side, nstep, batch=1024):
    """Run every candidate exactly; returns (best index, misfits, seconds)."""
    mis = np.empty(len(candidates))

    def keep(i, h):
        mis[i] = misfit(transform(h), target_t)

    t0 = time.perf_counter()
    exact_runs(candidates, seed, side, nstep, batch=batch, on_result=keep)
    secs = time.perf_counter() - t0
    return int(np.argmin(mis)), mis, secs


def proposal_search(prop, target_t, seed, candidates, side, nstep, top_k=32):
    """Rank candidates by predicted misfit, run the top_k exactly; returns
    (best index, exact misfits of the top_k, top_k indices, predicted
    misfits, seconds)."""
    t0 = time.perf_counter()
    terrain1 = generate_terrain(side, side, int(seed)).reshape(side, side).astype(np.float32)
    pn = normalise(candidates[:, 0], candidates[:, 1], candidates[:, 2])
    pred = prop.rank(terrain1, pn, target_t)
    top = np.argsort(pred, kind="stable")[:top_k]
    mis = np.empty(len(top))

    def keep(i, h):
        mis[i] = misfit(transform(h), target_t)

    exact_runs(candidates[top], seed, side, nstep, batch=len(top), on_result=keep)
    secs = time.perf_counter() - t0
    j = int(np.argmin(mis))
    return int(top[j]), mis, top, pred, secs


def relative_param_error(a, b):
    """Per-parameter relative difference |a - b| / |b| of (K, U, m) rows."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b) / np.abs(b)
