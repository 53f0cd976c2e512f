"""The learned proposal against the exact-only search, end to end on the GPU.

Scaled-down inversion: 64 x 64 grids, 300 time-steps, 512 exact training
runs, 8^3 candidate (K, U, m) sets on a grid, hidden target parameters off
the grid with a terrain seed outside the training set.  The exact solver
judges every reported result.  What the trained network gains:
  * it finds a candidate whose exact misfit is within 5% of the exact-only
    search's best while running 16 of the 512 candidates exactly;
  * a proposal without the network (16 candidates drawn at random) does
    markedly worse, so the ranking carries the information.
"""
import sys

import numpy as np
import pytest

import fastscape_amd as fa
from fastscape_amd import inversion as inv

pytestmark = pytest.mark.gpu

SIDE, NSTEP, NTRAIN, TOPK = 64, 300, 512, 16


def note(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def test_proposal_recovers_exact_only_result():
    terrain, pn, t, params, secs = inv.make_training_set(NTRAIN, SIDE, NSTEP, seed0=1000,
                                                         master_seed=5)
    note("training set: %d exact runs in %.2f s" % (NTRAIN, secs))
    prop = inv.Proposal(SIDE, max_batch=32, seed=2)
    try:
        hist = prop.train(terrain, pn, t, epochs=40, batch=32, lr=2e-3, seed=4)
        note("training loss %s" % np.array2string(hist[::8], precision=4))
        assert hist[-1] < 0.2 * hist[0]
        cands = inv.candidate_grid(8)
        hidden = np.array([[3.1e-5, 1.3e-3, 0.47]])
        seed = 77
        target = {}
        inv.exact_runs(hidden, seed, SIDE, NSTEP, on_result=lambda i, h: target.update(h=h))
        tt = inv.transform(target["h"])
        best_e, mis_e, t_exact = inv.exact_only_search(tt, seed, cands, SIDE, NSTEP)
        best_p, mis_p, top, pred, t_prop = inv.proposal_search(prop, tt, seed, cands, SIDE,
                                                               NSTEP, top_k=TOPK)
        note("exact-only best %d misfit %.5f (%.2f s); proposal best %d misfit %.5f (%.2f s)"
             % (best_e, mis_e[best_e], t_exact, best_p, mis_p.min(), t_prop))
        assert mis_p.min() <= 1.05 * mis_e[best_e]
        # the predicted ranking correlates with the exact one
        rank_corr = np.corrcoef(np.argsort(np.argsort(pred)), np.argsort(np.argsort(mis_e)))[0, 1]
        note("rank correlation predicted vs exact misfit %.3f" % rank_corr)
        assert rank_corr > 0.5
        # without the network: best of TOPK random candidates, averaged over draws
        rng = np.random.default_rng(0)
        rand_best = np.array([mis_e[rng.choice(len(cands), TOPK, replace=False)].min()
                              for _ in range(200)])
        note("random-%d proposal: median best misfit %.5f" % (TOPK, np.median(rand_best)))
        assert np.median(rand_best) > 1.5 * mis_p.min()
        assert t_prop < t_exact
    finally:
        prop.close()
