"""Ensembles: many FastScape grids stepped by one persistent launch per step.

Binding of the fs_ensemble_* entry points of include/fastscape.h.  Each
realization has its own size, parameter block (fs_params, RB:58-65) and
terrain seed, and its time-step is the reference's (fastscape_RB.cpp:366-376).
"""
import ctypes

import numpy as np

from ._lib import (FS_OK, FastScapeError, FsParams, FsRealization, FsSchedule, _check, _lib,
                   _ptr)

_u64 = ctypes.c_uint64


def realization(width, height, params, engine_seed=0):
    """An fs_realization from a size, an FsParams block and a terrain seed."""
    if not isinstance(params, FsParams):
        raise TypeError("params must be an FsParams")
    return FsRealization(int(width), int(height), params, int(engine_seed))


def schedule(recv_rows=64, tiles_per_claim=1, group_cells=0):
    return FsSchedule(int(recv_rows), int(tiles_per_claim), int(group_cells))


class Ensemble:
    """Realizations advanced together on one GPU.

    ``reals`` is a sequence of FsRealization (see ``realization``).  Every
    accessor takes the realization's index in that sequence.
    """

    def __init__(self, reals, device=0):
        self._e = None
        reals = list(reals)
        if not reals:
            raise ValueError("an ensemble needs at least one realization")
        arr = (FsRealization * len(reals))(*reals)
        self.shapes = [(r.height, r.width) for r in reals]
        e = ctypes.c_void_p()
        _check(_lib().fs_ensemble_create(arr, len(reals), int(device), ctypes.byref(e)),
               "fs_ensemble_create")
        self._e = e

    def close(self):
        if self._e is not None:
            _lib().fs_ensemble_destroy(self._e)
            self._e = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.shapes)

    @property
    def cells(self):
        n = ctypes.c_int(0)
        c = ctypes.c_size_t(0)
        _check(_lib().fs_ensemble_size(self._e, ctypes.byref(n), ctypes.byref(c)),
               "fs_ensemble_size")
        return c.value

    def run(self, nstep):
        """nstep+1 time-steps of every realization (RB:366)."""
        _check(_lib().fs_ensemble_run(self._e, int(nstep)), "fs_ensemble_run")

    def step(self):
        _check(_lib().fs_ensemble_step(self._e), "fs_ensemble_step")

    def _n(self, r):
        # an index out of range reaches the library, which rejects it
        if not 0 <= int(r) < len(self.shapes):
            return 0
        H, W = self.shapes[r]
        return W * H

    def getH(self, r):
        """Elevations of realization r as a (height, width) array."""
        h = np.empty(self._n(r), dtype=np.float64)
        _check(_lib().fs_ensemble_get_h(self._e, int(r), _ptr(h, ctypes.c_double), h.size),
               "fs_ensemble_get_h")
        return h.reshape(self.shapes[r])

    def setH(self, r, h):
        h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        _check(_lib().fs_ensemble_set_h(self._e, int(r), _ptr(h, ctypes.c_double), h.size),
               "fs_ensemble_set_h")

    def receivers(self, r):
        a = np.empty(self._n(r), dtype=np.int32)
        _check(_lib().fs_ensemble_get_receivers(self._e, int(r), _ptr(a, ctypes.c_int32), a.size),
               "fs_ensemble_get_receivers")
        return a

    def accum(self, r):
        a = np.empty(self._n(r), dtype=np.float64)
        _check(_lib().fs_ensemble_get_accum(self._e, int(r), _ptr(a, ctypes.c_double), a.size),
               "fs_ensemble_get_accum")
        return a

    def levels(self, r):
        n = ctypes.c_int32(0)
        _check(_lib().fs_ensemble_get_levels(self._e, int(r), None, 0, ctypes.byref(n)),
               "fs_ensemble_get_levels")
        a = np.empty(max(n.value, 1), dtype=np.int32)
        _check(_lib().fs_ensemble_get_levels(self._e, int(r), _ptr(a, ctypes.c_int32), a.size,
                                             ctypes.byref(n)), "fs_ensemble_get_levels")
        return a[:n.value].copy()

    def path_stats(self, r):
        t, l = _u64(0), _u64(0)
        _check(_lib().fs_ensemble_path_stats(self._e, int(r), ctypes.byref(t), ctypes.byref(l)),
               "fs_ensemble_path_stats")
        return t.value, l.value

    def set_schedule(self, s):
        _check(_lib().fs_ensemble_set_schedule(self._e, ctypes.byref(s)),
               "fs_ensemble_set_schedule")

    def get_schedule(self):
        s = FsSchedule()
        _check(_lib().fs_ensemble_get_schedule(self._e, ctypes.byref(s)),
               "fs_ensemble_get_schedule")
        return s

    def autotune(self, trial_steps, interval=0):
        _check(_lib().fs_ensemble_autotune(self._e, int(trial_steps), int(interval)),
               "fs_ensemble_autotune")

    def tune_log(self):
        n = ctypes.c_int(0)
        _check(_lib().fs_ensemble_tune_log(self._e, None, None, 0, ctypes.byref(n)),
               "fs_ensemble_tune_log")
        cand = (FsSchedule * max(n.value, 1))()
        ms = np.zeros(max(n.value, 1), dtype=np.float64)
        _check(_lib().fs_ensemble_tune_log(self._e, cand, _ptr(ms, ctypes.c_double), n.value,
                                           ctypes.byref(n)), "fs_ensemble_tune_log")
        return [(cand[i].as_dict(), float(ms[i])) for i in range(n.value)]

    def stats(self):
        s = _u64(0)
        ms = ctypes.c_double(0)
        _check(_lib().fs_ensemble_stats(self._e, ctypes.byref(s), ctypes.byref(ms)),
               "fs_ensemble_stats")
        return s.value, ms.value

    def synchronize(self):
        _check(_lib().fs_ensemble_synchronize(self._e), "fs_ensemble_synchronize")


def config4_realizations(count=256, master_seed=0, sides=(512, 768, 1024, 1536, 2048)):
    """The ensemble workload: square sides drawn uniformly from `sides`,
    K log-uniform [1e-6, 1e-4], U log-uniform [5e-4, 5e-3] m/yr, m uniform
    [0.3, 0.6], n = 1, dx = 100 m (cell area 1e4 m^2), dt = 1000 yr; realization r
    has terrain seed r.  Parameters come from numpy's PCG64 seeded with
    master_seed, drawn in realization order (side, K, U, m)."""
    from ._lib import PRESET_BENCHMARK, preset
    rng = np.random.Generator(np.random.PCG64(master_seed))
    out, meta = [], []
    for r in range(count):
        side = int(rng.choice(sides))
        K = float(10.0 ** rng.uniform(-6.0, -4.0))
        U = float(10.0 ** rng.uniform(np.log10(5e-4), np.log10(5e-3)))
        m = float(rng.uniform(0.3, 0.6))
        p = preset(PRESET_BENCHMARK, keq=K, ueq=U, meq=m, neq=1.0)
        out.append(realization(side, side, p, engine_seed=r))
        meta.append(dict(side=side, K=K, U=U, m=m, seed=r))
    return out, meta
