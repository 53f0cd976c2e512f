"""ctypes binding of the CPU oracle (liboracle.so, fastscape_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product path.
Stage functions follow fastscape_RB.cpp (RB:111-328); see the C source for
the line-by-line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CLI_PATH = os.path.join(HERE, "oracle_cli")
REF_DIR = os.path.join(HERE, "_ref")

POW_LIBM = 0
POW_EXACT = 1
NO_FLOW = -1


class Params(ctypes.Structure):
    _fields_ = [("keq", ctypes.c_double), ("neq", ctypes.c_double), ("meq", ctypes.c_double),
                ("ueq", ctypes.c_double), ("dt", ctypes.c_double), ("dx", ctypes.c_double),
                ("cell_area", ctypes.c_double), ("tol", ctypes.c_double)]


_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile liboracle.so and oracle_cli (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.fso_preset.argtypes = [ctypes.c_int, ctypes.POINTER(Params)]
        L.fso_mt_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.fso_mt_next.argtypes = [ctypes.c_void_p]
        L.fso_mt_next.restype = ctypes.c_uint32
        L.fso_canonical.argtypes = [ctypes.c_void_p]
        L.fso_canonical.restype = ctypes.c_double
        L.fso_init_terrain.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
        L.fso_receivers.argtypes = [_dp, _ip, ctypes.c_int, ctypes.c_int, _dp]
        L.fso_donors.argtypes = [_ip, _ip, _ip, ctypes.c_int, ctypes.c_int]
        L.fso_order.argtypes = [_ip, _ip, _ip, _ip, _ip, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _ip]
        L.fso_flowacc.argtypes = [_ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.fso_uplift.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double]
        L.fso_erode.argtypes = [_ip, _ip, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(Params), ctypes.c_int]
        L.fso_erode_levels_omp.argtypes = [_ip, _ip, ctypes.c_int, _ip, _dp, _dp, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(Params), ctypes.c_int]
        L.fso_write_dem.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int]
        L.fso_fill_depressions.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        _LIB = L
    return _LIB


def preset(which, **overrides):
    p = Params()
    if lib().fso_preset(which, ctypes.byref(p)) != 0:
        raise ValueError("bad preset %r" % which)
    for k, v in overrides.items():
        setattr(p, k, float(v))
    return p


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class MT19937:
    """The oracle's mt19937 (random.hpp:22)."""

    def __init__(self, seed):
        self._buf = ctypes.create_string_buffer(624 * 4 + 8)
        lib().fso_mt_seed(self._buf, seed)

    def next(self):
        return lib().fso_mt_next(self._buf)

    def canonical(self):
        return lib().fso_canonical(self._buf)


def dr(p):
    s2 = 1.414213562373095048801688724209698078569671875376948
    return np.array([p.dx if i % 2 == 0 else p.dx * s2 for i in range(8)], dtype=np.float64)


def init_terrain(W, H, engine_seed=0):
    h = np.empty(W * H, dtype=np.float64)
    lib().fso_init_terrain(_d(h), W, H, engine_seed)
    return h


def receivers(h, W, H, p, rec=None):
    """ComputeReceivers (RB:160-182); rings 0/1 keep NO_FLOW."""
    rec = np.full(W * H, NO_FLOW, dtype=np.int32) if rec is None else rec.copy()
    d = dr(p)
    lib().fso_receivers(_d(np.ascontiguousarray(h, dtype=np.float64)), _i(rec), W, H, _d(d))
    return rec


def donors(rec, W, H):
    """ComputeDonors (RB:188-203); unused donor slots are 0."""
    ndon = np.zeros(W * H, dtype=np.int32)
    donor = np.zeros((W * H, 8), dtype=np.int32)
    lib().fso_donors(_i(np.ascontiguousarray(rec, dtype=np.int32)), _i(ndon), _i(donor), W, H)
    return ndon, donor


def order(rec, ndon, donor, W, H):
    """GenerateOrder (RB:213-254) -> (stack, levels)."""
    stack = np.zeros(W * H, dtype=np.int32)
    levels = np.zeros(W * H + 3, dtype=np.int32)
    nlevel = ctypes.c_int(0)
    rc = lib().fso_order(_i(np.ascontiguousarray(rec, dtype=np.int32)),
                         _i(np.ascontiguousarray(ndon, dtype=np.int32)),
                         _i(np.ascontiguousarray(donor, dtype=np.int32)), _i(stack), _i(levels),
                         levels.size, W, H, ctypes.byref(nlevel))
    if rc != 0:
        raise RuntimeError("fso_order failed: %d" % rc)
    return stack, levels[:nlevel.value].copy()


def flowacc(stack, rec, W, H, cell_area):
    """ComputeFlowAcc (RB:262-279)."""
    accum = np.zeros(W * H, dtype=np.float64)
    lib().fso_flowacc(_i(np.ascontiguousarray(stack, dtype=np.int32)),
                      _i(np.ascontiguousarray(rec, dtype=np.int32)), _d(accum), W, H, cell_area)
    return accum


def uplift(h, W, H, p):
    """AddUplift (RB:285-296)."""
    h = np.array(h, dtype=np.float64, copy=True)
    lib().fso_uplift(_d(h), W, H, p.ueq, p.dt)
    return h


def erode(stack, rec, accum, h, W, H, p, powmode=POW_LIBM):
    """Erode (RB:306-328), on a copy of h."""
    h = np.array(h, dtype=np.float64, copy=True)
    rc = lib().fso_erode(_i(np.ascontiguousarray(stack, dtype=np.int32)),
                         _i(np.ascontiguousarray(rec, dtype=np.int32)),
                         _d(np.ascontiguousarray(accum, dtype=np.float64)), _d(h), W, H,
                         ctypes.byref(p), powmode)
    if rc != 0:
        raise RuntimeError("Newton iteration cap reached")
    return h


def erode_levels_omp(stack, levels, rec, accum, h, W, H, p, powmode=POW_LIBM):
    """Erode of fastscape_RB+P.cpp (P:313-351), on a copy of h."""
    h = np.array(h, dtype=np.float64, copy=True)
    lv = np.ascontiguousarray(levels, dtype=np.int32)
    rc = lib().fso_erode_levels_omp(_i(np.ascontiguousarray(stack, dtype=np.int32)), _i(lv),
                                    lv.size, _i(np.ascontiguousarray(rec, dtype=np.int32)),
                                    _d(np.ascontiguousarray(accum, dtype=np.float64)), _d(h),
                                    W, H, ctypes.byref(p), powmode)
    if rc != 0:
        raise RuntimeError("Newton iteration cap reached")
    return h


def step(h, W, H, p, powmode=POW_LIBM):
    """One time-step from elevations h (RB:367-372); returns every stage."""
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    rec = receivers(h, W, H, p)
    ndon, donor = donors(rec, W, H)
    stack, levels = order(rec, ndon, donor, W, H)
    accum = flowacc(stack, rec, W, H, p.cell_area)
    hup = uplift(h, W, H, p)
    hnew = erode(stack, rec, accum, hup, W, H, p, powmode)
    return dict(rec=rec, ndon=ndon, donor=donor, stack=stack, levels=levels, accum=accum,
                hup=hup, h=hnew)


def run(W, H, nstep, p, engine_seed=0, powmode=POW_LIBM, h=None):
    """Construct + run(nstep): nstep+1 time-steps (RB:366)."""
    h = init_terrain(W, H, engine_seed) if h is None else np.array(h, dtype=np.float64).ravel()
    for _ in range(nstep + 1):
        h = step(h, W, H, p, powmode)["h"]
    return h


def fill_depressions(h, W, H):
    """Priority-Flood+epsilon (Barnes, Lehman & Mulla 2014) on a copy of h:
    the depression-filled surface, rings 0 and 1 unchanged.  No reference
    counterpart (the reference has no depression handling)."""
    h = np.array(h, dtype=np.float64, copy=True).ravel()
    lib().fso_fill_depressions(_d(h), W, H)
    return h


def write_dem(filename, h, W, H):
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    if lib().fso_write_dem(os.fsencode(filename), _d(h), W, H) != 0:
        raise OSError("cannot write " + filename)
