"""ctypes binding of libfastscape_hip.so (the C ABI in include/fastscape.h).

The Python surface mirrors the reference class (fastscape_RB.cpp:46-406):
``FastScape(width, height)``, ``run(nstep)``, ``getH()``, plus the per-stage
readouts used by the parity tests.  There is no fallback: if the HIP
library is missing or no GPU is present, construction raises.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfastscape_hip.so")

FS_OK = 0
FS_NO_FLOW = -1
FS_MAX_LEVELS = 65533
PRESET_REFERENCE = 0
PRESET_BENCHMARK = 1
TIMER_NAMES = ["Step1: Initialize", "Step2: DetermineReceivers", "Step3: DetermineDonors",
               "Step4: GenerateOrder", "Step5: FlowAcc", "Step6: Uplift", "Step7: Erosion",
               "Overall"]
KERNELS = ["k_receivers", "k_topology", "k_roots", "k_bfs", "k_flowacc", "k_uplift_roots",
           "k_erode", "k_jump", "k_tile_acc", "k_tile_erode", "k_deep_local", "k_deep_nodes",
           "k_deep_counts", "k_deep_erode"]
PATH_AUTO = 0
PATH_LEVELS = 1
PATH_TILE_GRAPH = 2
TILE_HALO = 16
TILE_HALO_DEEP = 32

# Every symbol declared in include/fastscape.h.
EXPORTED = [
    "fs_params_preset", "fs_create", "fs_create_from_h", "fs_destroy", "fs_run", "fs_step",
    "fs_step_noalloc",
    "fs_get_h", "fs_set_h", "fs_get_timers", "fs_reset_timers", "fs_get_receivers",
    "fs_get_donors", "fs_get_stack", "fs_get_levels", "fs_get_accum", "fs_device_h",
    "fs_stream", "fs_synchronize", "fs_fill_depressions", "fs_set_depression_fill",
    "fs_set_step_path", "fs_path_stats", "fs_tile_stats", "fs_graph_stats", "fs_set_kernel_timing", "fs_kernel_stats",
    "fs_kernel_name", "fs_calibrate_copy", "fs_ensemble_create", "fs_ensemble_destroy",
    "fs_ensemble_size", "fs_ensemble_run", "fs_ensemble_step", "fs_ensemble_get_h",
    "fs_ensemble_set_h", "fs_ensemble_get_receivers", "fs_ensemble_get_accum",
    "fs_ensemble_get_levels", "fs_ensemble_path_stats", "fs_ensemble_set_schedule",
    "fs_ensemble_get_schedule", "fs_ensemble_autotune", "fs_ensemble_tune_log",
    "fs_ensemble_stats", "fs_ensemble_synchronize",
    "fs_unet_create", "fs_unet_destroy", "fs_unet_param_count", "fs_unet_layer_shapes",
    "fs_unet_get_params", "fs_unet_set_params", "fs_unet_forward", "fs_unet_loss_grad",
    "fs_unet_set_data", "fs_unet_train", "fs_unet_rank", "fs_unet_stats",
    "fs_write_dem", "fs_generate_terrain", "fs_strerror", "fs_last_error",
]


class FsParams(ctypes.Structure):
    """fs_params: keq, neq, meq, ueq, dt, dx, cell_area, tol (RB:58-65)."""
    _fields_ = [("keq", ctypes.c_double), ("neq", ctypes.c_double), ("meq", ctypes.c_double),
                ("ueq", ctypes.c_double), ("dt", ctypes.c_double), ("dx", ctypes.c_double),
                ("cell_area", ctypes.c_double), ("tol", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FsRealization(ctypes.Structure):
    """fs_realization: grid size, parameter block, terrain seed."""
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("params", FsParams),
                ("engine_seed", ctypes.c_uint32)]


class FsSchedule(ctypes.Structure):
    """fs_schedule: receiver rows per item, tiles per claim, cells per level-path workgroup."""
    _fields_ = [("recv_rows", ctypes.c_int), ("tiles_per_claim", ctypes.c_int),
                ("group_cells", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FastScapeError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        detail = _lib().fs_last_error().decode()
        msg = "%s: %s" % (where, _lib().fs_strerror(status).decode())
        super().__init__(msg + (" (%s)" % detail if detail else ""))


_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            "HIP library not built: %s (run barnes2019-landscape_amd/build.py)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "fs_params_preset": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(FsParams)]),
        "fs_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(FsParams),
                                     ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(_vp)]),
        "fs_create_from_h": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(FsParams),
                                            _dp, ctypes.c_int, ctypes.POINTER(_vp)]),
        "fs_destroy": (ctypes.c_int, [_vp]),
        "fs_run": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fs_step": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fs_step_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(FsParams), ctypes.c_int, _vp]),
        "fs_get_h": (ctypes.c_int, [_vp, _dp, ctypes.c_size_t]),
        "fs_set_h": (ctypes.c_int, [_vp, _dp, ctypes.c_size_t]),
        "fs_get_timers": (ctypes.c_int, [_vp, _u64p]),
        "fs_reset_timers": (ctypes.c_int, [_vp]),
        "fs_get_receivers": (ctypes.c_int, [_vp, _ip, ctypes.c_size_t]),
        "fs_get_donors": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_size_t]),
        "fs_get_stack": (ctypes.c_int, [_vp, _ip, ctypes.c_size_t]),
        "fs_get_levels": (ctypes.c_int, [_vp, _ip, ctypes.c_size_t, _ip]),
        "fs_get_accum": (ctypes.c_int, [_vp, _dp, ctypes.c_size_t]),
        "fs_device_h": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
        "fs_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
        "fs_synchronize": (ctypes.c_int, [_vp]),
        "fs_fill_depressions": (ctypes.c_int, [_vp]),
        "fs_set_depression_fill": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fs_set_step_path": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fs_path_stats": (ctypes.c_int, [_vp, _u64p, _u64p]),
        "fs_tile_stats": (ctypes.c_int, [_vp, _u64p, _u64p]),
        "fs_graph_stats": (ctypes.c_int, [_vp, _u64p, ctypes.POINTER(ctypes.c_int32)]),
        "fs_set_kernel_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fs_kernel_stats": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _u64p]),
        "fs_kernel_name": (ctypes.c_char_p, [ctypes.c_int]),
        "fs_calibrate_copy": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, _dp]),
        "fs_ensemble_create": (ctypes.c_int, [ctypes.POINTER(FsRealization), ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(_vp)]),
        "fs_ensemble_destroy": (ctypes.c_int, [_vp]),
        "fs_ensemble_size": (ctypes.c_int, [_vp, _ip, ctypes.POINTER(ctypes.c_size_t)]),
        "fs_ensemble_run": (ctypes.c_int, [_vp, ctypes.c_int]),
        "fs_ensemble_step": (ctypes.c_int, [_vp]),
        "fs_ensemble_get_h": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "fs_ensemble_set_h": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "fs_ensemble_get_receivers": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_size_t]),
        "fs_ensemble_get_accum": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_size_t]),
        "fs_ensemble_get_levels": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_size_t, _ip]),
        "fs_ensemble_path_stats": (ctypes.c_int, [_vp, ctypes.c_int, _u64p, _u64p]),
        "fs_ensemble_set_schedule": (ctypes.c_int, [_vp, ctypes.POINTER(FsSchedule)]),
        "fs_ensemble_get_schedule": (ctypes.c_int, [_vp, ctypes.POINTER(FsSchedule)]),
        "fs_ensemble_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "fs_ensemble_tune_log": (ctypes.c_int, [_vp, ctypes.POINTER(FsSchedule), _dp,
                                                ctypes.c_int, _ip]),
        "fs_ensemble_stats": (ctypes.c_int, [_vp, _u64p, _dp]),
        "fs_ensemble_synchronize": (ctypes.c_int, [_vp]),
        "fs_unet_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_vp)]),
        "fs_unet_destroy": (ctypes.c_int, [_vp]),
        "fs_unet_param_count": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
        "fs_unet_layer_shapes": (ctypes.c_int, [_vp, _ip, _ip, _ip, ctypes.c_int, _ip]),
        "fs_unet_get_params": (ctypes.c_int, [_vp, _fp, ctypes.c_size_t]),
        "fs_unet_set_params": (ctypes.c_int, [_vp, _fp, ctypes.c_size_t]),
        "fs_unet_forward": (ctypes.c_int, [_vp, _fp, _fp, ctypes.c_int, _fp]),
        "fs_unet_loss_grad": (ctypes.c_int, [_vp, _fp, _fp, _fp, ctypes.c_int, _fp, _fp,
                                             ctypes.c_size_t]),
        "fs_unet_set_data": (ctypes.c_int, [_vp, _fp, _fp, _fp, ctypes.c_int]),
        "fs_unet_train": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_uint64, _fp]),
        "fs_unet_rank": (ctypes.c_int, [_vp, _fp, _fp, ctypes.c_int, _fp, _fp]),
        "fs_unet_stats": (ctypes.c_int, [_vp, _dp, _dp, _u64p]),
        "fs_write_dem": (ctypes.c_int, [ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int]),
        "fs_generate_terrain": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]),
        "fs_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "fs_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def library():
    """The loaded ctypes library (raises if it is not built)."""
    return _lib()


def _check(status, where):
    if status != FS_OK:
        raise FastScapeError(status, where)


def preset(which=PRESET_REFERENCE, **overrides):
    p = FsParams()
    _check(_lib().fs_params_preset(which, ctypes.byref(p)), "fs_params_preset")
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, float(v))
    return p


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def generate_terrain(width, height, engine_seed=0):
    """GenerateRandomTerrain's recipe (RB:111-130) on the host."""
    h = np.empty(width * height, dtype=np.float64)
    _check(_lib().fs_generate_terrain(_ptr(h, ctypes.c_double), width, height, engine_seed),
           "fs_generate_terrain")
    return h


def calibrate_copy(nbytes, device=0):
    """Time a plain device copy of nbytes (calibrates traffic counters)."""
    ms = ctypes.c_double(0)
    _check(_lib().fs_calibrate_copy(int(nbytes), int(device), ctypes.byref(ms)),
           "fs_calibrate_copy")
    return ms.value


def write_dem(filename, h, width, height):
    """PrintDEM (RB:20-40)."""
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    if h.size != width * height:
        raise ValueError("h has %d cells, expected %d" % (h.size, width * height))
    _check(_lib().fs_write_dem(os.fsencode(filename), _ptr(h, ctypes.c_double), width, height),
           "fs_write_dem")


class FastScape:
    """The reference model class (fastscape_RB.cpp:46-406) on one GPU.

    ``engine_seed`` seeds the mt19937 terrain generator directly; the
    reference's CLI seed maps to engine seed 0 for every nonzero value
    (random.cpp:31).
    """

    def __init__(self, width, height, params=None, engine_seed=0, device=0, h=None):
        self.width, self.height = int(width), int(height)
        self.size = self.width * self.height
        self.params = params if params is not None else preset(PRESET_REFERENCE)
        self._m = _vp()
        lib = _lib()
        if h is None:
            st = lib.fs_create(self.width, self.height, ctypes.byref(self.params),
                               int(engine_seed), int(device), ctypes.byref(self._m))
        else:
            h = np.ascontiguousarray(h, dtype=np.float64).ravel()
            if h.size != self.size:
                raise ValueError("h has %d cells, expected %d" % (h.size, self.size))
            st = lib.fs_create_from_h(self.width, self.height, ctypes.byref(self.params),
                                      _ptr(h, ctypes.c_double), int(device),
                                      ctypes.byref(self._m))
        _check(st, "fs_create")

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if self._m:
            _lib().fs_destroy(self._m)
            self._m = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- the reference surface ------------------------------------------------
    def run(self, nstep):
        """run(nstep): nstep+1 time-steps (RB:341-398)."""
        _check(_lib().fs_run(self._m, int(nstep)), "fs_run")

    def step(self, iterations=1):
        """Exactly ``iterations`` time-steps."""
        _check(_lib().fs_step(self._m, int(iterations)), "fs_step")

    def getH(self):
        """Elevations as a (height, width) array (RB:403-405)."""
        h = np.empty(self.size, dtype=np.float64)
        _check(_lib().fs_get_h(self._m, _ptr(h, ctypes.c_double), self.size), "fs_get_h")
        return h.reshape(self.height, self.width)

    def setH(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        if h.size != self.size:
            raise ValueError("h has %d cells, expected %d" % (h.size, self.size))
        _check(_lib().fs_set_h(self._m, _ptr(h, ctypes.c_double), self.size), "fs_set_h")

    def timers(self):
        """Cumulative stage timers in microseconds (RB:380-387)."""
        t = (ctypes.c_uint64 * 8)()
        _check(_lib().fs_get_timers(self._m, t), "fs_get_timers")
        return dict(zip(TIMER_NAMES, [int(v) for v in t]))

    def reset_timers(self):
        _check(_lib().fs_reset_timers(self._m), "fs_reset_timers")

    # -- state of the last time-step --------------------------------------------
    def receivers(self):
        r = np.empty(self.size, dtype=np.int32)
        _check(_lib().fs_get_receivers(self._m, _ptr(r, ctypes.c_int32), self.size),
               "fs_get_receivers")
        return r

    def donors(self):
        ndon = np.empty(self.size, dtype=np.int32)
        donor = np.empty((self.size, 8), dtype=np.int32)
        _check(_lib().fs_get_donors(self._m, _ptr(ndon, ctypes.c_int32),
                                    _ptr(donor, ctypes.c_int32), self.size), "fs_get_donors")
        return ndon, donor

    def stack(self):
        s = np.empty(self.size, dtype=np.int32)
        _check(_lib().fs_get_stack(self._m, _ptr(s, ctypes.c_int32), self.size), "fs_get_stack")
        return s

    def levels(self):
        n = ctypes.c_int32(0)
        _check(_lib().fs_get_levels(self._m, None, 0, ctypes.byref(n)), "fs_get_levels")
        lv = np.empty(max(n.value, 1), dtype=np.int32)
        _check(_lib().fs_get_levels(self._m, _ptr(lv, ctypes.c_int32), lv.size,
                                    ctypes.byref(n)), "fs_get_levels")
        return lv[:n.value]

    def accum(self):
        a = np.empty(self.size, dtype=np.float64)
        _check(_lib().fs_get_accum(self._m, _ptr(a, ctypes.c_double), self.size), "fs_get_accum")
        return a

    # -- device access ---------------------------------------------------------------
    def synchronize(self):
        _check(_lib().fs_synchronize(self._m), "fs_synchronize")

    def stream(self):
        s = _vp()
        _check(_lib().fs_stream(self._m, ctypes.byref(s)), "fs_stream")
        return s.value

    def device_h(self):
        p = _vp()
        _check(_lib().fs_device_h(self._m, ctypes.byref(p)), "fs_device_h")
        return p.value

    def fill_depressions(self):
        """Fill the current field's depressions once (Priority-Flood+epsilon)."""
        _check(_lib().fs_fill_depressions(self._m), "fs_fill_depressions")

    def set_depression_fill(self, enable):
        """Fill depressions before the receivers of every following step."""
        _check(_lib().fs_set_depression_fill(self._m, 1 if enable else 0),
               "fs_set_depression_fill")

    def set_path(self, path):
        """PATH_AUTO (halo tiles when levels allow, else the tile graph),
        PATH_LEVELS or PATH_TILE_GRAPH."""
        _check(_lib().fs_set_step_path(self._m, int(path)), "fs_set_step_path")

    def path_stats(self):
        """Steps taken by each path: {"tile": n, "levels": n, "graph": n}."""
        t = ctypes.c_uint64(0)
        lv = ctypes.c_uint64(0)
        _check(_lib().fs_path_stats(self._m, ctypes.byref(t), ctypes.byref(lv)), "fs_path_stats")
        return {"tile": t.value, "levels": lv.value, "graph": self.graph_stats()["steps"]}

    def graph_stats(self):
        """Tile-graph steps and the most tile borders one drainage path
        crossed in the last of them: {"steps": n, "tiers_last_step": k}."""
        n = ctypes.c_uint64(0)
        k = ctypes.c_int32(0)
        _check(_lib().fs_graph_stats(self._m, ctypes.byref(n), ctypes.byref(k)), "fs_graph_stats")
        return {"steps": n.value, "tiers_last_step": k.value}

    def tile_stats(self):
        """Tile-path steps split by halo: {"halo16": n, "halo32": n}."""
        t16 = ctypes.c_uint64(0)
        t32 = ctypes.c_uint64(0)
        _check(_lib().fs_tile_stats(self._m, ctypes.byref(t16), ctypes.byref(t32)), "fs_tile_stats")
        return {"halo16": t16.value, "halo32": t32.value}

    def set_kernel_timing(self, enable):
        _check(_lib().fs_set_kernel_timing(self._m, 1 if enable else 0), "fs_set_kernel_timing")

    def kernel_stats(self):
        out = {}
        for k, name in enumerate(KERNELS):
            ms = ctypes.c_double(0)
            n = ctypes.c_uint64(0)
            _check(_lib().fs_kernel_stats(self._m, k, ctypes.byref(ms), ctypes.byref(n)),
                   "fs_kernel_stats")
            out[name] = (ms.value, n.value)
        return out
