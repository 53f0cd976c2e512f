"""MI355X FastScape time-step: Python host binding of libfastscape_hip.so.

The compute path is the HIP library (csrc/); this package only loads it and
mirrors the reference class surface (fastscape_RB.cpp:46-406) for tests,
benchmarks and scripting.
"""
from ._lib import (EXPORTED, FS_MAX_LEVELS, FS_NO_FLOW, KERNELS, LIB_PATH, PATH_AUTO,
                   PATH_LEVELS, PATH_TILE_GRAPH, PRESET_BENCHMARK, PRESET_REFERENCE, TILE_HALO,
                   TILE_HALO_DEEP,
                   TIMER_NAMES, FastScape, FastScapeError, FsParams,
                   FsRealization, FsSchedule, calibrate_copy, generate_terrain, library, preset,
                   write_dem)
from .ensemble import Ensemble, config4_realizations, realization, schedule
from .unet import UNet

__all__ = ["EXPORTED", "FS_MAX_LEVELS", "FS_NO_FLOW", "KERNELS", "LIB_PATH", "PATH_AUTO",
           "PATH_LEVELS", "PATH_TILE_GRAPH", "TILE_HALO", "TILE_HALO_DEEP",
           "PRESET_BENCHMARK", "PRESET_REFERENCE", "TIMER_NAMES", "FastScape",
           "FastScapeError", "FsParams", "FsRealization", "FsSchedule", "Ensemble", "UNet",
           "config4_realizations", "realization", "schedule", "calibrate_copy", "generate_terrain", "library", "preset", "write_dem"]
