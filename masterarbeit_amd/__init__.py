"""MI355X-native SU(2) pure-gauge link-field engine.

The compute path is hand-written HIP for gfx950 behind a C ABI
(include/lgt_su2.h, built into masterarbeit_amd/liblgt_su2.so); the reference's
own C entry points are exported on top of it by liblgt_compat.so
(include/lgt_compat.h).  This package only binds the library for Python.
"""
from .engine import (APE_ALL, APE_SPATIAL, APE_TEMPORAL, FP32, FP64, Q_CLOVER, Q_IMPROVED, Engine, LgtError,
                     heatbath_harness, load_library, metropolis_harness, overrelax_harness)

__all__ = [
    "Engine", "LgtError", "load_library", "heatbath_harness", "overrelax_harness", "metropolis_harness",
    "FP64", "FP32", "APE_SPATIAL", "APE_ALL", "APE_TEMPORAL", "Q_IMPROVED", "Q_CLOVER",
]
