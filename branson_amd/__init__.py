"""branson_amd -- MI355X-native replicated-mode IMC transport engine.

The engine is the HIP/C++ library ``libbranson_hip.so`` behind the C-ABI in
``include/branson_hip.h``; this package holds its sources (``csrc/``), the
in-tree build (``build.py``) and a ctypes binding (``capi.py``).  Nothing here
computes transport in Python: when the library is missing, ``capi.lib()``
raises instead of falling back.
"""
from .capi import (ALGO_HISTORY, ALGO_PERSISTENT, PHOTON_DTYPE, PRECISION_FP32_TRACK,
                   PRECISION_FP64, BhError, Engine, lib)

__all__ = ["Engine", "BhError", "lib", "PHOTON_DTYPE", "PRECISION_FP64",
           "PRECISION_FP32_TRACK", "ALGO_HISTORY", "ALGO_PERSISTENT"]
