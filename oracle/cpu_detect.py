"""ctypes binding of oracle/cpu_detect.cpp: the CPU detect path (Otsu, 8-connected components,
per-label log-quadratic) in C++ on every host core.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the CPU baseline that bench.py times beside
the GPU; it restates oracle/segment.detect_centers_logquad.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "libcpudetect.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run `make -C oracle`")
        lib = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        lib.cpu_otsu_u16.restype = ctypes.c_int32
        lib.cpu_otsu_u16.argtypes = [P, I64, I64, I64, ctypes.c_int32]
        lib.cpu_detect_logquad.restype = I64
        lib.cpu_detect_logquad.argtypes = [P, I64, I64, I64, ctypes.c_int32, I64, I64, ctypes.c_int32,
                                           ctypes.c_int32, P, I64, P]
        _lib = lib
    return _lib


def otsu_u16(img, threads: int = 0) -> int:
    g = np.ascontiguousarray(img, dtype=np.uint16)
    return int(_load().cpu_otsu_u16(g.ctypes.data_as(ctypes.c_void_p), g.shape[0], g.shape[1], g.shape[1],
                                    int(threads)))


def detect_centers_logquad(img, *, threshold_value=None, area_min=1, area_max=50, pad=3, threads: int = 0):
    """(centres (N, 2) ordered by each component's first pixel, threshold) of the CPU path."""
    lib = _load()
    g = np.ascontiguousarray(img, dtype=np.uint16)
    H, W = g.shape
    cap = max(16, H * W // 4)
    out = np.zeros((cap, 2))
    t = ctypes.c_int32(0)
    n = lib.cpu_detect_logquad(g.ctypes.data_as(ctypes.c_void_p), H, W, W,
                               -1 if threshold_value is None else int(threshold_value), int(area_min), int(area_max),
                               int(pad), int(threads), out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(t))
    if n < 0:
        raise RuntimeError("cpu_detect_logquad: output capacity exceeded")
    return out[:n].copy(), int(t.value)
