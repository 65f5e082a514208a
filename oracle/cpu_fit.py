"""ctypes binding of oracle/cpu_fit.cpp: the FP64 erf-disk / erf-ellipse LM on every host core.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the CPU baseline that bench.py times beside
the GPU; it restates oracle/erfdisk.fit_batch (plain fits) in C++ with OpenMP.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "libcpufit.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run `make -C oracle`")
        lib = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        lib.cpu_fit_batch.restype = ctypes.c_int32
        lib.cpu_fit_batch.argtypes = [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, P, P, P, P, ctypes.c_int32,
                                      ctypes.c_int64, P, ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_int32, P, P, P, P]
        _lib = lib
    return _lib


def host_cores() -> int:
    """Cores this process may run on."""
    return len(os.sched_getaffinity(0))


def fit_batch(img, seeds, half, model, neighbours=None, only=None, threads: int = 0, max_iter: int = 30,
              step_tol: float = 1e-6):
    """(params (K, 7), rms, iterations, status, threads used) of the plain LM from the moment start,
    for the features `only` (all by default); the others stay NaN with status 4."""
    lib = _load()
    img = np.ascontiguousarray(img, dtype=np.uint16)
    seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 2)
    K = seeds.shape[0]
    half = np.ascontiguousarray(np.broadcast_to(np.asarray(half, np.int16), (K,)))
    model = np.ascontiguousarray(np.broadcast_to(np.asarray(model, np.uint8), (K,)))
    nbr = None if neighbours is None else np.ascontiguousarray(neighbours, dtype=np.int32)
    idx = None if only is None else np.ascontiguousarray(only, dtype=np.int64)
    params = np.full((K, 7), np.nan)
    rms = np.full(K, np.nan)
    iters = np.zeros(K, np.int32)
    status = np.full(K, 4, np.uint8)
    p = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    used = lib.cpu_fit_batch(p(img), img.shape[0], img.shape[1], img.strides[0] // 2, p(seeds), p(half), p(model),
                             p(nbr), 0 if nbr is None else nbr.shape[1], K, p(idx), 0 if idx is None else idx.size,
                             int(max_iter), float(step_tol), int(threads), p(params), p(rms), p(iters), p(status))
    return params, rms, iters, status, int(used)
