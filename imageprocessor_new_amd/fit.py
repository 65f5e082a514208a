"""Batched erf-disk / erf-ellipse LM fits of one patch per feature on the GPU.

No reference function fits this model; the batched per-feature LM replaces the role of
the reference's `lm_fit` (subpx/_gpu/deblend.py:55-152) for contact-hole centres, radii
(CD = 2 r) and edge widths. The kernel is spx_fit_batch (include/subpx_hip.h).

Neighbour policy. Where a patch is larger than the hole pitch (the 15-63 px patches of a
staggered array at ~39 px pitch), neighbouring holes fall inside it. With ``cells=True`` a
patch pixel takes part only when it lies in the feature's Voronoi cell against its 8
nearest seeds (spx_seed_neighbours + spx_fit_batch_cells), the cell rule of the reference's
Voronoi refinements (subpx/_cpu/cells.py:8-84). A patch that lies inside its cell is fitted
exactly as without the mask.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, stream_handle, to_device_u16
from .types import ErfFitResult

MODELS = {"disk": _lib.MODEL_ERF_DISK6, "erf_disk": _lib.MODEL_ERF_DISK6,
          "ellipse": _lib.MODEL_ERF_ELLIPSE7, "erf_ellipse": _lib.MODEL_ERF_ELLIPSE7}
MAX_HALF = 31


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else int(t.data_ptr())


NEIGHBOURS = 8


def seed_neighbours_device(seeds: torch.Tensor, shape, radius: float, nb: int = NEIGHBOURS,
                           stream: torch.cuda.Stream | None = None) -> torch.Tensor:
    """(K, nb) int32 device tensor of the nb nearest other seeds within ``radius`` px
    (-1 pads), ordered by distance then index (spx_seed_neighbours)."""
    K = int(seeds.shape[0])
    if seeds.shape != (K, 2) or seeds.dtype != torch.float32 or not seeds.is_contiguous():
        raise ValueError("seeds must be a contiguous (K, 2) float32 tensor")
    out = torch.empty((K, nb), dtype=torch.int32, device=seeds.device)
    if K:
        H, W = int(shape[0]), int(shape[1])
        _lib.call("spx_seed_neighbours", _ptr(seeds), K, H, W, float(radius), int(nb), _ptr(out),
                  stream_handle(stream))
    return out


def cell_radius(half_max: int) -> float:
    """Seeds farther than this cannot move the bisector into a patch of half side half_max, nor
    into the 4-neighbours the moment start tests around it. A patch pixel q lies within
    sqrt(2) (h + 0.5) of the seed s (the patch is centred on rint(s)) and its 4-neighbours within
    sqrt(2) (h + 1.5); a seed t can take q only when |t - s| <= |t - q| + |q - s| <= 2 |q - s|.
    The 1e-3 px covers the float32 rounding of the radius."""
    return 2.0 * float(np.sqrt(2.0)) * (float(half_max) + 1.5) + 1e-3


# Launch configurations chosen by autotune_launch, per (device, K class).
_TUNED: dict = {}
LAUNCH_CANDIDATES = ((4, 1), (2, 1), (1, 1), (4, 0))  # (waves per workgroup, persistent)


def _tune_key(dev: torch.device, K: int):
    return (dev.index if dev.index is not None else torch.cuda.current_device(), max(int(K), 1).bit_length())


def _with_launch(o: _lib.SpxFitOptions, tile_waves: int, persistent: int,
                 bucket_waves: int | None = None) -> _lib.SpxFitOptions:
    return _lib.fit_options(max_iter=o.max_iter, step_tol=o.step_tol, lambda0=o.lambda0, lambda_up=o.lambda_up,
                            lambda_down=o.lambda_down, init_sigma=o.init_sigma, tile_waves=int(tile_waves),
                            persistent=int(persistent), bucket_mask=o.bucket_mask, learned_iters=o.learned_iters,
                            bucket_waves=o.bucket_waves if bucket_waves is None else int(bucket_waves))


WAVES_PER_SIMD = (2, 3, 4)  # register-allocation variants of every bucket kernel


def _median_ms(fn, repeats: int) -> float:
    times = []
    for _ in range(max(1, int(repeats))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times))


def autotune_launch(img: torch.Tensor, seeds: torch.Tensor, half: torch.Tensor, model: torch.Tensor, *,
                    options: _lib.SpxFitOptions | None = None, neighbours: torch.Tensor | None = None,
                    candidates=LAUNCH_CANDIDATES, repeats: int = 3, tune_buckets: bool = True) -> dict:
    """Time the fit's launch configurations on these features with HIP events and remember the
    fastest for this device and batch size class; fit_erf_disks_device (and FitGraph) use it
    when options.tile_waves is 0. Results do not depend on the configuration.

    First the whole launch over ``candidates`` (waves per workgroup, persistent grid or one
    wave per feature); then, with ``tune_buckets``, every bucket the features fill on its own
    (bucket_mask = that bucket) at 2, 3 and 4 waves per SIMD, keeping each bucket's fastest
    register allocation (SpxFitOptions.bucket_waves). Returns {"best": (waves, persistent,
    bucket_waves), "ms": {...}, "bucket_ms": {bucket: {waves_per_simd: ms}}}."""
    base = options if options is not None else _lib.fit_options()
    K = int(seeds.shape[0])
    out = None
    ms = {}
    for cand in candidates:
        o = _with_launch(base, *cand)
        out = fit_erf_disks_device(img, seeds, half, model, options=o, out=out, neighbours=neighbours)
        ms[f"{cand[0]}w_{'persistent' if cand[1] else 'per_feature'}"] = _median_ms(
            lambda: fit_erf_disks_device(img, seeds, half, model, options=o, out=out, neighbours=neighbours),
            repeats)
    best = min(zip(candidates, ms.values()), key=lambda t: t[1])[0]
    bucket_waves = int(base.bucket_waves)
    bucket_ms = {}
    if tune_buckets:
        mask = int(base.bucket_mask) or _lib.bucket_mask_of(half.cpu().numpy(), model.cpu().numpy())
        for b in range(10):
            if not (mask >> b) & 1:
                continue
            per = {}
            for v, wps in enumerate(WAVES_PER_SIMD):
                o = _with_launch(base, *best, bucket_waves=(bucket_waves & ~(3 << (2 * b))) | ((v + 1) << (2 * b)))
                o.bucket_mask = 1 << b
                fit_erf_disks_device(img, seeds, half, model, options=o, out=out, neighbours=neighbours)
                per[wps] = _median_ms(
                    lambda: fit_erf_disks_device(img, seeds, half, model, options=o, out=out, neighbours=neighbours),
                    repeats)
            pick = min(per, key=per.get)
            bucket_waves = (bucket_waves & ~(3 << (2 * b))) | ((WAVES_PER_SIMD.index(pick) + 1) << (2 * b))
            bucket_ms[b] = per
    chosen = (int(best[0]), int(best[1]), int(bucket_waves))
    _TUNED[_tune_key(img.device, K)] = chosen
    return {"best": chosen, "ms": ms, "bucket_ms": bucket_ms}


def _check_fit_inputs(img, seeds, half, model, init, neighbours) -> None:
    if img.dim() != 2 or img.dtype != torch.uint16:
        raise ValueError("img must be a 2-D uint16 tensor")
    K = int(seeds.shape[0])
    if seeds.shape != (K, 2) or seeds.dtype != torch.float32:
        raise ValueError("seeds must be (K, 2) float32")
    if half.shape != (K,) or half.dtype != torch.int16 or model.shape != (K,) or model.dtype != torch.uint8:
        raise ValueError("half must be (K,) int16 and model (K,) uint8")
    if init is not None and (init.shape != (K, 7) or init.dtype != torch.float32):
        raise ValueError("init must be (K, 7) float32")
    dev = img.device
    for t in (seeds, half, model) + ((init,) if init is not None else ()):
        if t.device != dev or not t.is_contiguous():
            raise ValueError("all inputs must be contiguous tensors on the image's device")
    if neighbours is not None and (neighbours.dim() != 2 or neighbours.shape[0] != K
                                   or neighbours.dtype != torch.int32 or neighbours.device != dev
                                   or not neighbours.is_contiguous()):
        raise ValueError("neighbours must be a contiguous (K, nb) int32 tensor on the image's device")


def _new_fit_outputs(K: int, dev) -> dict:
    return {
        "params": torch.empty((K, 7), dtype=torch.float32, device=dev),
        "centers": torch.empty((K, 2), dtype=torch.float64, device=dev),
        "rms": torch.empty((K,), dtype=torch.float32, device=dev),
        "iters": torch.empty((K,), dtype=torch.int16, device=dev),
        "status": torch.empty((K,), dtype=torch.uint8, device=dev),
    }


def fit_erf_disks_device(img: torch.Tensor, seeds: torch.Tensor, half: torch.Tensor,
                         model: torch.Tensor, init: torch.Tensor | None = None, *,
                         options: _lib.SpxFitOptions | None = None, out: dict | None = None,
                         stream: torch.cuda.Stream | None = None,
                         neighbours: torch.Tensor | None = None) -> dict:
    """Enqueue spx_fit_batch (spx_fit_batch_cells with ``neighbours``) on device tensors; no
    host synchronisation.

    img (H, W) uint16, seeds (K, 2) float32, half (K,) int16, model (K,) uint8,
    init (K, 7) float32 or None, neighbours (K, nb) int32 or None. Returns dict of device
    tensors params/centers/rms/iters/status (``centers`` (K, 2) float64: the centres unrounded;
    an ``out`` dict without it leaves them unwritten).
    """
    _check_fit_inputs(img, seeds, half, model, init, None)
    K = int(seeds.shape[0])
    dev = img.device
    if out is None:
        out = _new_fit_outputs(K, dev)
    opts = options if options is not None else _lib.fit_options()
    if opts.tile_waves == 0:
        tuned = _TUNED.get(_tune_key(dev, K))
        if tuned is not None:
            opts = _with_launch(opts, *tuned[:2], bucket_waves=tuned[2] if opts.bucket_waves == 0 else None)
    H, W = img.shape
    pitch = img.stride(0)
    centers = out.get("centers")
    if centers is not None and (centers.shape != (K, 2) or centers.dtype != torch.float64
                                or centers.device != dev or not centers.is_contiguous()):
        raise ValueError("out['centers'] must be a contiguous (K, 2) float64 tensor on the image's device")
    if neighbours is None:
        _lib.call("spx_fit_batch", _ptr(img), H, W, pitch, _ptr(seeds), _ptr(half), _ptr(model),
                  _ptr(init), ctypes.byref(opts), K, _ptr(out["params"]), _ptr(centers), _ptr(out["rms"]),
                  _ptr(out["iters"]), _ptr(out["status"]), stream_handle(stream))
        return out
    if (neighbours.dim() != 2 or neighbours.shape[0] != K or neighbours.dtype != torch.int32
            or neighbours.device != dev or not neighbours.is_contiguous()):
        raise ValueError("neighbours must be a contiguous (K, nb) int32 tensor on the image's device")
    _lib.call("spx_fit_batch_cells", _ptr(img), H, W, pitch, _ptr(seeds), _ptr(half), _ptr(model),
              _ptr(init), ctypes.byref(opts), K, _ptr(neighbours), int(neighbours.shape[1]),
              _ptr(out["params"]), _ptr(centers), _ptr(out["rms"]), _ptr(out["iters"]), _ptr(out["status"]),
              stream_handle(stream))
    return out


class FitGraph:
    """One batched fit recorded as a HIP graph (spx_fit_graph_create) and replayed by
    ``launch``: bucketing, every bucket's persistent kernel on the four forked streams and the
    joins in one hipGraphLaunch. The graph is bound to these tensors (it keeps references to
    them): each launch fits what they hold when it runs, so a loop over batches of the same
    size writes its next seeds, patch sizes or image into them (``copy_``) and launches again.
    Results equal ``fit_erf_disks_device``'s on the same inputs. The graph owns its scratch and
    streams; ``close`` (or garbage collection) waits for its launches and frees it.
    """

    def __init__(self, img: torch.Tensor, seeds: torch.Tensor, half: torch.Tensor, model: torch.Tensor,
                 init: torch.Tensor | None = None, *, options: _lib.SpxFitOptions | None = None,
                 out: dict | None = None, neighbours: torch.Tensor | None = None):
        K = int(seeds.shape[0])
        if K < 1:
            raise ValueError("a fit graph needs at least one feature")
        _check_fit_inputs(img, seeds, half, model, init, neighbours)
        dev = img.device
        if out is None:
            out = _new_fit_outputs(K, dev)
        opts = options if options is not None else _lib.fit_options()
        if opts.tile_waves == 0:
            tuned = _TUNED.get(_tune_key(dev, K))
            if tuned is not None:
                opts = _with_launch(opts, *tuned[:2], bucket_waves=tuned[2] if opts.bucket_waves == 0 else None)
        self._keep = (img, seeds, half, model, init, neighbours)
        self.out = out
        self.options = opts
        self.device = dev
        self._handle = ctypes.c_void_p()
        H, W = img.shape
        nb = int(neighbours.shape[1]) if neighbours is not None else 0
        with torch.cuda.device(dev):
            _lib.call("spx_fit_graph_create", _ptr(img), H, W, img.stride(0), _ptr(seeds), _ptr(half),
                      _ptr(model), _ptr(init), ctypes.byref(opts), K, _ptr(neighbours), nb,
                      _ptr(out["params"]), _ptr(out.get("centers")), _ptr(out["rms"]), _ptr(out["iters"]),
                      _ptr(out["status"]), ctypes.byref(self._handle))

    def launch(self, stream: torch.cuda.Stream | None = None) -> dict:
        """Enqueue one replay on ``stream`` (default: the current stream); returns ``out``."""
        if not self._handle:
            raise RuntimeError("the fit graph is closed")
        with torch.cuda.device(self.device):
            _lib.call("spx_fit_graph_launch", self._handle, stream_handle(stream))
        return self.out

    def close(self) -> None:
        if self._handle:
            _lib.call("spx_fit_graph_destroy", self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_erf_disks(image, seeds, *, half=15, model="disk", init=None, max_iter: int = 30,
                  step_tol: float = 1e-6, lambda0: float = 1e-3, lambda_up: float = 10.0,
                  lambda_down: float = 0.1, init_sigma: float = 1.0, cells: bool = False,
                  learned_iters: int = 0, backend: str = "gpu", device: int = 0) -> ErfFitResult:
    """Fit an erf-edged disk (6 parameters) or ellipse (7) to a patch around each seed.

    ``seeds`` (K, 2) [x, y]; the patch is (2 half + 1)^2 pixels centred on rint(seed)
    (half per feature or scalar, 2..31). ``model`` is "disk" / "ellipse" or a per-feature
    array of 0/1. ``init`` (K, 7) [x0, y0, r, sigma, A, B, q] replaces the moment start.
    Fits stop after ``max_iter`` step attempts or when a step below ``step_tol`` px is
    taken. ``cells=True`` restricts every patch to the feature's Voronoi cell among the
    seeds (the neighbour policy above).
    ``learned_iters`` > 0 makes ``init`` a learned proposal (SpxFitOptions.learned_iters): at
    most that many attempts from it, the residual judge, and the plain fit from the moment
    start for the features it does not accept.
    """
    resolve_backend(backend)
    seeds_np = np.asarray(seeds, dtype=np.float32).reshape(-1, 2)
    K = seeds_np.shape[0]
    half_np = np.broadcast_to(np.asarray(half, dtype=np.int16), (K,)).copy()
    if K and (half_np.min() < 2 or half_np.max() > MAX_HALF):
        raise ValueError(f"half must lie in [2, {MAX_HALF}]")
    if isinstance(model, str):
        if model not in MODELS:
            raise ValueError(f"model must be one of {sorted(MODELS)}")
        model_np = np.full(K, MODELS[model], dtype=np.uint8)
    else:
        model_np = np.broadcast_to(np.asarray(model, dtype=np.uint8), (K,)).copy()
    opts = _lib.fit_options(max_iter=max_iter, step_tol=step_tol, lambda0=lambda0,
                            lambda_up=lambda_up, lambda_down=lambda_down, init_sigma=init_sigma,
                            bucket_mask=_lib.bucket_mask_of(half_np, model_np), learned_iters=learned_iters)
    with gpu_device(device) as dev:
        img_t = to_device_u16(image, dev)
        if K == 0:
            empty = np.zeros((0, 7), np.float32)
            return ErfFitResult(empty, np.zeros(0, np.float32), np.zeros(0, np.int16),
                                np.zeros(0, np.uint8), model_np, centers=np.zeros((0, 2)))
        seeds_t = torch.from_numpy(seeds_np).to(dev)
        half_t = torch.from_numpy(half_np).to(dev)
        model_t = torch.from_numpy(model_np).to(dev)
        init_t = None
        if init is not None:
            init_t = torch.from_numpy(np.asarray(init, dtype=np.float32).reshape(K, 7)).to(dev)
        nbr = None
        if cells:
            nbr = seed_neighbours_device(seeds_t, img_t.shape, cell_radius(int(half_np.max())))
        out = fit_erf_disks_device(img_t, seeds_t, half_t, model_t, init_t, options=opts, neighbours=nbr)
        torch.cuda.current_stream().synchronize()
        return ErfFitResult(out["params"].cpu().numpy(), out["rms"].cpu().numpy(),
                            out["iters"].cpu().numpy(), out["status"].cpu().numpy(), model_np,
                            backend="gpu",
                            meta={"max_iter": int(max_iter), "step_tol": float(step_tol), "cells": bool(cells)},
                            centers=out["centers"].cpu().numpy())
