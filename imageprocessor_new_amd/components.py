"""Threshold + connected components with stats on the GPU.

Replaces ``connected_components_stats`` (reference subpx/components.py:15,
subpx/_cpu/components.py:18-34, subpx/_gpu/components.py:64-134) and the Otsu threshold
and triangle thresholds and the morphology of ``_segment_binary`` (reference
subpx/_cpu/centers.py:247-266) with spx_otsu_u16 / spx_triangle_u8 / spx_threshold_mask /
spx_morph_cross / spx_threshold_cc (include/subpx_hip.h).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, stream_handle, to_device_u16
from .types import ComponentStatsResult

DEFAULT_CAPACITY = 1 << 16


def otsu_threshold_device(img: torch.Tensor, step: int = 1) -> int:
    """Otsu threshold of a (H, W) uint16 device image (view img[::step, ::step])."""
    if img.dim() != 2 or img.dtype != torch.uint16:
        raise ValueError("img must be a 2-D uint16 tensor")
    out = ctypes.c_int32(0)
    H, W = img.shape
    _lib.call("spx_otsu_u16", int(img.data_ptr()), H, W, img.stride(0), int(step),
              ctypes.byref(out), stream_handle())
    return int(out.value)


def triangle_threshold_device(img: torch.Tensor, step: int = 1) -> int:
    """Triangle threshold of a (H, W) uint16 device image holding 8-bit grey levels (OpenCV's
    THRESH_TRIANGLE is 8-bit only; grey levels above 255 raise ValueError)."""
    if img.dim() != 2 or img.dtype != torch.uint16:
        raise ValueError("img must be a 2-D uint16 tensor")
    out = ctypes.c_int32(0)
    H, W = img.shape
    _lib.call("spx_triangle_u8", int(img.data_ptr()), H, W, img.stride(0), int(step), ctypes.byref(out),
              stream_handle())
    return int(out.value)


def binary_mask_device(img: torch.Tensor, thresh: int, *, invert: bool = False, morph_open: int = 0,
                       morph_close: int = 0) -> torch.Tensor:
    """Foreground mask (uint16 0 / 1) of ``_segment_binary`` (reference subpx/_cpu/centers.py:257-265):
    binarise, then ``morph_open`` erosions followed by as many dilations, then ``morph_close``
    dilations followed by as many erosions, all with the 3x3 cross."""
    if img.dim() != 2 or img.dtype != torch.uint16 or img.stride(1) != 1:
        raise ValueError("img must be a 2-D uint16 tensor with unit column stride")
    H, W = img.shape
    a = torch.empty((H, W), dtype=torch.uint16, device=img.device)
    b = torch.empty_like(a)
    _lib.call("spx_threshold_mask", int(img.data_ptr()), H, W, img.stride(0), int(thresh), int(bool(invert)),
              int(a.data_ptr()), stream_handle())
    ops = [0] * int(morph_open) + [1] * int(morph_open) + [1] * int(morph_close) + [0] * int(morph_close)
    for op in ops:
        _lib.call("spx_morph_cross", int(a.data_ptr()), H, W, W, op, int(b.data_ptr()), stream_handle())
        a, b = b, a
    return a


def threshold_components_device(img: torch.Tensor, thresh: int, *, invert: bool = False,
                                connectivity: int = 8, capacity: int | None = None, labels_out: bool = True):
    """Foreground (g > thresh, or g <= thresh with invert) + components.

    Returns (num_labels, labels (H, W) int32 device tensor, stats (num, 5) int32 device,
    centroids (num, 2) float64 device). Rows follow OpenCV's layout, row 0 = background.
    ``labels_out=False`` skips the label image (labels is then None; the stats are the same).
    """
    if connectivity not in (4, 8):
        raise ValueError("connectivity must be 4 or 8")
    if img.dim() != 2 or img.dtype != torch.uint16 or img.stride(1) != 1:
        raise ValueError("img must be a 2-D uint16 tensor with unit column stride")
    H, W = img.shape
    dev = img.device
    labels = torch.empty((H, W), dtype=torch.int32, device=dev) if labels_out else None
    cap = int(capacity or DEFAULT_CAPACITY)
    for _ in range(2):
        stats = torch.empty((cap, 5), dtype=torch.int32, device=dev)
        cents = torch.empty((cap, 2), dtype=torch.float64, device=dev)
        n = ctypes.c_int32(0)
        try:
            _lib.call("spx_threshold_cc", int(img.data_ptr()), H, W, img.stride(0), int(thresh),
                      int(bool(invert)), int(connectivity), None if labels is None else int(labels.data_ptr()),
                      int(stats.data_ptr()), int(cents.data_ptr()), cap, ctypes.byref(n),
                      stream_handle())
        except _lib.SpxError as e:
            if e.code != _lib.SPX_E_CAPACITY:
                raise
            cap = int(n.value)
            continue
        num = int(n.value)
        return num, labels, stats[:num], cents[:num]
    raise RuntimeError("spx_threshold_cc: capacity retry failed")


def connected_components_stats(mask, *, connectivity: int = 8, backend: str = "gpu",
                               device: int = 0) -> ComponentStatsResult:
    """Connected components of ``mask > 0`` (reference subpx/components.py:15)."""
    m = np.asarray(mask)
    if m.ndim != 2:
        raise ValueError("connected_components_stats expects a 2D mask.")
    with gpu_device(device) as dev:
        img = to_device_u16((m > 0).astype(np.uint16), dev)
        num, labels, stats, cents = threshold_components_device(img, 0, connectivity=connectivity)
        torch.cuda.current_stream().synchronize()
        return ComponentStatsResult(
            labels=labels.cpu().numpy(), stats=stats.cpu().numpy(), centroids=cents.cpu().numpy(),
            num_labels=num, backend="gpu",
            meta={"connectivity": int(connectivity), "implementation": "spx_threshold_cc"},
        )
