"""Edge-profile centre refinements on the GPU (reference subpx/_gpu/centers.py:518-929).

``refine_centers_edge_erf_gpu`` and ``refine_centers_edge_moment_gpu`` keep the reference's
signatures, defaults and return value: ``(centers (N, 2) float64, ok (N,) bool)`` for rows
``[x, y, w, h, cx, cy]``. Each runs in one HIP kernel, one thread per feature
(``spx_edge_erf``, ``spx_edge_moment``); the reference's per-profile fit
``_edge_from_profile_erf_gpu`` (:665-779) is exposed as ``edge_from_profile_erf`` through
``spx_edge_profile_erf``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, stream_handle

DEFAULTS = dict(band_rad=4, edge_rad=10, smooth_passes=2, loc_rad=3, grad_power=8.0, iters=2,
                erf_fit_rad=5, erf_gn_iters=5, erf_init_sigma=1.25, erf_sigma_min=0.35, erf_sigma_max=6.0,
                erf_max_shift=2.5, erf_damp=1e-4, erf_cond_max=1e8)


def edge_options(**kw) -> _lib.SpxEdgeErfOptions:
    o = dict(DEFAULTS)
    unknown = set(kw) - set(o)
    if unknown:
        raise TypeError(f"unknown edge-erf options: {sorted(unknown)}")
    o.update(kw)
    return _lib.SpxEdgeErfOptions(
        int(o["band_rad"]), int(o["edge_rad"]), int(o["smooth_passes"]), int(o["loc_rad"]), float(o["grad_power"]),
        int(o["iters"]), int(o["erf_fit_rad"]), int(o["erf_gn_iters"]), float(o["erf_init_sigma"]),
        float(o["erf_sigma_min"]), float(o["erf_sigma_max"]), float(o["erf_max_shift"]), float(o["erf_damp"]),
        float(o["erf_cond_max"]))


def refine_edge_erf_device(img_t: torch.Tensor, stats6_t: torch.Tensor, **kw):
    """Refine on a device image (uint16 or float32, (H, W), unit column stride) and device rows
    (K, 6) float32. Returns NumPy (centers (K, 2) float64, ok (K,) bool)."""
    return _edge_call("spx_edge_erf", img_t, stats6_t, **kw)


def refine_edge_moment_device(img_t: torch.Tensor, stats6_t: torch.Tensor, **kw):
    """Gradient-moment edges only (no erf fits); arguments and result as refine_edge_erf_device."""
    return _edge_call("spx_edge_moment", img_t, stats6_t, **kw)


def _edge_call(name, img_t, stats6_t, **kw):
    if img_t.dim() != 2 or img_t.stride(1) != 1:
        raise ValueError("image must be a 2-D device tensor with unit column stride")
    if img_t.dtype == torch.uint16:
        dtype = _lib.DTYPE_U16
    elif img_t.dtype == torch.float32:
        dtype = _lib.DTYPE_F32
    else:
        raise TypeError(f"image dtype must be uint16 or float32, got {img_t.dtype}")
    dev = img_t.device
    rows = stats6_t.to(device=dev, dtype=torch.float32)[:, :6].contiguous()
    K = int(rows.shape[0])
    centers = torch.empty((K, 2), dtype=torch.float64, device=dev)
    ok = torch.empty((K,), dtype=torch.uint8, device=dev)
    if K:
        opt = edge_options(**kw)
        H, W = img_t.shape
        _lib.call(name, img_t.data_ptr(), dtype, int(H), int(W), int(img_t.stride(0)), rows.data_ptr(), K,
                  ctypes.byref(opt), centers.data_ptr(), ok.data_ptr(), stream_handle())
    torch.cuda.current_stream().synchronize()
    return centers.cpu().numpy(), ok.cpu().numpy().astype(bool)


def refine_centers_edge_erf_gpu(gray, stats_xywh_cc, *, band_rad: int = 4, edge_rad: int = 10,
                                smooth_passes: int = 2, loc_rad: int = 3, grad_power: float = 8.0, iters: int = 2,
                                erf_fit_rad: int = 5, erf_gn_iters: int = 5, erf_init_sigma: float = 1.25,
                                erf_sigma_min: float = 0.35, erf_sigma_max: float = 6.0, erf_max_shift: float = 2.5,
                                erf_damp: float = 1e-4, erf_cond_max: float = 1e8, device: int = 0):
    """Reference refine_centers_edge_erf_gpu (subpx/_gpu/centers.py:882-929): gradient-moment
    seeds then 1-D erf-edge fits of the four edge profiles of each feature."""
    resolve_backend("gpu")
    g_np = np.asarray(gray, dtype=np.float32)
    stats_np = np.asarray(stats_xywh_cc, dtype=np.float32)
    if g_np.ndim != 2:
        raise ValueError("gray must be a 2D grayscale image")
    if stats_np.size == 0:
        return np.zeros((0, 2), dtype=np.float64), np.zeros((0,), dtype=bool)
    if stats_np.ndim != 2 or stats_np.shape[1] < 6:
        raise ValueError("stats_xywh_cc must have shape (N, >=6) with columns [x, y, w, h, cx, cy]")
    with gpu_device(device) as dev:
        img_t = torch.from_numpy(np.ascontiguousarray(g_np)).to(dev)
        rows_t = torch.from_numpy(np.ascontiguousarray(stats_np[:, :6])).to(dev)
        return refine_edge_erf_device(
            img_t, rows_t, band_rad=band_rad, edge_rad=edge_rad, smooth_passes=smooth_passes, loc_rad=loc_rad,
            grad_power=grad_power, iters=iters, erf_fit_rad=erf_fit_rad, erf_gn_iters=erf_gn_iters,
            erf_init_sigma=erf_init_sigma, erf_sigma_min=erf_sigma_min, erf_sigma_max=erf_sigma_max,
            erf_max_shift=erf_max_shift, erf_damp=erf_damp, erf_cond_max=erf_cond_max)


def refine_centers_edge_moment_gpu(gray, stats_xywh_cc, *, band_rad: int = 4, edge_rad: int = 10,
                                   smooth_passes: int = 2, loc_rad: int = 3, grad_power: float = 8.0, iters: int = 2,
                                   device: int = 0):
    """Reference refine_centers_edge_moment_gpu (subpx/_gpu/centers.py:607-636): centre =
    midpoints of the gradient-moment edges of the four edge profiles, two passes. The
    reference computes the profiles and moments in float32; this kernel uses float64."""
    resolve_backend("gpu")
    g_np = np.asarray(gray, dtype=np.float32)
    stats_np = np.asarray(stats_xywh_cc, dtype=np.float32)
    if g_np.ndim != 2:
        raise ValueError("gray must be a 2D grayscale image")
    if stats_np.size == 0:
        return np.zeros((0, 2), dtype=np.float64), np.zeros((0,), dtype=bool)
    if stats_np.ndim != 2 or stats_np.shape[1] < 6:
        raise ValueError("stats_xywh_cc must have shape (N, >=6) with columns [x, y, w, h, cx, cy]")
    with gpu_device(device) as dev:
        img_t = torch.from_numpy(np.ascontiguousarray(g_np)).to(dev)
        rows_t = torch.from_numpy(np.ascontiguousarray(stats_np[:, :6])).to(dev)
        return refine_edge_moment_device(img_t, rows_t, band_rad=band_rad, edge_rad=edge_rad,
                                         smooth_passes=smooth_passes, loc_rad=loc_rad, grad_power=grad_power,
                                         iters=iters)


def edge_from_profile_erf(prof, base_x, x_seed, *, fit_rad: int = 5, gn_iters: int = 5, init_sigma: float = 1.25,
                          sigma_min: float = 0.35, sigma_max: float = 6.0, max_shift: float = 2.5,
                          damp: float = 1e-4, cond_max: float = 1e8, loc_rad: int = 2, device: int = 0):
    """Reference _edge_from_profile_erf_gpu (subpx/_gpu/centers.py:665-779): one erf-edge
    Gauss-Newton fit per profile row. Returns NumPy (x_edge (B,) float64, NaN where refused;
    ok (B,) bool)."""
    resolve_backend("gpu")
    prof = np.ascontiguousarray(np.asarray(prof, dtype=np.float64))
    if prof.ndim != 2:
        raise ValueError("prof must have shape (B, L)")
    B, L = prof.shape
    base = np.array(np.broadcast_to(np.asarray(base_x, dtype=np.float64), (B,)))
    seed = np.array(np.broadcast_to(np.asarray(x_seed, dtype=np.float64), (B,)))
    if L < 5:
        return seed.copy(), np.zeros(B, dtype=bool)
    if L > 31:
        raise ValueError("profiles longer than 31 samples are not supported")
    opt = edge_options(loc_rad=loc_rad, erf_fit_rad=fit_rad, erf_gn_iters=gn_iters, erf_init_sigma=init_sigma,
                       erf_sigma_min=sigma_min, erf_sigma_max=sigma_max, erf_max_shift=max_shift, erf_damp=damp,
                       erf_cond_max=cond_max)
    with gpu_device(device) as dev:
        p_t = torch.from_numpy(prof).to(dev)
        b_t = torch.from_numpy(base).to(dev)
        s_t = torch.from_numpy(seed).to(dev)
        x_t = torch.empty((B,), dtype=torch.float64, device=dev)
        ok_t = torch.empty((B,), dtype=torch.uint8, device=dev)
        if B:
            _lib.call("spx_edge_profile_erf", p_t.data_ptr(), B, L, b_t.data_ptr(), s_t.data_ptr(), ctypes.byref(opt),
                      x_t.data_ptr(), ok_t.data_ptr(), stream_handle())
        torch.cuda.current_stream().synchronize()
        return x_t.cpu().numpy(), ok_t.cpu().numpy().astype(bool)
