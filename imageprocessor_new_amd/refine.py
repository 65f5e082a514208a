"""Per-feature centre refinements on labelled components (GPU kernels behind the C-ABI).

* ``refine_centers_logquad_gpu_match_cpu``: the GPU backend's "logquad" method (reference
  subpx/_gpu/centers.py:288-517), kernel ``spx_refine_logquad_window``;
* ``refine_logquad_rows_device``: the CPU backend's log-quadratic fit per label (reference
  subpx/_cpu/centers.py:67-107, 376-391), kernel ``spx_refine_logquad``;
* ``refine_weighted_rows_device``: the GPU detect path's "weighted" / "none" centroid
  (reference subpx/_gpu/centers.py:43-66, 931-955), kernel ``spx_refine_weighted``.

All return NumPy ``(centers (K, 2) float64, ok (K,) bool)``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, stream_handle

LOGQUAD_WINDOW_DEFAULTS = dict(win_rad=4, core_frac=0.20, weight_power=1.0, bg_dilate=2, bg_win_pad=2,
                               bg_tail_frac=0.05, plane_bg=True, min_bg_pts=12, min_plane_pts=20, max_cond=1e10,
                               negdef_eps=1e-12, fit_dilate=1, eps=1e-12)


def logquad_window_options(**kw) -> _lib.SpxLogquadWindowOptions:
    o = dict(LOGQUAD_WINDOW_DEFAULTS)
    unknown = set(kw) - set(o)
    if unknown:
        raise TypeError(f"unknown logquad options: {sorted(unknown)}")
    o.update(kw)
    return _lib.SpxLogquadWindowOptions(
        int(o["win_rad"]), int(o["bg_win_pad"]), int(o["bg_dilate"]), int(o["fit_dilate"]), int(o["min_bg_pts"]),
        int(o["min_plane_pts"]), 1 if o["plane_bg"] else 0, 0, float(o["core_frac"]), float(o["weight_power"]),
        float(o["bg_tail_frac"]), float(o["max_cond"]), float(o["negdef_eps"]), float(o["eps"]))


def _image_dtype(img_t: torch.Tensor) -> int:
    if img_t.dim() != 2 or img_t.stride(1) != 1:
        raise ValueError("image must be a 2-D device tensor with unit column stride")
    if img_t.dtype == torch.uint16:
        return _lib.DTYPE_U16
    if img_t.dtype == torch.float32:
        return _lib.DTYPE_F32
    raise TypeError(f"image dtype must be uint16 or float32, got {img_t.dtype}")


def _labels_checked(labels_t: torch.Tensor, img_t: torch.Tensor) -> torch.Tensor:
    if labels_t.dtype != torch.int32 or tuple(labels_t.shape) != tuple(img_t.shape):
        raise ValueError("labels must be an int32 tensor of the image's shape")
    return labels_t.contiguous()


def _outputs(K: int, dev):
    return (torch.empty((K, 2), dtype=torch.float64, device=dev), torch.empty((K,), dtype=torch.uint8, device=dev))


def _finish(centers: torch.Tensor, ok: torch.Tensor):
    torch.cuda.current_stream().synchronize()
    return centers.cpu().numpy(), ok.cpu().numpy().astype(bool)


def refine_logquad_window_device(img_t, labels_t, rows6_t, lab_ids_t, *, pad: int = 3, **kw):
    """Window log-quadratic refinement on device tensors: image (H, W) uint16 / float32,
    labels (H, W) int32, rows (K, 6) float32 [x, y, w, h, cx, cy], lab_ids (K,) int32."""
    dtype = _image_dtype(img_t)
    dev = img_t.device
    labels_t = _labels_checked(labels_t, img_t)
    rows = rows6_t.to(device=dev, dtype=torch.float32).contiguous()
    ids = lab_ids_t.to(device=dev, dtype=torch.int32).contiguous()
    K = int(rows.shape[0])
    if rows.dim() != 2 or rows.shape[1] != 6 or ids.shape != (K,):
        raise ValueError("rows must be (K, 6) and lab_ids (K,)")
    centers, ok = _outputs(K, dev)
    if K:
        opt = logquad_window_options(**kw)
        H, W = img_t.shape
        _lib.call("spx_refine_logquad_window", img_t.data_ptr(), dtype, int(H), int(W), int(img_t.stride(0)),
                  labels_t.data_ptr(), rows.data_ptr(), ids.data_ptr(), K, int(pad), ctypes.byref(opt),
                  centers.data_ptr(), ok.data_ptr(), stream_handle())
    return _finish(centers, ok)


def refine_centers_logquad_gpu_match_cpu(g_gpu, lab_gpu, stats_gpu, lab_ids_gpu, *, pad: int = 3, win_rad: int = 4,
                                         core_frac: float = 0.20, weight_power: float = 1.0, bg_dilate: int = 2,
                                         bg_win_pad: int = 2, bg_tail_frac: float = 0.05, plane_bg: bool = True,
                                         min_bg_pts: int = 12, min_plane_pts: int = 20, max_cond: float = 1e10,
                                         negdef_eps: float = 1e-12, fit_dilate: int = 1, eps: float = 1e-12,
                                         batch: int = 50_000, use_float64: bool = True, device: int = 0):
    """Reference refine_centers_logquad_gpu_match_cpu (subpx/_gpu/centers.py:288-517).

    Same arguments and validation. ``g_gpu`` is a 2-D float32 image and ``lab_gpu`` int32
    labels (torch tensors or NumPy arrays); ``stats_gpu`` has 4 columns [x, y, w, h] (centre
    x + w // 2, y + h // 2) or 6 [x, y, w, h, cx, cy]. The kernel computes in FP64 and needs no
    batching, so ``batch`` and ``use_float64`` do not change the result. Returns NumPy
    ``(refined (K, 2) float64, ok (K,) bool)``; refused rows hold the centroid.
    """
    resolve_backend("gpu")
    if len(np.shape(g_gpu)) != 2 or _dtype_name(g_gpu) != "float32":
        raise ValueError(f"g_gpu must be 2D float32, got shape={tuple(np.shape(g_gpu))} dtype={_dtype_name(g_gpu)}")
    if len(np.shape(lab_gpu)) != 2 or _dtype_name(lab_gpu) != "int32":
        raise ValueError(f"lab_gpu must be 2D int32, got shape={tuple(np.shape(lab_gpu))} dtype={_dtype_name(lab_gpu)}")
    stats = _to_numpy(stats_gpu)
    ids = _to_numpy(lab_ids_gpu).astype(np.int32).ravel()
    if stats.ndim != 2 or stats.shape[0] != ids.shape[0]:
        raise ValueError(f"stats_gpu rows ({stats.shape[0]}) must match lab_ids_gpu length ({ids.shape[0]})")
    if stats.shape[1] not in (4, 6):
        raise ValueError(f"stats_gpu must have 4 or 6 columns, got {stats.shape[1]}")
    if stats.shape[1] == 4:
        xywh = stats.astype(np.int32)
        stats = np.concatenate([stats.astype(np.float32), (xywh[:, :2] + xywh[:, 2:4] // 2).astype(np.float32)], 1)
    with gpu_device(device) as dev:
        img_t = torch.as_tensor(g_gpu).to(dev).contiguous()
        lab_t = torch.as_tensor(lab_gpu).to(dev).contiguous()
        return refine_logquad_window_device(
            img_t, lab_t, torch.from_numpy(np.ascontiguousarray(stats, np.float32)).to(dev),
            torch.from_numpy(ids).to(dev), pad=pad, win_rad=win_rad, core_frac=core_frac,
            weight_power=weight_power, bg_dilate=bg_dilate, bg_win_pad=bg_win_pad, bg_tail_frac=bg_tail_frac,
            plane_bg=plane_bg, min_bg_pts=min_bg_pts, min_plane_pts=min_plane_pts, max_cond=max_cond,
            negdef_eps=negdef_eps, fit_dilate=fit_dilate, eps=eps)


def _rows5_call(name, img_t, labels_t, rows5, pad):
    if img_t.dtype != torch.uint16:
        raise TypeError(f"{name}: image must be uint16")
    _image_dtype(img_t)
    dev = img_t.device
    labels_t = _labels_checked(labels_t, img_t)
    rows_t = torch.as_tensor(np.ascontiguousarray(np.asarray(rows5, dtype=np.int32).reshape(-1, 5))).to(dev)
    K = int(rows_t.shape[0])
    centers, ok = _outputs(K, dev)
    if K:
        H, W = img_t.shape
        _lib.call(name, img_t.data_ptr(), int(H), int(W), int(img_t.stride(0)), labels_t.data_ptr(), rows_t.data_ptr(),
                  K, int(pad), centers.data_ptr(), ok.data_ptr(), stream_handle())
    return _finish(centers, ok)


def refine_logquad_rows_device(img_t, labels_t, rows5, *, pad: int = 3):
    """CPU-backend log-quadratic fit per label; rows (K, 5) [x, y, w, h, label]."""
    return _rows5_call("spx_refine_logquad", img_t, labels_t, rows5, pad)


def refine_weighted_rows_device(img_t, labels_t, rows5, *, pad: int = 3):
    """Weighted centroid per label; rows (K, 5) [x, y, w, h, label]."""
    return _rows5_call("spx_refine_weighted", img_t, labels_t, rows5, pad)


def _dtype_name(a) -> str:
    if isinstance(a, torch.Tensor):
        return str(a.dtype).replace("torch.", "")
    return str(np.asarray(a).dtype)


def _to_numpy(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)
