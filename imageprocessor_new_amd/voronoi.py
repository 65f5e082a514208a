"""Voronoi-cell refinements: ``refine="radial_symmetry"`` and ``refine="isophote_curvature"``.

Keeps the reference's functions and behaviour (subpx/_gpu/centers.py:1545-1630, 978-1469;
subpx/_gpu/voronoi.py:116-145; subpx/_cpu/cells.py:34-84): components whose bounding box
touches the image border are left out; every remaining component's centroid seeds a Voronoi
cell of the image; the cell's ROI (bbox +- pad, fill = the 10th percentile of the cell's
border outside the cell) is upsampled 4x and refined by Parthasarathy's radial symmetry (the
point nearest to every gradient line) or by isophote-curvature votes; non-finite centres are
dropped; the residual / spread and the cell areas go to ``meta``.

Every step runs on the GPU (libsubpx_hip.so): spx_voronoi_labels, spx_voronoi_rois,
spx_radial_symmetry / spx_isophote_curvature; one workgroup per ROI.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, stream_handle

VORONOI_METHODS = ("radial_symmetry", "isophote_curvature")
QUALITY_KEYS = {"radial_symmetry": "radial_symmetry_residual", "isophote_curvature": "isophote_curvature_spread"}


def _ptr(t):
    return None if t is None else int(t.data_ptr())


def voronoi_labels_device(seeds: torch.Tensor, shape, *, labels: bool = True, areas: bool = True):
    """(labels (H, W) int32 or None, areas (n,) int32 or None) of the nearest seeds (device (n, 2)
    float64 [x, y]), exact ties to the lowest index."""
    H, W = int(shape[0]), int(shape[1])
    dev = seeds.device
    n = int(seeds.shape[0])
    lab = torch.empty((H, W), dtype=torch.int32, device=dev) if labels else None
    ar = torch.empty((n,), dtype=torch.int32, device=dev) if areas else None
    _lib.call("spx_voronoi_labels", _ptr(seeds.contiguous()), n, H, W, _ptr(lab), _ptr(ar), None, stream_handle())
    return lab, ar


def compute_voronoi_labels_gpu(seeds_xy, image_shape, *, device: int = 0) -> np.ndarray:
    """(H, W) int32 NumPy label map of the nearest seed (reference subpx/_gpu/voronoi.py:116-145)."""
    seeds = np.asarray(seeds_xy, dtype=np.float64)
    if seeds.ndim != 2 or seeds.shape[1] != 2:
        raise ValueError(f"seeds_xy must be (N, 2), got {seeds.shape}")
    with gpu_device(device) as dev:
        lab, _ = voronoi_labels_device(torch.from_numpy(seeds).to(dev), image_shape, areas=False)
        return lab.cpu().numpy()


def _stack_shape(rows: np.ndarray, H: int, W: int, pad: int):
    x0 = np.maximum(0, rows[:, 0].astype(np.int32) - pad)
    y0 = np.maximum(0, rows[:, 1].astype(np.int32) - pad)
    x1 = np.minimum(W, (rows[:, 0] + rows[:, 2]).astype(np.int32) + pad)
    y1 = np.minimum(H, (rows[:, 1] + rows[:, 3]).astype(np.int32) + pad)
    return int((y1 - y0).max()), int((x1 - x0).max())


def _estimate(rois_t, masks_t, refine, upsample_factor, boundary_margin, pre_smooth_sigma):
    N, Hm, Wm = rois_t.shape
    dev = rois_t.device
    centers = torch.empty((N, 2), dtype=torch.float64, device=dev)
    quality = torch.empty((N,), dtype=torch.float64, device=dev)
    f = int(upsample_factor)
    if refine == "radial_symmetry":
        m = max(1, f // 2) if boundary_margin is None else int(boundary_margin)
        _lib.call("spx_radial_symmetry", _ptr(rois_t), _ptr(masks_t), N, Hm, Wm, f, m, _ptr(centers),
                  _ptr(quality), stream_handle())
    else:
        m = f if boundary_margin is None else int(boundary_margin)
        _lib.call("spx_isophote_curvature", _ptr(rois_t), _ptr(masks_t), N, Hm, Wm, f, m, float(pre_smooth_sigma),
                  _ptr(centers), _ptr(quality), stream_handle())
    return centers, quality


def refine_rows_device(img_t: torch.Tensor, rows: np.ndarray, refine: str, *, pad: int = 3,
                       upsample_factor: int = 4):
    """Voronoi refinement of component rows [x, y, w, h, cx, cy] (host float64) on a device
    image: (centres (K, 2) image px for every row, NaN where refused, quality (K,), cell areas
    (K,)) as host arrays."""
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 6)
    K = rows.shape[0]
    if K == 0:
        return np.zeros((0, 2)), np.zeros(0), np.zeros(0, np.int64)
    H, W = img_t.shape
    dev = img_t.device
    rows_t = torch.from_numpy(np.ascontiguousarray(rows)).to(dev)
    labels, areas = voronoi_labels_device(rows_t[:, 4:6].contiguous(), (H, W))
    Hm, Wm = _stack_shape(rows, H, W, int(pad))
    rois = torch.empty((K, Hm, Wm), dtype=torch.float32, device=dev)
    masks = torch.empty((K, Hm, Wm), dtype=torch.uint8, device=dev)
    origins = torch.empty((K, 2), dtype=torch.int32, device=dev)
    _lib.call("spx_voronoi_rois", _ptr(img_t), H, W, img_t.stride(0), _ptr(labels), _ptr(rows_t), K, int(pad), Hm,
              Wm, _ptr(rois), _ptr(masks), _ptr(origins), stream_handle())
    local, quality = _estimate(rois, masks, refine, upsample_factor, None, 0.5)
    centers = local + origins.to(torch.float64)
    return centers.cpu().numpy(), quality.cpu().numpy(), areas.cpu().numpy().astype(np.int64)


def _batch(rois, masks, refine, upsample_factor, boundary_margin, pre_smooth_sigma, device):
    rois = np.asarray(rois, dtype=np.float64)
    masks = np.asarray(masks, dtype=bool)
    if rois.ndim != 3 or masks.shape != rois.shape:
        raise ValueError("rois and masks must be (N, H, W) arrays of one shape")
    N = rois.shape[0]
    if N == 0:
        return np.empty((0, 2)), np.empty((0,))
    with gpu_device(device) as dev:
        r = torch.from_numpy(np.ascontiguousarray(rois.astype(np.float32))).to(dev)
        m = torch.from_numpy(np.ascontiguousarray(masks.astype(np.uint8))).to(dev)
        c, q = _estimate(r, m, refine, upsample_factor, boundary_margin, pre_smooth_sigma)
        return c.cpu().numpy(), q.cpu().numpy()


def _radial_symmetry_batch_gpu(rois, masks, *, upsample_factor: int = 4, boundary_margin: int | None = None,
                               residual_max: float | None = None, gpu_batch: int = 2048, device: int = 0):
    """(centres (N, 2) ROI px [x, y], residuals (N,)) of a ROI stack (reference
    subpx/_gpu/centers.py:1170-1230); ``residual_max`` refuses centres above it (NaN, inf).
    ``gpu_batch`` is accepted for the signature: one launch takes any batch."""
    c, r = _batch(rois, masks, "radial_symmetry", upsample_factor, boundary_margin, 0.0, device)
    if residual_max is not None:
        bad = r > residual_max
        c[bad] = np.nan
        r[bad] = np.inf
    return c, r


def _isophote_curvature_batch_gpu(rois, masks, *, upsample_factor: int = 4, boundary_margin: int | None = None,
                                  pre_smooth_sigma: float = 0.5, gpu_batch: int = 2048, device: int = 0):
    """(centres (N, 2) ROI px [x, y], spreads (N,)) of a ROI stack (reference
    subpx/_gpu/centers.py:1416-1469)."""
    return _batch(rois, masks, "isophote_curvature", upsample_factor, boundary_margin, pre_smooth_sigma, device)
