"""Public centre detection API on the GPU.

Keeps the signature and result of the reference ``detect_centers`` (subpx/centers.py:130-220)
and its GPU path (subpx/_gpu/centers.py:1472-1862): threshold -> connected components ->
area filter -> per-feature refinement, optionally over horizontal strips with a global
Otsu threshold, an ownership band and a final de-duplication. Every stage runs on the GPU
through libsubpx_hip.so; the host only filters component rows and merges strips.

Refinement methods (reference _detect_single_tile_gpu, subpx/_gpu/centers.py:1632-1710):
  * "logquad": the GPU backend's window log-quadratic fit with plane background
    (refine.py, reference :288-517); ``logquad_variant="cpu"`` selects the CPU backend's
    per-label fit instead (subpx/_cpu/centers.py:67-107);
  * "edge_gradmoment" (aliases "edge_moment", "edges_centered"): gradient-moment edges
    (reference :547-636);
  * "edge_erf" (aliases "edge_gradmoment_erf", "edge_erf_centered"): erf-edge fits
    (reference :782-929);
  * "weighted" and "none": the background-subtracted weighted centroid (reference
    :1687-1710, both through _weighted_from_batch_gpu);
  * "auto": per component, "logquad" when max(w, h) <= small_feature_max, else
    "edge_gradmoment" (reference subpx/_cpu/centers.py:22-30);
  * "erf_disk" / "erf_ellipse": batched erf-edge disk / ellipse LM fits seeded at the
    component centroid (no reference counterpart; see fit.py);
  * "radial_symmetry" / "isophote_curvature": the Voronoi-cell methods (voronoi.py, reference
    :1545-1630): components off the image border, one cell ROI each, ``meta`` carries the
    residual / spread and the cell areas.
Centres are emitted method group by method group in the reference's order (logquad,
edge_gradmoment, edge_erf, weighted, none), each group in component order, keeping only
accepted, finite centres.
"""

from __future__ import annotations

import time
from typing import Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, to_device_u16
from .components import (binary_mask_device, otsu_threshold_device, threshold_components_device,
                         triangle_threshold_device)
from .types import CenterResult

_ERF_METHODS = {"erf_disk": _lib.MODEL_ERF_DISK6, "erf_ellipse": _lib.MODEL_ERF_ELLIPSE7}
_ALIASES = {"logquadratic": "logquad", "edge_moment": "edge_gradmoment", "edges_centered": "edge_gradmoment",
            "edge_gradmoment_erf": "edge_erf", "edge_erf_centered": "edge_erf"}
_GROUP_ORDER = ("logquad", "edge_gradmoment", "edge_erf", "weighted", "none") + tuple(_ERF_METHODS)
_SUPPORTED = set(_GROUP_ORDER) | set(_ALIASES) | {"auto"}
_VORONOI_METHODS = {"radial_symmetry", "isophote_curvature"}


def choose_refine_method_for_bbox(width: float, height: float, *, small_feature_max: float = 12.0) -> str:
    """Reference subpx/_cpu/centers.py:22-30."""
    return "logquad" if float(max(width, height)) <= float(small_feature_max) else "edge_gradmoment"


def _row_methods(refine, rows_np, small_feature_max):
    """Method of every row as an index into _GROUP_ORDER (int8)."""
    K = rows_np.shape[0]
    if refine == "auto":
        size = np.maximum(rows_np[:, 2], rows_np[:, 3])
        lq, em = _GROUP_ORDER.index("logquad"), _GROUP_ORDER.index("edge_gradmoment")
        return np.where(size <= float(small_feature_max), lq, em).astype(np.int8)
    return np.full(K, _GROUP_ORDER.index(_ALIASES.get(refine, refine)), dtype=np.int8)


def _refine_rows(img_t, labels, rows_np, cents_np, method, *, pad, fit_half, fit_options, logquad_variant,
                 fit_cells=True):
    """Refine components of one method on one image / strip. rows_np (K, 5) [x, y, w, h, label],
    cents_np (K, 2) CC centroids. Returns (centers (K, 2), ok (K,), extra)."""
    from . import edge_erf, refine as rf

    dev = img_t.device
    K = rows_np.shape[0]
    extra = {}
    stats6 = np.concatenate([rows_np[:, :4].astype(np.float32), cents_np.astype(np.float32)], axis=1)
    if method == "logquad":
        if logquad_variant == "cpu":
            c, ok = rf.refine_logquad_rows_device(img_t, labels, rows_np, pad=pad)
        else:
            c, ok = rf.refine_logquad_window_device(img_t, labels, torch.from_numpy(stats6).to(dev),
                                                    torch.from_numpy(rows_np[:, 4].astype(np.int32)).to(dev),
                                                    pad=pad)
        return c, ok, extra
    if method in ("weighted", "none"):
        c, ok = rf.refine_weighted_rows_device(img_t, labels, rows_np, pad=pad)
        return c, ok, extra
    if method == "edge_gradmoment":
        c, ok = edge_erf.refine_edge_moment_device(img_t, torch.from_numpy(stats6).to(dev))
        return c, ok, extra
    if method == "edge_erf":
        c, ok = edge_erf.refine_edge_erf_device(img_t, torch.from_numpy(stats6).to(dev))
        return c, ok, extra
    if method in _ERF_METHODS:
        from .fit import cell_radius, fit_erf_disks_device, seed_neighbours_device

        seeds = torch.from_numpy(cents_np.astype(np.float32)).to(dev)
        if fit_half is None:
            half_np = np.clip(np.ceil(0.5 * rows_np[:, 2:4].max(axis=1)).astype(np.int64) + 4, 7, 31)
        else:
            half_np = np.broadcast_to(np.asarray(fit_half, dtype=np.int64), (K,))
        half_t = torch.from_numpy(half_np.astype(np.int16)).to(dev)
        model_t = torch.full((K,), _ERF_METHODS[method], dtype=torch.uint8, device=dev)
        nbr = None
        if fit_cells:
            nbr = seed_neighbours_device(seeds, img_t.shape, cell_radius(int(half_np.max())))
        res = fit_erf_disks_device(img_t, seeds, half_t, model_t, options=fit_options, neighbours=nbr)
        torch.cuda.current_stream().synchronize()
        params = res["params"].cpu().numpy()
        status = res["status"].cpu().numpy()
        extra = {"params": params, "status": status, "rms": res["rms"].cpu().numpy(),
                 "iterations": res["iters"].cpu().numpy()}
        ok = (status <= 1) & np.isfinite(params[:, :2]).all(axis=1)
        return res["centers"].cpu().numpy(), ok, extra
    raise ValueError(f"Unsupported refine method: {method}")


def _detect_single_tile(img_t, *, threshold, invert, area_min, area_max, pad, refine, connectivity,
                        pre_threshold, fit_half, fit_options, small_feature_max=12.0, logquad_variant="gpu",
                        morph_open=0, morph_close=0, fit_cells=True, upsample_factor=4):
    """Segmentation -> CC -> area filter -> refinement on one device image (or strip view).
    info["stage_ms"] holds wall times of the stages (each ends with its results on the host)."""
    t0 = time.perf_counter()
    if pre_threshold is None:
        if threshold == "otsu":
            thresh = otsu_threshold_device(img_t)
        elif threshold == "triangle":
            thresh = triangle_threshold_device(img_t)
        else:
            raise ValueError(f"Unsupported threshold method: {threshold!r}. Supported: ['otsu', 'triangle']")
    else:
        thresh = int(np.floor(pre_threshold))
    t1 = time.perf_counter()
    # capacity for a component per 256 px up front: one labelling pass on hole arrays
    cap = int(min(max(65536, img_t.shape[0] * img_t.shape[1] // 256), 1 << 24))
    if morph_open > 0 or morph_close > 0:
        mask = binary_mask_device(img_t, thresh, invert=invert, morph_open=morph_open, morph_close=morph_close)
        num, labels, stats, cents = threshold_components_device(mask, 0, connectivity=connectivity, capacity=cap)
    else:
        num, labels, stats, cents = threshold_components_device(img_t, thresh, invert=invert,
                                                                connectivity=connectivity, capacity=cap)
    stats_np = stats.cpu().numpy()
    cents_np = cents.cpu().numpy()
    t2 = time.perf_counter()
    area = stats_np[:, 4]
    keep = np.nonzero((area >= int(area_min)) & (area <= int(area_max)))[0]
    keep = keep[keep >= 1]
    if refine in _VORONOI_METHODS:
        return _voronoi_tile(img_t, stats_np, cents_np, keep, refine, pad=pad, upsample_factor=upsample_factor,
                             info={"threshold_value": thresh, "num_labels": num}, t=(t0, t1, t2))
    rows = np.concatenate([stats_np[keep, :4], keep[:, None]], axis=1).astype(np.int64)
    methods = _row_methods(refine, rows, small_feature_max)
    out = []
    info = {"threshold_value": thresh, "num_labels": num, "n_candidates": int(keep.size)}
    t3 = time.perf_counter()
    n_refined = 0
    counts = np.bincount(methods, minlength=len(_GROUP_ORDER))
    for code, method in enumerate(_GROUP_ORDER):
        if counts[code] == 0:
            continue
        idx = np.arange(methods.size) if counts[code] == methods.size else np.nonzero(methods == code)[0]
        centers, ok, extra = _refine_rows(img_t, labels, rows[idx], cents_np[keep][idx], method, pad=pad,
                                          fit_half=fit_half, fit_options=fit_options,
                                          logquad_variant=logquad_variant, fit_cells=fit_cells)
        good = ok & np.isfinite(centers).all(axis=1)
        out.append(centers[good])
        n_refined += int(good.sum())
        if extra:
            info["fit"] = {k: v[good] for k, v in extra.items()}
            st = extra["status"]
            info["fit_status_counts"] = {int(s): int((st == s).sum()) for s in np.unique(st)}
    info["n_refined"] = n_refined
    centers = np.concatenate(out, axis=0) if out else np.zeros((0, 2), dtype=np.float64)
    t4 = time.perf_counter()
    info["stage_ms"] = {"threshold": 1e3 * (t1 - t0), "components": 1e3 * (t2 - t1), "rows": 1e3 * (t3 - t2),
                        "refine": 1e3 * (t4 - t3)}
    return centers, info


def _voronoi_tile(img_t, stats_np, cents_np, keep, refine, *, pad, upsample_factor, info, t):
    """The Voronoi dispatch of the reference's single-tile path (subpx/_gpu/centers.py:1545-1630):
    components off the image border seed the cells; centres, the quality measure and the cell
    areas of the finite centres, in component order."""
    from .voronoi import QUALITY_KEYS, refine_rows_device

    t0, t1, t2 = t
    H, W = img_t.shape
    st = stats_np[keep].astype(np.int64)
    inside = (st[:, 0] > 0) & (st[:, 1] > 0) & (st[:, 0] + st[:, 2] < W) & (st[:, 1] + st[:, 3] < H)
    keep = keep[inside]
    rows = np.concatenate([stats_np[keep, :4].astype(np.float64), cents_np[keep].astype(np.float64)], axis=1)
    t3 = time.perf_counter()
    centers, quality, areas = refine_rows_device(img_t, rows, refine, pad=int(pad),
                                                 upsample_factor=int(upsample_factor))
    fin = np.isfinite(centers).all(axis=1)
    t4 = time.perf_counter()
    info = dict(info)
    info.update({"n_candidates": int(keep.size), "n_refined": int(fin.sum()), "upsample_factor": int(upsample_factor),
                 QUALITY_KEYS[refine]: quality[fin], "voronoi_cell_area": areas[fin],
                 "stage_ms": {"threshold": 1e3 * (t1 - t0), "components": 1e3 * (t2 - t1), "rows": 1e3 * (t3 - t2),
                              "refine": 1e3 * (t4 - t3)}})
    return centers[fin], info


def _erf_strip_device(img_t, thresh, *, invert, area_min, area_max, connectivity, fit_half, fit_options,
                      model_code, fit_cells, morph_open=0, morph_close=0, events=None, net=None, lm_iters=4,
                      fallbacks=None, capture=None):
    """Segmentation -> CC -> area filter -> erf fit of one strip, kept on the device: the rows of
    _detect_single_tile's erf group (label order, same seeds, half sides, neighbour lists and
    fit), with every component fitted in one launch and the filtered-out rows (background,
    area outside the limits) given half side 0, which the fit skips. Returns (centres (num, 2)
    float64, ok (num,) bool) device tensors; `events` collects (stage, start, end) CUDA events.
    The only host synchronisation is the component count. With ``net`` (a cnn.LearnedStart)
    the fit starts from the network's proposal for at most ``lm_iters`` steps, and the rows not
    converged then are fitted again from the moment start by the same wave in the same launch
    (``fallbacks`` collects their count per strip as device tensors). ``capture``: a list that
    receives the strip's fit inputs and results as host arrays (seeds, half, model, nbr, params,
    centers, status, ok; synchronising), for checking the fits against a reference."""
    from .fit import cell_radius, fit_erf_disks_device, seed_neighbours_device

    def mark(name, e0):
        if events is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            events.append((name, e0, e1))
            return e1
        return None

    e0 = None
    if events is not None:
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
    cap = int(min(max(65536, img_t.shape[0] * img_t.shape[1] // 256), 1 << 24))
    if morph_open > 0 or morph_close > 0:
        mask = binary_mask_device(img_t, thresh, invert=invert, morph_open=morph_open, morph_close=morph_close)
        num, _, stats, cents = threshold_components_device(mask, 0, connectivity=connectivity, capacity=cap,
                                                           labels_out=False)
    else:
        num, _, stats, cents = threshold_components_device(img_t, thresh, invert=invert, connectivity=connectivity,
                                                           capacity=cap, labels_out=False)
    e0 = mark("components", e0)
    dev = img_t.device
    area = stats[:, 4]
    valid = (area >= int(area_min)) & (area <= int(area_max))
    valid[0] = False
    # seeds of filtered-out rows are NaN: the neighbour lists skip them
    seeds = torch.where(valid[:, None], cents.to(torch.float32),
                        torch.full_like(cents, float("nan"), dtype=torch.float32))
    seeds = seeds.contiguous()
    if fit_half is None:
        size = torch.maximum(stats[:, 2], stats[:, 3])
        half = torch.clamp((size + 1) // 2 + 4, 7, 31)
    else:
        half = torch.full((num,), int(fit_half), dtype=torch.int32, device=dev)
    half = torch.where(valid, half, torch.zeros_like(half)).to(torch.int16).contiguous()
    model = torch.full((num,), int(model_code), dtype=torch.uint8, device=dev)
    e0 = mark("rows", e0)
    nbr = None
    if fit_cells:
        hmax = int(half.max().item()) if num else 0
        nbr = seed_neighbours_device(seeds, img_t.shape, cell_radius(max(hmax, 7)))
    if net is None:
        res = fit_erf_disks_device(img_t, seeds, half, model, options=fit_options, neighbours=nbr)
        params, status, centers = res["params"], res["status"], res["centers"]
    else:
        init = net.propose(img_t, torch.nan_to_num(seeds, nan=0.0).contiguous(), half)
        e0 = mark("propose", e0)
        o = fit_options
        fused = _lib.fit_options(max_iter=o.max_iter, step_tol=o.step_tol, lambda0=o.lambda0, lambda_up=o.lambda_up,
                                 lambda_down=o.lambda_down, init_sigma=o.init_sigma, tile_waves=o.tile_waves,
                                 persistent=o.persistent, bucket_mask=o.bucket_mask, learned_iters=int(lm_iters))
        res = fit_erf_disks_device(img_t, seeds, half, model, init=init, options=fused, neighbours=nbr)
        params, status, centers = res["params"], res["status"], res["centers"]
        if fallbacks is not None:
            fallbacks.append((valid & (res["iters"] > int(lm_iters))).sum())
    ok = valid & (status <= 1) & torch.isfinite(params[:, :2]).all(dim=1)
    mark("refine", e0)
    if capture is not None:
        capture.append({"seeds": seeds.cpu().numpy(), "half": half.cpu().numpy(), "model": model.cpu().numpy(),
                        "nbr": None if nbr is None else nbr.cpu().numpy(), "params": params.cpu().numpy(),
                        "centers": centers.cpu().numpy(), "status": status.cpu().numpy(), "ok": ok.cpu().numpy()})
    return centers, ok


def detect_centers(
    image,
    *,
    backend: str = "auto",
    threshold: str = "triangle",
    invert: bool = False,
    area_min: int = 1,
    area_max: int = 50,
    morph_open: int = 0,
    morph_close: int = 0,
    pad: int = 3,
    refine: str = "logquad",
    connectivity: int = 8,
    gpu_batch: int = 4096,
    device: int = 0,
    use_float64: bool = True,
    components_backend: str = "cpu",
    small_feature_max: float = 12.0,
    upsample_factor: int = 4,
    tile_h: int | None = None,
    overlap: int = 128,
    threshold_mode: str = "auto",
    otsu_downsample: int = 4,
    thr_scale: float = 0.8,
    dedupe_eps: float = 1.5,
    background: str | None = None,
    deblend_iterations: int = 30,
    fit_half=None,
    fit_max_iter: int = 30,
    fit_step_tol: float = 1e-6,
    fit_cells: bool = True,
    logquad_variant: str = "gpu",
    learned_start=None,
    learned_iters: int = 4,
    fit_inputs_out: list | None = None,
) -> CenterResult:
    """Detect and refine centres (reference subpx/centers.py:130-220).

    Same keywords and CenterResult as the reference; ``upsample_factor`` sets the Voronoi
    refinements' upsampling; ``gpu_batch``, ``use_float64`` and ``components_backend`` are
    accepted for signature compatibility
    (the kernels need no batching and always refine in FP64 / FP32-Jacobian + FP64 solve;
    components always come from the GPU kernel, numbered as OpenCV numbers them).
    ``small_feature_max`` splits ``refine="auto"``. ``fit_half``, ``fit_max_iter``,
    ``fit_step_tol`` and ``fit_cells`` (Voronoi cell mask against neighbouring components,
    see fit.py) configure the erf fits; ``logquad_variant`` picks the GPU backend's
    ("gpu", default) or the CPU backend's ("cpu") log-quadratic algorithm.
    ``learned_start`` (a trained cnn.LearnedStart, erf methods only) starts every fit from the
    network's proposal for at most ``learned_iters`` LM steps and refits the features not
    converged then from the moment start. ``fit_inputs_out`` (a list; erf refinement with a
    global threshold) receives, per strip, its row offset ``y0`` and its fits' inputs and results
    as host arrays: a diagnostic for checking the fits against a reference.
    """
    b = resolve_backend(backend)
    if logquad_variant not in ("gpu", "cpu"):
        raise ValueError("logquad_variant must be 'gpu' or 'cpu'")
    if background not in (None, "model"):
        raise ValueError("background must be None or 'model'")
    if background == "model":
        from .deblend import DEBLEND_METHODS, refine_deblended

        if refine not in DEBLEND_METHODS:
            raise NotImplementedError(
                f"background='model' is not supported for refine='{refine}': edge methods read the full "
                "image rather than per-cell ROIs. Use one of " + ", ".join(sorted(DEBLEND_METHODS)) + ".")
        first = detect_centers(
            image, backend=backend, threshold=threshold, invert=invert, area_min=area_min, area_max=area_max,
            morph_open=morph_open, morph_close=morph_close, pad=pad, refine=refine, connectivity=connectivity,
            gpu_batch=gpu_batch, device=device, use_float64=use_float64, components_backend=components_backend,
            small_feature_max=small_feature_max, upsample_factor=upsample_factor, tile_h=tile_h, overlap=overlap,
            threshold_mode=threshold_mode, otsu_downsample=otsu_downsample, thr_scale=thr_scale,
            dedupe_eps=dedupe_eps, logquad_variant=logquad_variant)
        with gpu_device(device) as dev:
            return refine_deblended(to_device_u16(image, dev), first, refine=refine, invert=bool(invert), pad=int(pad),
                                    iterations=int(deblend_iterations), upsample_factor=int(upsample_factor),
                                    tile_h=tile_h, overlap=int(overlap))
    if refine not in _SUPPORTED and refine not in _VORONOI_METHODS:
        raise ValueError(f"Unsupported refine method: {refine}")
    if isinstance(image, torch.Tensor):
        shape = tuple(image.shape)
    else:
        shape = np.shape(image)
    if len(shape) != 2:
        raise ValueError("Expected 2D grayscale image")
    fit_options = _lib.fit_options(max_iter=fit_max_iter, step_tol=fit_step_tol)
    common = dict(threshold=threshold, invert=invert, area_min=area_min, area_max=area_max, pad=pad,
                  refine=refine, connectivity=connectivity, fit_half=fit_half, fit_options=fit_options,
                  small_feature_max=small_feature_max, logquad_variant=logquad_variant,
                  morph_open=int(morph_open), morph_close=int(morph_close), fit_cells=bool(fit_cells),
                  upsample_factor=int(upsample_factor))
    base_meta = {
        "invert": bool(invert), "area_min": int(area_min), "area_max": int(area_max),
        "morph_open": int(morph_open), "morph_close": int(morph_close), "pad": int(pad),
        "connectivity": int(connectivity), "gpu_batch": int(gpu_batch), "device": int(device),
        "use_float64": bool(use_float64), "components_backend": str(components_backend),
        "small_feature_max": float(small_feature_max), "logquad_variant": str(logquad_variant),
    }
    with gpu_device(device) as dev:
        img_t = to_device_u16(image, dev)
        H, W = img_t.shape
        if learned_start is not None and refine not in _ERF_METHODS:
            raise ValueError("learned_start applies to the erf refinements (erf_disk, erf_ellipse)")
        if tile_h is None and learned_start is not None:
            thresh = otsu_threshold_device(img_t) if threshold == "otsu" else triangle_threshold_device(img_t)
            fb = []
            c_t, ok_t = _erf_strip_device(
                img_t, int(thresh), invert=invert, area_min=area_min, area_max=area_max, connectivity=connectivity,
                fit_half=fit_half, fit_options=fit_options, model_code=_ERF_METHODS[refine], fit_cells=bool(fit_cells),
                morph_open=int(morph_open), morph_close=int(morph_close), net=learned_start,
                lm_iters=int(learned_iters), fallbacks=fb)
            meta = dict(base_meta)
            meta.update({"threshold_value": int(thresh), "learned_fallbacks": int(sum(int(v) for v in fb))})
            return CenterResult(centers_xy=c_t[ok_t].cpu().numpy(),
                                method=f"threshold={threshold}, refine={refine}, learned start", backend=b, meta=meta)
        if tile_h is None:
            centers, info = _detect_single_tile(img_t, pre_threshold=None, **common)
            meta = dict(base_meta)
            meta.update(info)
            return CenterResult(centers_xy=centers, method=f"threshold={threshold}, refine={refine}",
                                backend=b, meta=meta)

        tile_h = int(tile_h)
        if overlap >= tile_h:
            raise ValueError("overlap must be < tile_h")
        pre_threshold = None
        stage_ms = {"global_threshold": 0.0, "threshold": 0.0, "components": 0.0, "rows": 0.0, "refine": 0.0,
                    "band": 0.0, "dedupe": 0.0}
        t0 = time.perf_counter()
        if threshold_mode in ("global", "auto"):
            ds = max(1, int(otsu_downsample))
            pre_threshold = float(otsu_threshold_device(img_t, ds)) * float(thr_scale)
        stage_ms["global_threshold"] = 1e3 * (time.perf_counter() - t0)
        step = tile_h - int(overlap)
        all_centers = []
        n_tiles = 0
        y0 = 0
        device_path = refine in _ERF_METHODS and pre_threshold is not None
        learned_fallbacks = None
        if learned_start is not None and not device_path:
            raise ValueError("learned_start needs a global threshold (threshold_mode 'global' or 'auto')")
        pts_dev = None
        if device_path:
            # erf fits keep every strip on the device; one host copy of the centres at the end
            events = []
            dev_centers = []
            fallbacks = []
            tb = time.perf_counter()
            while y0 < H:
                y1 = min(H, y0 + tile_h)
                c_t, ok_t = _erf_strip_device(
                    img_t[y0:y1], int(np.floor(pre_threshold)), invert=invert, area_min=area_min,
                    area_max=area_max, connectivity=connectivity, fit_half=fit_half, fit_options=fit_options,
                    model_code=_ERF_METHODS[refine], fit_cells=bool(fit_cells), morph_open=int(morph_open),
                    morph_close=int(morph_close), events=events, net=learned_start, lm_iters=int(learned_iters),
                    fallbacks=fallbacks, capture=fit_inputs_out)
                if fit_inputs_out is not None:
                    fit_inputs_out[-1]["y0"] = y0
                n_tiles += 1
                half = int(overlap) // 2
                keep_lo = y0 if y0 == 0 else (y0 + half)
                keep_hi = y1 if y1 == H else (y1 - half)
                yy = c_t[:, 1] + y0
                m = ok_t & (yy >= keep_lo) & (yy < keep_hi)
                dev_centers.append(torch.stack([c_t[:, 0], yy], dim=1)[m])
                if y1 == H:
                    break
                y0 += step
            pts_t = torch.cat(dev_centers) if dev_centers else torch.zeros((0, 2), dtype=torch.float64)
            pts = pts_t.cpu().numpy()
            pts_dev = pts_t if pts_t.is_cuda else None
            stage_ms["strips_wall"] = 1e3 * (time.perf_counter() - tb)
            for name, e0, e1 in events:
                stage_ms[name] = stage_ms.get(name, 0.0) + e0.elapsed_time(e1)
            if fallbacks:
                learned_fallbacks = int(torch.stack(fallbacks).sum())
            all_centers = [pts] if len(pts) else []
        while not device_path and y0 < H:
            y1 = min(H, y0 + tile_h)
            tile = img_t[y0:y1]
            tile_thr = None if threshold_mode == "per_tile" else pre_threshold
            centers, tinfo = _detect_single_tile(tile, pre_threshold=tile_thr, **common)
            for k, v in tinfo["stage_ms"].items():
                stage_ms[k] += v
            tb = time.perf_counter()
            n_tiles += 1
            if centers.shape[0] > 0:
                centers = centers.copy()
                centers[:, 1] += y0
                half = int(overlap) // 2
                keep_lo = y0 if y0 == 0 else (y0 + half)
                keep_hi = y1 if y1 == H else (y1 - half)
                m = (centers[:, 1] >= keep_lo) & (centers[:, 1] < keep_hi)
                if np.any(m):
                    all_centers.append(centers[m])
            stage_ms["band"] += 1e3 * (time.perf_counter() - tb)
            if y1 == H:
                break
            y0 += step
        t0 = time.perf_counter()
        if all_centers:
            pts = np.vstack(all_centers)
            pairs = _candidate_pairs_device(pts if pts_dev is None else pts_dev, float(dedupe_eps), (H, W))
            if pairs is None:
                pairs = _candidate_pairs_kdtree(pts, float(dedupe_eps))
            centers = _merge_pairs(pts, pairs, float(dedupe_eps))
        else:
            centers = np.zeros((0, 2), dtype=np.float64)
        stage_ms["dedupe"] = 1e3 * (time.perf_counter() - t0)
    return CenterResult(
        centers_xy=centers,
        method=f"tiled_gpu(refine={refine})",
        backend=b,
        meta={
            "tile_h": int(tile_h), "overlap": int(overlap), "threshold_mode": str(threshold_mode),
            "otsu_downsample": int(otsu_downsample), "thr_scale": float(thr_scale),
            "dedupe_eps": float(dedupe_eps), "refine": str(refine), "invert": bool(invert),
            "area_min": int(area_min), "area_max": int(area_max), "device": int(device),
            "small_feature_max": float(small_feature_max), "upsample_factor": int(upsample_factor),
            "n_tiles": n_tiles, "threshold_value": pre_threshold, "stage_ms": stage_ms,
            "learned_fallbacks": learned_fallbacks,
        },
    )


def filter_centers(centers_xy, image_shape_hw: Union[np.ndarray, Tuple[int, int], Sequence[int]], *,
                   margin: float = 5.0) -> np.ndarray:
    """Keep centres at least ``margin`` px inside the image (reference subpx/centers.py:317-332)."""
    centers = np.asarray(centers_xy, dtype=np.float64)
    if centers.size == 0:
        return centers.reshape(0, 2)
    H, W = int(image_shape_hw[0]), int(image_shape_hw[1])
    m = float(margin)
    x, y = centers[:, 0], centers[:, 1]
    keep = (x >= m) & (x < (W - m)) & (y >= m) & (y < (H - m))
    return centers[keep]


def dedupe_centers(centers_xy, *, eps: float = 1.5) -> np.ndarray:
    """Merge detections closer than ``eps`` (reference subpx/centers.py:335-378).

    Same clusters, same cluster order and same means as the reference's O(N^2) loop: the
    pairs within eps are found with a k-d tree (candidate radius padded, then the
    reference's squared-distance test), and the reference's union-by-rank / path-halving
    union-find is replayed over them in the reference's (i, j) order.
    """
    pts = np.asarray(centers_xy, dtype=np.float64)
    if pts.shape[0] == 0:
        return pts.reshape(0, 2)
    return _merge_pairs(pts, _candidate_pairs_kdtree(pts, eps), eps)


def _candidate_pairs_kdtree(pts: np.ndarray, eps: float) -> np.ndarray:
    from scipy.spatial import cKDTree

    return cKDTree(pts).query_pairs(r=float(eps) * (1.0 + 1e-9) + 1e-12, output_type="ndarray")


def _candidate_pairs_device(pts, eps: float, shape) -> np.ndarray | None:
    """Candidate pairs (i < j) from the device neighbour grid (spx_seed_neighbours) on float32
    coordinates, with the radius padded past their rounding (< 0.01 px below 2^16 px); None
    when a neighbour list is full (it may have dropped a candidate). ``pts`` is a NumPy array
    or a device tensor; only the pairs come back to the host."""
    from .fit import seed_neighbours_device

    nb = 8
    dev = torch.device("cuda", torch.cuda.current_device())
    if isinstance(pts, torch.Tensor):
        s = pts.to(device=dev, dtype=torch.float32).contiguous()
    else:
        s = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float32)).to(dev)
    nbr = seed_neighbours_device(s, shape, float(eps) + 0.02, nb)
    if bool((nbr[:, nb - 1] >= 0).any()):
        return None
    ids = torch.arange(nbr.shape[0], device=dev, dtype=torch.int32)[:, None]
    i, k = torch.nonzero(nbr > ids, as_tuple=True)
    if i.numel() == 0:
        return np.zeros((0, 2), dtype=np.int64)
    return torch.stack([i.to(torch.int64), nbr[i, k].to(torch.int64)], dim=1).cpu().numpy()


def _merge_pairs(pts: np.ndarray, pairs: np.ndarray, eps: float) -> np.ndarray:
    """The reference's union-find over the exact pairs within eps, then cluster means."""
    n = pts.shape[0]
    eps2 = float(eps) ** 2
    parent = np.arange(n, dtype=np.int64)
    if pairs.size:
        pairs = np.sort(pairs, axis=1)
        d2 = np.sum((pts[pairs[:, 1]] - pts[pairs[:, 0]]) ** 2, axis=1)
        pairs = pairs[d2 <= eps2]
    if pairs.size == 0:
        return pts.copy()
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    pairs = pairs[order]
    par = parent.tolist()
    rk = [0] * n

    def find(a):
        while par[a] != a:
            par[a] = par[par[a]]
            a = par[a]
        return a

    for i, j in pairs.tolist():
        ra, rb = find(i), find(j)
        if ra == rb:
            continue
        if rk[ra] < rk[rb]:
            par[ra] = rb
        elif rk[ra] > rk[rb]:
            par[rb] = ra
        else:
            par[rb] = ra
            rk[ra] += 1
    roots = np.arange(n, dtype=np.int64)  # holes in no pair are their own roots
    for i in np.unique(pairs).tolist():
        roots[i] = find(i)
    unique_roots, inv = np.unique(roots, return_inverse=True)
    K = unique_roots.size
    sums = np.zeros((K, 2), dtype=np.float64)
    counts = np.zeros((K,), dtype=np.int64)
    np.add.at(sums, inv, pts)
    np.add.at(counts, inv, 1)
    return sums / np.maximum(counts[:, None], 1)
