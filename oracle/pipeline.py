"""CPU oracle of the GPU detect path with every refinement method.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates reference _detect_single_tile_gpu (subpx/_gpu/centers.py:1632-1710): Otsu threshold
and 8-connected components (oracle/segment.py), area filter, rows [x, y, w, h, cx, cy] in
float32, per-row method (``auto``: logquad when max(w, h) <= small_feature_max, else
edge_gradmoment, subpx/_cpu/centers.py:22-30; aliases as :1640-1645), and centres emitted
group by group in the order logquad, edge_gradmoment, edge_erf, weighted, none, keeping
accepted, finite centres. ``logquad_variant="cpu"`` uses the CPU backend's per-label fit.
"""

from __future__ import annotations

import numpy as np

from . import edge_erf, logquad_window, segment

ALIASES = {"logquadratic": "logquad", "edge_moment": "edge_gradmoment", "edges_centered": "edge_gradmoment",
           "edge_gradmoment_erf": "edge_erf", "edge_erf_centered": "edge_erf"}
ORDER = ("logquad", "edge_gradmoment", "edge_erf", "weighted", "none")


def component_rows(img, *, threshold_value=None, invert=False, area_min=1, area_max=50, connectivity=8,
                   threshold="otsu", morph_open=0, morph_close=0):
    if threshold_value is not None:
        t = threshold_value
    else:
        t = segment.otsu_u16(img) if threshold == "otsu" else segment.triangle_u8(img)
    mask = segment.binary_mask(img, t, invert, morph_open, morph_close)
    num, labels, stats, cents = segment.connected_components_fast(mask, connectivity)
    keep = np.array([k for k in range(1, num) if area_min <= stats[k, 4] <= area_max], dtype=np.int64)
    rows5 = np.concatenate([stats[keep, :4], keep[:, None]], axis=1).reshape(-1, 5).astype(np.int64)
    rows6 = np.concatenate([stats[keep, :4], cents[keep]], axis=1).reshape(-1, 6).astype(np.float32)
    return t, labels, rows5, rows6


def refine_method(img, labels, rows5, rows6, method, *, pad=3, logquad_variant="gpu"):
    g = np.asarray(img)
    if method == "logquad":
        if logquad_variant == "cpu":
            return segment.refine_rows_logquad(g, labels, rows5, pad)
        return logquad_window.refine_logquad_window(g.astype(np.float64), labels, rows6, rows5[:, 4], pad=pad)
    if method in ("weighted", "none"):
        return logquad_window.refine_weighted_rows(g, labels, rows5, pad=pad)
    if method == "edge_gradmoment":
        return edge_erf.refine_edge_moment(g.astype(np.float32), rows6)
    if method == "edge_erf":
        return edge_erf.refine_edge_erf(g.astype(np.float32), rows6)
    raise ValueError(method)


def detect_centers(img, *, threshold_value=None, invert=False, area_min=1, area_max=50, pad=3, connectivity=8,
                   refine="logquad", small_feature_max=12.0, logquad_variant="gpu", threshold="otsu",
                   morph_open=0, morph_close=0):
    """(centers (N, 2), threshold) of the GPU detect path on one image / strip."""
    t, labels, rows5, rows6 = component_rows(img, threshold_value=threshold_value, invert=invert,
                                             area_min=area_min, area_max=area_max, connectivity=connectivity,
                                             threshold=threshold, morph_open=morph_open, morph_close=morph_close)
    if refine == "auto":
        size = np.maximum(rows5[:, 2], rows5[:, 3]).astype(np.float64)
        methods = np.where(size <= float(small_feature_max), "logquad", "edge_gradmoment")
    else:
        methods = np.full(rows5.shape[0], ALIASES.get(refine, refine), dtype=object)
    out = []
    for m in ORDER:
        idx = np.nonzero(methods == m)[0]
        if idx.size == 0:
            continue
        c, ok = refine_method(img, labels, rows5[idx], rows6[idx], m, pad=pad, logquad_variant=logquad_variant)
        good = ok & np.isfinite(c).all(axis=1)
        out.append(c[good])
    return (np.concatenate(out, axis=0) if out else np.zeros((0, 2))), t
