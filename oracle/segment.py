"""CPU oracle of the integer front end and the CPU log-quadratic refinement.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker.

* ``otsu_u16``: the reference thresholds with ``cv2.threshold(g, 0, 255, BINARY | OTSU)``
  (reference subpx/_cpu/centers.py:258; global variant subpx/_gpu/centers.py:1789-1796).
  OpenCV (opencv-python >= 4.10, pyproject.toml:17; not installed here) runs, on a
  uint16 image, the Otsu search over the full 65,536-bin histogram; this restates its
  published algorithm step for step: p_i = h_i / N; mu = sum i p_i; for each bin in
  order: mu1 *= q1; q1 += p_i; q2 = 1 - q1; bins with min(q1, q2) < FLT_EPSILON or
  max(q1, q2) > 1 - FLT_EPSILON are skipped (before mu1 is updated); else
  mu1 = (mu1 + i p_i) / q1, mu2 = (mu - q1 mu1) / q2, sigma = q1 q2 (mu1 - mu2)^2 and the
  first strict maximum wins. The threshold is that bin; foreground = g > t
  (THRESH_BINARY) or g <= t (THRESH_BINARY_INV). Pinned by brute force over exact
  rational class moments on small images (tests/test_oracle_segment.py); exact cv2 ties
  are "parity unpinned" (cv2 absent).
* ``connected_components``: cv2.connectedComponentsWithStats(connectivity=8|4)
  (reference subpx/_cpu/components.py:18-34). Partition from scipy.ndimage.label (3x3 or
  cross structure); stats [x, y, w, h, area] and centroids (mean x, y) by brute force;
  row 0 = background pixels. Label numbering: OpenCV's default 8-connected algorithm
  scans 2x2 blocks in raster order and numbers components by their first block; the
  4-connected one by first pixel. The partition, stats and centroids are pinned by
  brute force; the numbering is "parity unpinned" (cv2 absent).
* ``refine_logquad``: reference subpx/_cpu/centers.py:67-107 (+ _bg_from_border :31-42,
  ROI rule :376-391), restated with numpy's own lstsq / cond / solve. Pinned by the exact
  centre of a sampled Gaussian and translation equivariance.
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np
from scipy import ndimage

FLT_EPSILON = float(np.finfo(np.float32).eps)


def histogram_u16(img: np.ndarray, step: int = 1) -> np.ndarray:
    g = np.asarray(img, dtype=np.uint16)[::step, ::step]
    return np.bincount(g.ravel(), minlength=65536).astype(np.int64)


def otsu_from_hist(h: np.ndarray) -> int:
    """OpenCV's Otsu recurrence over a 65,536-bin histogram (float64, sequential)."""
    h = np.asarray(h, dtype=np.int64)
    N = int(h.sum())
    scale = 1.0 / N
    mu = 0.0
    for i in range(h.size):
        mu += i * float(h[i])
    mu *= scale
    mu1 = 0.0
    q1 = 0.0
    max_sigma = 0.0
    max_val = 0
    hs = h.tolist()
    for i in range(len(hs)):
        p_i = hs[i] * scale
        mu1 *= q1
        q1 += p_i
        q2 = 1.0 - q1
        if min(q1, q2) < FLT_EPSILON or max(q1, q2) > 1.0 - FLT_EPSILON:
            continue
        mu1 = (mu1 + i * p_i) / q1
        mu2 = (mu - q1 * mu1) / q2
        sigma = q1 * q2 * (mu1 - mu2) * (mu1 - mu2)
        if sigma > max_sigma:
            max_sigma = sigma
            max_val = i
    return int(max_val)


def otsu_u16(img: np.ndarray, step: int = 1) -> int:
    return otsu_from_hist(histogram_u16(img, step))


def otsu_brute_force(img: np.ndarray) -> int:
    """Textbook Otsu by exact rationals: first t maximising w0 w1 (m0 - m1)^2, g > t."""
    g = np.asarray(img, dtype=np.int64).ravel()
    vals = np.unique(g)
    best, best_t = Fraction(-1), 0
    N = g.size
    for t in range(int(vals.min()), int(vals.max())):
        lo = g[g <= t]
        hi = g[g > t]
        if lo.size == 0 or hi.size == 0:
            continue
        w0 = Fraction(lo.size, N)
        w1 = Fraction(hi.size, N)
        m0 = Fraction(int(lo.sum()), lo.size)
        m1 = Fraction(int(hi.sum()), hi.size)
        s = w0 * w1 * (m0 - m1) ** 2
        if s > best:
            best, best_t = s, t
    return best_t


def triangle_u8(img: np.ndarray) -> int:
    """OpenCV's THRESH_TRIANGLE on 8-bit grey levels (reference call subpx/_cpu/centers.py:258;
    OpenCV defines it for CV_8UC1 only). Bounds = first / last non-empty bins widened by one;
    peak = first maximum; mirrored when the right side of the peak is longer; threshold = the
    bin farthest from the line (left bound, 0)-(peak, h_max) (first maximum of the
    unnormalised distance), minus one, mirrored back."""
    g = np.asarray(img)
    if g.size and int(g.max()) > 255:
        raise ValueError("triangle threshold needs 8-bit grey levels")
    h = np.bincount(g.ravel().astype(np.int64), minlength=256)[:256].astype(np.int64)
    nz = np.nonzero(h)[0]
    lo = int(nz[0]) if nz.size else 0
    hi = int(nz[-1]) if nz.size and nz[-1] > 0 else 0
    lo = lo - 1 if lo > 0 else lo
    hi = hi + 1 if hi < 255 else hi
    peak = int(np.argmax(h))
    flip = peak - lo < hi - peak
    if flip:
        h = h[::-1]
        lo, peak = 255 - hi, 255 - peak
    a, b = float(h[peak]), float(lo - peak)
    best, th = 0.0, lo
    for i in range(lo + 1, peak + 1):
        d = a * i + b * float(h[i])
        if d > best:
            best, th = d, i
    th -= 1
    return 255 - th if flip else th


def morph_cross(mask: np.ndarray, op: str) -> np.ndarray:
    """One erosion / dilation by the 3x3 cross (cv2 3x3 MORPH_ELLIPSE); pixels outside the
    image do not take part (OpenCV's default morphology border)."""
    m = np.asarray(mask, dtype=bool)
    fill = op == "erode"
    p = np.pad(m, 1, constant_values=fill)
    nb = [p[1:-1, 1:-1], p[:-2, 1:-1], p[2:, 1:-1], p[1:-1, :-2], p[1:-1, 2:]]
    if op == "erode":
        return np.logical_and.reduce(nb)
    return np.logical_or.reduce(nb)


def binary_mask(img: np.ndarray, t: float, invert: bool = False, morph_open: int = 0, morph_close: int = 0):
    """``_segment_binary``'s mask (reference subpx/_cpu/centers.py:257-265): threshold, then
    open (erode^n, dilate^n) and close (dilate^n, erode^n) with the 3x3 cross."""
    m = threshold(img, t, invert)
    for op in ["erode"] * morph_open + ["dilate"] * morph_open + ["dilate"] * morph_close + ["erode"] * morph_close:
        m = morph_cross(m, op)
    return m


def threshold(img: np.ndarray, t: float, invert: bool = False) -> np.ndarray:
    g = np.asarray(img)
    return (g <= t) if invert else (g > t)


def connected_components(mask: np.ndarray, connectivity: int = 8):
    """(num_labels, labels int32, stats (n,5) int32, centroids (n,2) f64), OpenCV layout."""
    m = np.asarray(mask) > 0
    H, W = m.shape
    structure = np.ones((3, 3), dtype=bool) if connectivity == 8 else ndimage.generate_binary_structure(2, 1)
    lab, n = ndimage.label(m, structure=structure)
    # canonical numbering
    if n:
        ys, xs = np.nonzero(lab)
        ids = lab[ys, xs]
        if connectivity == 8:
            key = (ys // 2).astype(np.int64) * ((W + 1) // 2) + (xs // 2)
        else:
            key = ys.astype(np.int64) * W + xs
        first = np.full(n + 1, np.iinfo(np.int64).max, dtype=np.int64)
        np.minimum.at(first, ids, key)
        order = np.argsort(first[1:], kind="stable")
        remap = np.zeros(n + 1, dtype=np.int32)
        remap[1 + order] = np.arange(1, n + 1, dtype=np.int32)
        lab = remap[lab]
    labels = lab.astype(np.int32)
    num = n + 1
    stats = np.zeros((num, 5), dtype=np.int32)
    cents = np.zeros((num, 2), dtype=np.float64)
    yy, xx = np.indices((H, W))
    for k in range(num):
        sel = labels == k
        cnt = int(sel.sum())
        if cnt == 0:
            stats[k] = [0, 0, 0, 0, 0]
            cents[k] = [np.nan, np.nan]
            continue
        x = xx[sel]
        y = yy[sel]
        stats[k] = [x.min(), y.min(), x.max() - x.min() + 1, y.max() - y.min() + 1, cnt]
        cents[k] = [x.sum() / cnt, y.sum() / cnt]
    return num, labels, stats, cents


def connected_components_fast(mask: np.ndarray, connectivity: int = 8):
    """Same result as connected_components, vectorised for large masks."""
    m = np.asarray(mask) > 0
    H, W = m.shape
    structure = np.ones((3, 3), dtype=bool) if connectivity == 8 else ndimage.generate_binary_structure(2, 1)
    lab, n = ndimage.label(m, structure=structure)
    ys, xs = np.nonzero(lab)
    ids = lab[ys, xs]
    if n:
        if connectivity == 8:
            key = (ys // 2).astype(np.int64) * ((W + 1) // 2) + (xs // 2)
        else:
            key = ys.astype(np.int64) * W + xs
        first = np.full(n + 1, np.iinfo(np.int64).max, dtype=np.int64)
        np.minimum.at(first, ids, key)
        order = np.argsort(first[1:], kind="stable")
        remap = np.zeros(n + 1, dtype=np.int32)
        remap[1 + order] = np.arange(1, n + 1, dtype=np.int32)
        lab = remap[lab]
        ids = lab[ys, xs]
    labels = lab.astype(np.int32)
    num = n + 1
    stats = np.zeros((num, 5), dtype=np.int64)
    cents = np.zeros((num, 2), dtype=np.float64)
    if n:
        area = np.bincount(ids, minlength=num)
        xmin = np.full(num, W, np.int64)
        ymin = np.full(num, H, np.int64)
        xmax = np.full(num, -1, np.int64)
        ymax = np.full(num, -1, np.int64)
        np.minimum.at(xmin, ids, xs)
        np.minimum.at(ymin, ids, ys)
        np.maximum.at(xmax, ids, xs)
        np.maximum.at(ymax, ids, ys)
        sx = np.bincount(ids, weights=xs.astype(np.float64), minlength=num)
        sy = np.bincount(ids, weights=ys.astype(np.float64), minlength=num)
        stats[1:] = np.stack([xmin, ymin, xmax - xmin + 1, ymax - ymin + 1, area], axis=1)[1:]
        cents[1:, 0] = sx[1:] / area[1:]
        cents[1:, 1] = sy[1:] / area[1:]
    bg = labels == 0
    cnt = int(bg.sum())
    if cnt:
        by, bx = np.nonzero(bg)
        stats[0] = [bx.min(), by.min(), bx.max() - bx.min() + 1, by.max() - by.min() + 1, cnt]
        cents[0] = [bx.sum() / cnt, by.sum() / cnt]
    else:
        cents[0] = [np.nan, np.nan]
    return num, labels, stats.astype(np.int32), cents


def bg_from_border(patch: np.ndarray, border: int = 2) -> float:
    """Median of the 2-px ring (reference subpx/_cpu/centers.py:31-42)."""
    g = np.asarray(patch)
    h, w = g.shape[:2]
    b = int(border)
    if h <= 2 * b or w <= 2 * b:
        return float(np.median(g))
    ring = np.concatenate([g[:b, :].ravel(), g[-b:, :].ravel(), g[:, :b].ravel(), g[:, -b:].ravel()])
    return float(np.median(ring))


def refine_logquadratic(gray_roi: np.ndarray, mask_roi: np.ndarray):
    """(cx, cy) ROI-local or NaN (reference subpx/_cpu/centers.py:67-107)."""
    g = np.asarray(gray_roi, dtype=np.float64)
    m = np.asarray(mask_roi, dtype=bool)
    if np.count_nonzero(m) < 6:
        return (np.nan, np.nan)
    bg = bg_from_border(g, border=2)
    v = g - bg
    v[~m] = 0.0
    vpos = v[m]
    if vpos.size == 0 or np.max(vpos) <= 0:
        return (np.nan, np.nan)
    z = np.log(np.maximum(vpos, 1e-9))
    ys, xs = np.nonzero(m)
    x = xs.astype(np.float64)
    y = ys.astype(np.float64)
    A = np.column_stack([x * x, y * y, x * y, x, y, np.ones_like(x)])
    coef, *_ = np.linalg.lstsq(A, z, rcond=None)
    a, b, c, d, e, _f = coef
    H = np.array([[2 * a, c], [c, 2 * b]], dtype=np.float64)
    if np.linalg.cond(H) > 1e8:
        return (np.nan, np.nan)
    try:
        xy = np.linalg.solve(H, np.array([-d, -e]))
    except np.linalg.LinAlgError:
        return (np.nan, np.nan)
    cx, cy = float(xy[0]), float(xy[1])
    h, w = g.shape
    if not (0 <= cx < w and 0 <= cy < h):
        return (np.nan, np.nan)
    return (cx, cy)


def refine_rows_logquad(img, labels, rows, pad: int = 3):
    """Per row [x, y, w, h, label]: image-coordinate centre + ok, ROI rule of
    reference subpx/_cpu/centers.py:376-391."""
    g = np.asarray(img)
    H, W = g.shape
    out = np.full((len(rows), 2), np.nan)
    ok = np.zeros(len(rows), dtype=bool)
    for k, (x, y, w, h, lab) in enumerate(np.asarray(rows, dtype=np.int64)):
        x0 = max(0, x - pad)
        y0 = max(0, y - pad)
        x1 = min(W, x + w + pad)
        y1 = min(H, y + h + pad)
        cx, cy = refine_logquadratic(g[y0:y1, x0:x1], labels[y0:y1, x0:x1] == lab)
        if np.isfinite(cx) and np.isfinite(cy):
            out[k] = [x0 + cx, y0 + cy]
            ok[k] = True
    return out, ok


def detect_centers_logquad(img, *, threshold_value=None, invert=False, area_min=1, area_max=50,
                           pad=3, connectivity=8):
    """Reference detect_centers_cpu(refine='logquad') with the threshold/CC above
    (reference subpx/_cpu/centers.py:281-408). Returns (centers (N,2), rows used)."""
    t = otsu_u16(img) if threshold_value is None else threshold_value
    bw = threshold(img, t, invert)
    num, labels, stats, _ = connected_components_fast(bw, connectivity)
    rows = [[*stats[k, :4], k] for k in range(1, num) if area_min <= stats[k, 4] <= area_max]
    rows = np.asarray(rows, dtype=np.int64).reshape(-1, 5)
    cent, ok = refine_rows_logquad(img, labels, rows, pad)
    return cent[ok], rows, t
