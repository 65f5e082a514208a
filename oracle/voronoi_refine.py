"""CPU oracle of the GPU detect path's Voronoi refinements ("radial_symmetry",
"isophote_curvature").

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, in NumPy with the reference's float32 / float64 split:

* the component rows of the Voronoi dispatch (subpx/_gpu/centers.py:1545-1575): area filter,
  rows whose bounding box touches the image border dropped, seeds = CC centroids;
* ``compute_voronoi_labels`` (subpx/_gpu/voronoi.py:24-113; subpx/_cpu/cells.py:8-31): every
  pixel takes its nearest seed by float64 squared distance, exact ties to the lowest index;
  the cell areas are the label counts over the whole image (:1585);
* ``_extract_voronoi_rois`` (subpx/_cpu/cells.py:34-84): ROI = bbox +- pad clipped to the
  image; the cell's border (cell pixels with a 4-neighbour outside the cell or on the ROI's
  frame) gives the fill value, its 10th percentile (NumPy linear); ROIs stacked at the
  largest ROI size, filled outside the cell and in the padding, mask = the cell;
* ``batch_bicubic_upsample_f32`` / ``batch_nn_upsample_f32`` (subpx/_gpu/upsample.py:131-231):
  edge-aligned mapping (output 0 -> input 0, last -> last), Keys' cubic a = -0.5 with clamped
  taps, nearest by int(o * scale + 0.5);
* ``_radial_symmetry_inner_gpu`` (subpx/_gpu/centers.py:978-1167) with its defaults
  (:1170-1230: factor 4, margin max(1, factor // 2)): mask eroded by a (2m + 1)^2 square
  (outside the array counts as empty), diagonal gradients at pixel midpoints smoothed by a 3x3
  box with edge padding, slopes and intercepts of the gradient lines, weights |grad|^2 /
  distance to the gradient centroid, the 2x2 weighted least squares of the lines' closest
  point, its weighted squared-distance residual; ROI coordinates back from the upsampled grid;
  centres outside the ROI stack's frame are NaN (residual inf);
* ``_isophote_curvature_inner_gpu`` (subpx/_gpu/centers.py:1233-1413) with its defaults
  (:1416-1469: factor 4, margin = factor, Gaussian pre-smoothing sigma 0.5 upsampled px):
  central first and second differences, displacement to the isophote's centre of curvature
  -I_x |grad|^2 / (I_y^2 I_xx - 2 I_x I_xy I_y + I_x^2 I_yy), votes of curvedness weight
  within half the ROI, their weighted mean and spread.

cupyx's ``gaussian_filter`` and ``binary_erosion`` follow SciPy's definitions; the tests pin
these helpers against scipy.ndimage. Pinned in tests/test_oracle_voronoi.py.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------------------------- cells
def voronoi_labels(seeds_xy, shape, chunk: int = 1 << 16) -> np.ndarray:
    """(H, W) int32 nearest-seed index (ties to the lowest index); -1 without seeds."""
    H, W = shape
    s = np.asarray(seeds_xy, dtype=np.float64).reshape(-1, 2)
    out = np.full(H * W, -1, dtype=np.int32)
    if s.shape[0] == 0:
        return out.reshape(H, W)
    idx = np.arange(H * W, dtype=np.int64)
    for a in range(0, H * W, chunk):
        b = min(H * W, a + chunk)
        px = (idx[a:b] % W).astype(np.float64)
        py = (idx[a:b] // W).astype(np.float64)
        dx = px[:, None] - s[None, :, 0]
        dy = py[:, None] - s[None, :, 1]
        d2 = dx * dx + dy * dy
        out[a:b] = np.argmin(d2, axis=1)  # first minimum = lowest index among ties
    return out.reshape(H, W)


def extract_rois(gray, labels, rows, pad: int = 3):
    """(rois (N, Hm, Wm) float64, masks (N, Hm, Wm) bool, origins [(x0, y0)])."""
    g = np.asarray(gray)
    H, W = g.shape
    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 6)
    N = rows.shape[0]
    if N == 0:
        return np.zeros((0, 0, 0)), np.zeros((0, 0, 0), dtype=bool), []
    x0 = np.maximum(0, rows[:, 0].astype(np.int32) - pad)
    y0 = np.maximum(0, rows[:, 1].astype(np.int32) - pad)
    x1 = np.minimum(W, (rows[:, 0] + rows[:, 2]).astype(np.int32) + pad)
    y1 = np.minimum(H, (rows[:, 1] + rows[:, 3]).astype(np.int32) + pad)
    Hm, Wm = int((y1 - y0).max()), int((x1 - x0).max())
    rois = np.zeros((N, Hm, Wm))
    masks = np.zeros((N, Hm, Wm), dtype=bool)
    for j in range(N):
        roi = g[y0[j]:y1[j], x0[j]:x1[j]].astype(np.float64)
        cell = labels[y0[j]:y1[j], x0[j]:x1[j]] == j
        h, w = cell.shape
        # interior: the cell pixel and its four neighbours inside the ROI and the cell
        inner = np.zeros_like(cell)
        if h > 2 and w > 2:
            inner[1:-1, 1:-1] = (cell[1:-1, 1:-1] & cell[:-2, 1:-1] & cell[2:, 1:-1]
                                 & cell[1:-1, :-2] & cell[1:-1, 2:])
        border = roi[cell & ~inner]
        bg = float(np.percentile(border, 10)) if border.size else 0.0
        rois[j] = bg
        rois[j, :h, :w] = np.where(cell, roi, bg)
        masks[j, :h, :w] = cell
    return rois, masks, list(zip(x0.tolist(), y0.tolist()))


# ----------------------------------------------------------------------------- resampling
def _keys(t):
    """Keys' cubic kernel (a = -0.5) at |t| in [0, 2], evaluated as the reference writes it."""
    t = np.asarray(t, dtype=F32)
    near = ((F32(1.5) * t - F32(2.5)) * t) * t + F32(1.0)
    far = ((F32(-0.5) * t + F32(2.5)) * t - F32(4.0)) * t + F32(2.0)
    return near, far


def upsample_bicubic(rois, factor: int) -> np.ndarray:
    r = np.asarray(rois, dtype=F32)
    N, Hin, Win = r.shape
    Ho, Wo = Hin * factor, Win * factor

    def axis(n_in, n_out):
        # the scale is a float32 quotient, as the reference's kernel divides
        scale = F32(n_in - 1.0) / F32(n_out - 1.0) if n_out > 1 else F32(0.0)
        s = np.arange(n_out, dtype=F32) * scale
        i = np.floor(s).astype(np.int64)
        f = (s - i.astype(F32)).astype(F32)
        w1, _ = _keys(f)
        w2, _ = _keys(F32(1.0) - f)
        _, w0 = _keys(F32(1.0) + f)
        _, w3 = _keys(F32(2.0) - f)
        taps = np.clip(i[:, None] + np.arange(-1, 3)[None, :], 0, n_in - 1)
        return np.stack([w0, w1, w2, w3], axis=1).astype(F32), taps

    wy, ty = axis(Hin, Ho)
    wx, tx = axis(Win, Wo)
    out = np.zeros((N, Ho, Wo), dtype=F32)
    for dy in range(4):
        for dx in range(4):  # the reference's accumulation order: dy outer, dx inner
            w = (wy[:, dy][:, None] * wx[:, dx][None, :]).astype(F32)
            out = (out + w[None] * r[:, ty[:, dy]][:, :, tx[:, dx]]).astype(F32)
    return out


def upsample_nearest(arr, factor: int) -> np.ndarray:
    a = np.asarray(arr)
    N, Hin, Win = a.shape
    Ho, Wo = Hin * factor, Win * factor

    def idx(n_in, n_out):
        scale = F32(n_in - 1.0) / F32(n_out - 1.0) if n_out > 1 else F32(0.0)
        return np.minimum((np.arange(n_out, dtype=F32) * scale + F32(0.5)).astype(np.int64), n_in - 1)

    return a[:, idx(Hin, Ho)][:, :, idx(Win, Wo)]


def erode_square(mask, margin: int) -> np.ndarray:
    """Binary erosion of (N, H, W) masks by a (2 margin + 1)^2 square; outside counts as empty."""
    m = np.asarray(mask, dtype=bool)
    if margin <= 0:
        return m.copy()
    N, H, W = m.shape
    p = np.zeros((N, H + 2 * margin, W + 2 * margin), dtype=bool)
    p[:, margin:margin + H, margin:margin + W] = m
    out = np.ones_like(m)
    for dy in range(2 * margin + 1):
        for dx in range(2 * margin + 1):
            out &= p[:, dy:dy + H, dx:dx + W]
    return out


def box3_edge(a) -> np.ndarray:
    """3x3 mean of (N, H, W) float32 with edge replication, divided by 9 as float32."""
    p = np.pad(a, ((0, 0), (1, 1), (1, 1)), mode="edge")
    s = (p[:, 0:-2, 0:-2] + p[:, 0:-2, 1:-1] + p[:, 0:-2, 2:] + p[:, 1:-1, 0:-2] + p[:, 1:-1, 1:-1]
         + p[:, 1:-1, 2:] + p[:, 2:, 0:-2] + p[:, 2:, 1:-1] + p[:, 2:, 2:])
    return (s / F32(9.0)).astype(F32)


def gaussian_weights(sigma: float, truncate: float = 4.0) -> np.ndarray:
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-0.5 / (float(sigma) ** 2) * x * x)
    return w / w.sum()


def gaussian_filter_yx(a, sigma: float) -> np.ndarray:
    """Separable Gaussian over the last two axes of float32 (N, H, W), mode "reflect"
    (d c b a | a b c d | d c b a), rows first then columns, float32 after each pass."""
    w = gaussian_weights(sigma)
    r = (len(w) - 1) // 2
    out = np.asarray(a, dtype=F32)
    for ax in (1, 2):
        n = out.shape[ax]
        src = np.arange(-r, n + r)
        # half-sample symmetric reflection, repeated for short axes
        period = 2 * n
        m = np.mod(src, period)
        src = np.where(m >= n, period - 1 - m, m)
        taken = np.take(out.astype(np.float64), src, axis=ax)
        acc = np.zeros(out.shape, dtype=np.float64)
        for k in range(len(w)):
            sl = [slice(None)] * 3
            sl[ax] = slice(k, k + n)
            acc += w[k] * taken[tuple(sl)]
        out = acc.astype(F32)
    return out


# ------------------------------------------------------------------------------ estimators
def _back_to_roi(xc, yc, Hm, Wm, up_h, up_w, factor):
    xp = xc + (up_w - 1.0) / 2.0
    yp = yc + (up_h - 1.0) / 2.0
    if factor > 1:
        x = xp * (Wm - 1.0) / max(up_w - 1.0, 1.0)
        y = yp * (Hm - 1.0) / max(up_h - 1.0, 1.0)
    else:
        x, y = xp, yp
    return x, y


def radial_symmetry(rois, masks, *, upsample_factor: int = 4, boundary_margin: int | None = None):
    """(centers (N, 2) ROI px [x, y], residuals (N,))."""
    rois = np.asarray(rois, dtype=np.float64)
    masks = np.asarray(masks, dtype=bool)
    N, Hm, Wm = rois.shape
    if N == 0:
        return np.zeros((0, 2)), np.zeros(0)
    factor = int(upsample_factor)
    margin = max(1, factor // 2) if boundary_margin is None else int(boundary_margin)
    if factor > 1:
        g = upsample_bicubic(rois.astype(F32), factor)
        m = upsample_nearest(masks.astype(F32), factor) > 0.5
    else:
        g = rois.astype(F32)
        m = masks.copy()
    up_h, up_w = g.shape[1], g.shape[2]
    eff = min(margin, max(0, (min(up_h, up_w) - 11) // 2))
    me = erode_square(m, eff)
    Hp, Wp = up_h - 1, up_w - 1
    if Hp < 1 or Wp < 1:
        return np.full((N, 2), np.nan), np.full(N, np.inf)
    dIdu = (g[:, :Hp, 1:Wp + 1] - g[:, 1:Hp + 1, :Wp]).astype(F32)
    dIdv = (g[:, :Hp, :Wp] - g[:, 1:Hp + 1, 1:Wp + 1]).astype(F32)
    if Hp >= 3 and Wp >= 3:
        dIdu, dIdv = box3_edge(dIdu), box3_edge(dIdv)
    valid = me[:, :Hp, :Wp] & me[:, :Hp, 1:Wp + 1] & me[:, 1:Hp + 1, :Wp] & me[:, 1:Hp + 1, 1:Wp + 1]
    denom = (dIdu - dIdv).astype(F32)
    num = (-(dIdv + dIdu)).astype(F32)
    near = np.abs(denom) < F32(1e-6)
    slope = (num / np.where(near, F32(1.0), denom)).astype(F32)
    slope = np.where(near, np.where(num >= 0, F32(1e9), F32(-1e9)), slope).astype(F32)
    xm = (np.arange(Wp, dtype=F32)[None, None, :] + F32(0.5)) - F32((up_w - 1.0) / 2.0)
    ym = (np.arange(Hp, dtype=F32)[None, :, None] + F32(0.5)) - F32((up_h - 1.0) / 2.0)
    xm = np.broadcast_to(xm, (N, Hp, Wp)).astype(F32)
    ym = np.broadcast_to(ym, (N, Hp, Wp)).astype(F32)
    b = (ym - slope * xm).astype(F32)
    g2 = (dIdu * dIdu + dIdv * dIdv).astype(F32)
    wb = np.where(valid, g2, F32(0.0)).astype(F32)
    wsum = np.maximum(wb.sum(axis=(1, 2), dtype=np.float64), 1e-30)
    gcx = ((wb * xm).sum(axis=(1, 2), dtype=np.float64) / wsum).astype(F32)
    gcy = ((wb * ym).sum(axis=(1, 2), dtype=np.float64) / wsum).astype(F32)
    d2 = ((xm - gcx[:, None, None]) ** 2 + (ym - gcy[:, None, None]) ** 2).astype(F32)
    dist = np.sqrt(np.maximum(d2, F32(1e-30))).astype(F32)
    w = np.where(valid, (g2 / dist).astype(F32), F32(0.0)).astype(np.float64)
    s = slope.astype(np.float64)
    bb = b.astype(np.float64)
    wm = w / (s * s + 1.0)
    sw = wm.sum(axis=(1, 2))
    smm = (s * s * wm).sum(axis=(1, 2))
    sm = (s * wm).sum(axis=(1, 2))
    smb = (s * bb * wm).sum(axis=(1, 2))
    sb = (bb * wm).sum(axis=(1, 2))
    det = smm * sw - sm * sm
    bad = np.abs(det) < 1e-30
    dets = np.where(bad, 1.0, det)
    xc = np.where(bad, np.nan, (-smb * sw + sm * sb) / dets)
    yc = np.where(bad, np.nan, (smm * sb - smb * sm) / dets)
    perp = (s * xc[:, None, None] - yc[:, None, None] + bb) ** 2 / (s * s + 1.0)
    resid = (w * perp).sum(axis=(1, 2)) / np.maximum(sw, 1e-30)
    x, y = _back_to_roi(xc, yc, Hm, Wm, up_h, up_w, factor)
    c = np.stack([x, y], axis=1)
    oob = (c[:, 0] < 0) | (c[:, 0] >= Wm) | (c[:, 1] < 0) | (c[:, 1] >= Hm)
    c[oob] = np.nan
    resid[oob] = np.inf
    return c, resid


def isophote_curvature(rois, masks, *, upsample_factor: int = 4, boundary_margin: int | None = None,
                       pre_smooth_sigma: float = 0.5):
    """(centers (N, 2) ROI px [x, y], spreads (N,))."""
    rois = np.asarray(rois, dtype=np.float64)
    masks = np.asarray(masks, dtype=bool)
    N, Hm, Wm = rois.shape
    if N == 0:
        return np.zeros((0, 2)), np.zeros(0)
    factor = int(upsample_factor)
    margin = factor if boundary_margin is None else int(boundary_margin)
    if factor > 1:
        g = upsample_bicubic(rois.astype(F32), factor)
        m = upsample_nearest(masks.astype(F32), factor) > 0.5
    else:
        g = rois.astype(F32)
        m = masks.copy()
    if pre_smooth_sigma > 0:
        g = gaussian_filter_yx(g, pre_smooth_sigma)
    up_h, up_w = g.shape[1], g.shape[2]
    eff = min(margin, max(0, (min(up_h, up_w) - 5) // 2))
    me = erode_square(m, eff)
    Hd, Wd = up_h - 2, up_w - 2
    if Hd < 1 or Wd < 1:
        return np.full((N, 2), np.nan), np.full(N, np.inf)
    c = g[:, 1:-1, 1:-1]
    Ix = ((g[:, 1:-1, 2:] - g[:, 1:-1, :-2]) / F32(2.0)).astype(F32)
    Iy = ((g[:, 2:, 1:-1] - g[:, :-2, 1:-1]) / F32(2.0)).astype(F32)
    Ixx = (g[:, 1:-1, 2:] - F32(2.0) * c + g[:, 1:-1, :-2]).astype(F32)
    Iyy = (g[:, 2:, 1:-1] - F32(2.0) * c + g[:, :-2, 1:-1]).astype(F32)
    Ixy = ((g[:, 2:, 2:] - g[:, 2:, :-2] - g[:, :-2, 2:] + g[:, :-2, :-2]) / F32(4.0)).astype(F32)
    eps = F32(1e-6)
    g2 = (Ix * Ix + Iy * Iy).astype(F32)
    den = (Iy * Iy * Ixx - F32(2.0) * Ix * Ixy * Iy + Ix * Ix * Iyy).astype(F32)
    valid = me[:, 1:-1, 1:-1] & (np.abs(den) > eps) & (g2 > eps)
    sden = np.where(np.abs(den) > eps, den, F32(1.0)).astype(F32)
    Dx = (-Ix * g2 / sden).astype(F32)
    Dy = (-Iy * g2 / sden).astype(F32)
    curv = np.sqrt(Ixx * Ixx + F32(2.0) * Ixy * Ixy + Iyy * Iyy).astype(F32)
    xp = (np.arange(Wd, dtype=F32)[None, None, :] + F32(1.0)) - F32((up_w - 1.0) / 2.0)
    yp = (np.arange(Hd, dtype=F32)[None, :, None] + F32(1.0)) - F32((up_h - 1.0) / 2.0)
    vx = (xp + Dx).astype(F32)
    vy = (yp + Dy).astype(F32)
    half = F32(max(up_w, up_h) / 2.0)
    valid = valid & ((Dx * Dx + Dy * Dy).astype(F32) < half * half)
    w = np.where(valid, curv, F32(0.0)).astype(np.float64)
    vx64, vy64 = vx.astype(np.float64), vy.astype(np.float64)
    ws = np.maximum(w.sum(axis=(1, 2)), 1e-30)
    xc = (w * vx64).sum(axis=(1, 2)) / ws
    yc = (w * vy64).sum(axis=(1, 2)) / ws
    spread = np.sqrt((w * ((vx64 - xc[:, None, None]) ** 2 + (vy64 - yc[:, None, None]) ** 2)).sum(axis=(1, 2))
                     / ws)
    x, y = _back_to_roi(xc, yc, Hm, Wm, up_h, up_w, factor)
    cc = np.stack([x, y], axis=1)
    oob = (cc[:, 0] < 0) | (cc[:, 0] >= Wm) | (cc[:, 1] < 0) | (cc[:, 1] >= Hm)
    cc[oob] = np.nan
    spread[oob] = np.inf
    return cc, spread


# ------------------------------------------------------------------------------ pipeline
def voronoi_rows(rows6, shape):
    """Rows [x, y, w, h, cx, cy] whose bounding box stays off the image border."""
    H, W = shape
    r = np.asarray(rows6, dtype=np.float64).reshape(-1, 6)
    x, y, w, h = r[:, 0], r[:, 1], r[:, 2], r[:, 3]
    keep = (x > 0) & (y > 0) & (x + w < W) & (y + h < H)
    return r[keep]


def detect_voronoi(img, refine, *, threshold="otsu", invert=False, area_min=1, area_max=50, pad=3,
                   upsample_factor=4, connectivity=8, threshold_value=None):
    """(centres (K, 2), meta) of the GPU detect path with a Voronoi refinement."""
    from . import segment

    if threshold_value is not None:
        t = threshold_value
    else:
        t = segment.otsu_u16(img) if threshold == "otsu" else segment.triangle_u8(img)
    mask = segment.binary_mask(img, t, invert, 0, 0)
    num, _, stats, cents = segment.connected_components_fast(mask, connectivity)
    keep = [k for k in range(1, num) if area_min <= stats[k, 4] <= area_max]
    rows6 = np.concatenate([stats[keep, :4].astype(np.float64), np.asarray(cents, dtype=np.float64)[keep]],
                           axis=1).reshape(-1, 6)
    rows = voronoi_rows(rows6, img.shape)
    key = "radial_symmetry_residual" if refine == "radial_symmetry" else "isophote_curvature_spread"
    if rows.shape[0] == 0:
        return np.zeros((0, 2)), {key: np.zeros(0), "voronoi_cell_area": np.zeros(0, np.int64)}
    labels = voronoi_labels(rows[:, 4:6], img.shape)
    areas = np.bincount(labels.ravel(), minlength=rows.shape[0])
    rois, masks, origins = extract_rois(img, labels, rows, pad=pad)
    fn = radial_symmetry if refine == "radial_symmetry" else isophote_curvature
    local, q = fn(rois, masks, upsample_factor=upsample_factor)
    g = local + np.asarray(origins, dtype=np.float64)
    fin = np.isfinite(g).all(axis=1)
    return g[fin], {key: q[fin], "voronoi_cell_area": areas[:rows.shape[0]][fin].astype(np.int64),
                    "rois": (rois, masks, origins), "labels": labels}
