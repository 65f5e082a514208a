"""FP64 CPU oracle of the GPU backend's "logquad" refinement and of its weighted centroid.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``refine_logquad_window`` restates, one feature at a time, the reference's
``refine_centers_logquad_gpu_match_cpu`` (subpx/_gpu/centers.py:288-517) with its helpers
``_offsets_square`` (:125-137), ``_masked_median_lower`` (:145-153) and ``_plane_fit_batched``
(:264-285):

1. integer centre = rint of the component centroid; outer window (2R+1)^2, R = win_rad +
   bg_win_pad, pixels outside the image read 0 with label 0 (:337-352);
2. ROI = bbox +- pad clipped to the image (:327-330); blob = label match in the outer window,
   dilated by a (2 bg_dilate + 1)^2 maximum filter clipped to the window (:361-368: cupyx
   ``maximum_filter`` with mode "reflect", which for a maximum is the clipped window);
3. background = lower median (sorted index (n - 1) // 2) of the ROI pixels outside the
   dilated blob, valid with >= min_bg_pts points (:370-373); tail re-median over the
   background pixels with g - bg0 <= bg_tail_frac * max over the fit window of g - bg0
   (:383-393);
4. plane a x + b y + c fitted by 3x3 normal equations in absolute image coordinates over
   the background pixels used, accepted with >= min_plane_pts points, |det| > 1e-18 and
   >= 3 points; otherwise the scalar background (:395-410);
5. fit window (2 win_rad + 1)^2: blob & ROI dilated by (2 fit_dilate + 1)^2 (clipped) & ROI;
   fit = blob & v > 0; core = fit & v >= core_frac * max(v); sel = core if it has >= 6
   pixels, else fit; >= 6 selected pixels needed (:415-433);
6. weights v^weight_power on sel, clipped at their lower-order statistic of index
   floor(0.99 (n - 1)) (:435-443); z = log(max(v, eps));
7. weighted least squares of z on [xn^2, yn^2, xn yn, xn, yn, 1], xn = (x - mean) / scale with
   scale = max(std x, std y, 1) over sel (:445-473); LU solve;
8. H = [[2a, c], [c, 2b]]: both eigenvalues < -negdef_eps, |l2| / |l1| <= max_cond,
   |det| > 1e-20; centre = mean + scale H^-1 (-d, -e) within +-win_rad of the float
   centroid; otherwise ok = False and the centroid is returned (:475-515).

``refine_weighted_rows`` restates the GPU detect path's "weighted" / "none" refinement
(:1687-1710 -> ``_refine_batch_gpu`` :931-955 -> ``_weighted_from_batch_gpu`` :43-66) with the
CPU border median ``_bg_from_border`` (subpx/_cpu/centers.py:31-42) that ``_pad_rois`` (:33-46)
applies per ROI.

Pinned in tests/test_oracle_logquad_window.py: exact recovery of log-quadratic spots
(circular, elliptical, rotated) on constant and sloped backgrounds, the lower median and the
99th-percentile index against numpy's "lower" quantile, the clipped dilation against
scipy.ndimage.maximum_filter(mode="reflect"), the plane against numpy lstsq, and the
weighted centroid against closed forms.
"""

from __future__ import annotations

import math

import numpy as np

from .segment import bg_from_border

DEFAULTS = dict(pad=3, win_rad=4, core_frac=0.20, weight_power=1.0, bg_dilate=2, bg_win_pad=2, bg_tail_frac=0.05,
                plane_bg=True, min_bg_pts=12, min_plane_pts=20, max_cond=1e10, negdef_eps=1e-12, fit_dilate=1,
                eps=1e-12)


def lower_median(vals) -> float:
    v = np.sort(np.asarray(vals, dtype=np.float64).ravel())
    if v.size == 0:
        return math.inf
    return float(v[(v.size - 1) // 2])


def dilate_clipped(mask: np.ndarray, rad: int) -> np.ndarray:
    """Binary dilation by a (2 rad + 1)^2 square, neighbourhood clipped to the array."""
    m = np.asarray(mask, dtype=bool)
    out = np.zeros_like(m)
    H, W = m.shape
    for dy in range(-rad, rad + 1):
        for dx in range(-rad, rad + 1):
            ys0, ys1 = max(0, dy), min(H, H + dy)
            xs0, xs1 = max(0, dx), min(W, W + dx)
            out[ys0 - dy:ys1 - dy, xs0 - dx:xs1 - dx] |= m[ys0:ys1, xs0:xs1]
    return out


def plane_fit(x, y, g):
    """(a, b, c, ok) of g ~ a x + b y + c by 3x3 normal equations (reference :264-285)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    g = np.asarray(g, np.float64)
    M = np.array([[np.sum(x * x), np.sum(x * y), np.sum(x)],
                  [np.sum(x * y), np.sum(y * y), np.sum(y)],
                  [np.sum(x), np.sum(y), float(x.size)]])
    rhs = np.array([np.sum(x * g), np.sum(y * g), np.sum(g)])
    det = np.linalg.det(M)
    ok = bool(np.isfinite(det) and abs(det) > 1e-18 and x.size >= 3)
    if not ok:
        return 0.0, 0.0, 0.0, False
    a, b, c = np.linalg.solve(M, rhs)
    return float(a), float(b), float(c), True


def _solve6(M, rhs):
    try:
        return np.linalg.solve(M, rhs)
    except np.linalg.LinAlgError:
        return np.full(6, np.nan)


def refine_one_logquad_window(g, labels, row6, lab, o):
    """One feature. g (H, W) float64, labels (H, W) int, row6 [x, y, w, h, cx, cy] (float32
    values), lab = component label. Returns (rx, ry, ok)."""
    H, W = g.shape
    x, y, w, h = (int(np.float32(v)) for v in row6[:4])
    pad = int(o["pad"])
    x0, y0 = max(0, x - pad), max(0, y - pad)
    x1, y1 = min(W, x + w + pad), min(H, y + h + pad)
    cxf, cyf = float(np.float32(row6[4])), float(np.float32(row6[5]))
    cxi, cyi = int(np.rint(cxf)), int(np.rint(cyf))
    wr = int(o["win_rad"])
    R = wr + int(o["bg_win_pad"])
    off = np.arange(-R, R + 1)
    gy = cyi + off[:, None] + 0 * off[None, :]
    gx = cxi + off[None, :] + 0 * off[:, None]
    inside = (gx >= 0) & (gx < W) & (gy >= 0) & (gy < H)
    gyc, gxc = np.clip(gy, 0, H - 1), np.clip(gx, 0, W - 1)
    gout = np.where(inside, g[gyc, gxc], 0.0)
    lout = np.where(inside, labels[gyc, gxc], 0)
    roi = (gx >= x0) & (gx < x1) & (gy >= y0) & (gy < y1)
    m0 = lout == lab
    dil = dilate_clipped(m0, int(o["bg_dilate"])) if int(o["bg_dilate"]) > 0 else m0
    bgm = roi & ~dil
    bg0 = lower_median(gout[bgm])
    ok_bg = np.isfinite(bg0) and int(bgm.sum()) >= int(o["min_bg_pts"])

    s = slice(int(o["bg_win_pad"]), int(o["bg_win_pad"]) + 2 * wr + 1)
    gw, lw, roiw = gout[s, s], lout[s, s], roi[s, s]
    xw, yw = gx[s, s].astype(np.float64), gy[s, s].astype(np.float64)
    v0 = gw - bg0
    vmax0 = float(v0[roiw].max()) if roiw.any() else -math.inf
    used = bgm
    if float(o["bg_tail_frac"]) > 0.0:
        with np.errstate(invalid="ignore"):
            bgm2 = bgm & ((gout - bg0) <= float(o["bg_tail_frac"]) * vmax0)
        bg02 = lower_median(gout[bgm2])
        if np.isfinite(bg02) and int(bgm2.sum()) >= int(o["min_bg_pts"]):
            bg0, used = bg02, bgm2
    if o["plane_bg"]:
        ok_plane = ok_bg and int(used.sum()) >= int(o["min_plane_pts"])
        a, b, c, okp = plane_fit(gx[used], gy[used], gout[used])
        ok_plane = ok_plane and okp
        v = gw - (a * xw + b * yw + c) if ok_plane else gw - bg0
    else:
        v = gw - bg0

    inb = (lw == lab) & roiw
    if int(o["fit_dilate"]) > 0:
        inb = dilate_clipped(inb, int(o["fit_dilate"])) & roiw
    with np.errstate(invalid="ignore"):
        fit = inb & (v > 0) & np.isfinite(v)
    vmax = float(v[fit].max()) if fit.any() else -math.inf
    ok_fit = ok_bg and np.isfinite(vmax) and vmax > 0
    core = fit & (v >= float(o["core_frac"]) * vmax)
    sel = core if int(core.sum()) >= 6 else fit
    n_sel = int(sel.sum())
    ok_sel = ok_fit and n_sel >= 6

    vv = np.maximum(v, float(o["eps"]))
    z = np.log(vv)
    wts = np.where(sel, vv ** float(o["weight_power"]), 0.0)
    if n_sel > 0:
        k = max(int(math.floor(0.99 * (n_sel - 1.0))), 0)
        thr99 = float(np.sort(wts[sel])[k])
    else:
        thr99 = 0.0
    wts = np.minimum(wts, thr99)

    n = max(float(n_sel), 1.0)
    xm, ym = float(xw[sel].sum()) / n, float(yw[sel].sum()) / n
    varx = max(float((xw[sel] ** 2).sum()) / n - xm * xm, 0.0)
    vary = max(float((yw[sel] ** 2).sum()) / n - ym * ym, 0.0)
    scl = max(math.sqrt(varx), math.sqrt(vary), 1.0)
    xn, yn = ((xw - xm) / scl).ravel(), ((yw - ym) / scl).ravel()
    P = np.stack([xn * xn, yn * yn, xn * yn, xn, yn, np.ones_like(xn)], axis=1)
    wv = wts.ravel()
    if ok_sel:
        coef = _solve6((P * wv[:, None]).T @ P, P.T @ (wv * z.ravel()))
    else:
        coef = np.zeros(6)
    a2, b2, c2, d2, e2 = coef[:5]
    H00, H11, H01 = 2.0 * a2, 2.0 * b2, c2
    tr = H00 + H11
    disc = math.sqrt(max((H00 - H11) ** 2 + 4.0 * H01 * H01, 0.0)) if np.isfinite(tr) else math.nan
    lam1, lam2 = 0.5 * (tr - disc), 0.5 * (tr + disc)
    ok = ok_sel and lam1 < -float(o["negdef_eps"]) and lam2 < -float(o["negdef_eps"])
    ok = ok and abs(lam2) / max(abs(lam1), 1e-30) <= float(o["max_cond"])
    det = H00 * H11 - H01 * H01
    ok = ok and abs(det) > 1e-20
    if ok:
        cxg = xm + scl * ((-d2 * H11 + e2 * H01) / det)
        cyg = ym + scl * ((d2 * H01 - e2 * H00) / det)
        ok = (cxf - wr <= cxg <= cxf + wr) and (cyf - wr <= cyg <= cyf + wr)
    if ok:
        return cxg, cyg, True
    return cxf, cyf, False


def refine_logquad_window(gray, labels, rows6, lab_ids, **kw):
    """(centers (K, 2), ok (K,)) for rows [x, y, w, h, cx, cy] and their labels."""
    o = dict(DEFAULTS)
    o.update(kw)
    g = np.asarray(gray, dtype=np.float64)
    rows6 = np.asarray(rows6, dtype=np.float32)
    K = rows6.shape[0]
    out = np.zeros((K, 2))
    ok = np.zeros(K, bool)
    for k in range(K):
        rx, ry, good = refine_one_logquad_window(g, labels, rows6[k], int(lab_ids[k]), o)
        out[k] = rx, ry
        ok[k] = good
    return out, ok


def refine_weighted_rows(gray, labels, rows5, pad=3):
    """(centers (K, 2), ok (K,)) of the GPU detect path's weighted centroid, rows [x, y, w, h, label]."""
    g = np.asarray(gray)
    H, W = g.shape
    rows5 = np.asarray(rows5, dtype=np.int64).reshape(-1, 5)
    out = np.full((rows5.shape[0], 2), np.nan)
    for k, (x, y, w, h, lab) in enumerate(rows5):
        x0, y0 = max(0, x - pad), max(0, y - pad)
        x1, y1 = min(W, x + w + pad), min(H, y + h + pad)
        roi = g[y0:y1, x0:x1]
        m = labels[y0:y1, x0:x1] == lab
        bg = bg_from_border(roi)
        gr = roi.astype(np.float64)
        wts = np.where(m, np.maximum(gr - bg, 0.0), 0.0)
        s = wts.sum()
        yy, xx = np.indices(gr.shape, dtype=np.float64)
        cx = (wts * xx).sum() / max(s, 1e-12)
        cy = (wts * yy).sum() / max(s, 1e-12)
        ms = float(m.sum())
        if not (s > 0) or not np.isfinite(cx) or not np.isfinite(cy):
            cx = (m * xx).sum() / max(ms, 1e-12)
            cy = (m * yy).sum() / max(ms, 1e-12)
        if 0 <= cx < gr.shape[1] and 0 <= cy < gr.shape[0]:
            out[k] = x0 + cx, y0 + cy
    return out, np.isfinite(out).all(axis=1)
