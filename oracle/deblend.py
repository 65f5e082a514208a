"""CPU oracle of the deblending behind ``detect_centers(background="model")``.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates in NumPy / SciPy:

* ``deblend`` (reference subpx/_cpu/deblend.py:421-522) with its helpers: the image as a smooth
  background map plus one 2-D Gaussian per hole, ROI-local parameters [A, x0, y0, u1, u2, u3],
  precision L L^T with L = [[e^u1, 0], [u2, e^u3]] (:48-71); the first fit on the cell border's
  10th-percentile background (:467-476, 152-171); then, per iteration, the background map from
  the residual's tile medians over free pixels (far from every seed and below 5 % of the median
  amplitude in the model) with the measured / interpolated / inferred rules (:280-393), the
  Gaussian LM of every cell on its neighbour-cleaned pixels (:174-254), the relaxed update, the
  status rules (:262-277) and the 99th-percentile stop (:497-522); truncated renders
  (:74-107) and cleaned cells (:136-149);
* ``_deblended_cells_centers`` / ``_refine_deblended`` (reference subpx/centers.py:225-316): the
  square cell ROIs of half ceil(0.6 pitch), the per-strip bands with overlap, and the final
  refinement of every cleaned cell by the GPU backend's batch methods
  (subpx/_gpu/centers.py:33-111, 931-956: the 2-px border median as background, the weighted
  centroid, the unweighted log-quadratic with its 1e-9 ridge and gates) or by the Voronoi
  estimators on the cell's window near the fitted Gaussian.

Pinned in tests/test_oracle_deblend.py by the reference's own known answers
(tests/test_deblend.py:46-57: noise-free staggered Gaussian lattices at four pitch ratios ->
background within 1 % of the amplitude, amplitudes within 1 %, centres within 0.01 px;
:177-192: detect_centers(background="model", refine="logquad") median centre error below
0.005 px) and by the Gaussian Jacobian against central differences.
"""

from __future__ import annotations

import numpy as np

from . import voronoi_refine as vr
from .segment import bg_from_border

RENDER_SIGMAS = 4.0
FWHM_K = float(np.sqrt(2 * np.log(2)))
SIGMA_MIN = 0.5
STATUS_OK, STATUS_DIVERGED, STATUS_NON_GAUSSIAN, STATUS_OUTSIDE = 0, 1, 2, 3
STATUS_NAMES = ("ok", "fit_diverged", "non_gaussian_fit", "center_outside_cell")


# ----------------------------------------------------------------------------- the model
def precision(p):
    e1, u2, e3 = np.exp(p[:, 3]), p[:, 4], np.exp(p[:, 5])
    return e1 * e1, e1 * u2, u2 * u2 + e3 * e3  # (P00, P01, P11)


def sigma_major(p):
    """1 / sqrt of the precision's smaller eigenvalue (the Gaussian's long axis)."""
    a, b, c = precision(np.asarray(p, dtype=np.float64).reshape(-1, 6))
    lam = np.linalg.eigvalsh(np.stack([np.stack([a, b], -1), np.stack([b, c], -1)], -2))[:, 0]
    return 1.0 / np.sqrt(np.maximum(lam, 1e-300))


def gaussian(p, dx, dy):
    """(model, shape, a, b, e1, u2, e3) of parameters p (N, 6) at offsets dx, dy (N, K)."""
    e1 = np.exp(p[:, 3])[:, None]
    u2 = p[:, 4][:, None]
    e3 = np.exp(p[:, 5])[:, None]
    a = e1 * dx + u2 * dy
    b = e3 * dy
    shape = np.exp(-0.5 * (a * a + b * b))
    return p[:, 0][:, None] * shape, shape, a, b, e1, u2, e3


def jacobian(p, dx, dy):
    """d model / d p, (N, K, 6)."""
    model, shape, a, b, e1, u2, e3 = gaussian(p, dx, dy)
    half = -0.5 * model
    return np.stack((shape, half * (-2 * a * e1), half * (-2 * a * u2 - 2 * b * e3), half * (2 * a * e1 * dx),
                     half * (2 * a * dy), half * (2 * b * e3 * dy)), axis=2)


def render_radius(p, cap):
    return np.clip(np.ceil(RENDER_SIGMAS * sigma_major(p)), 2, cap)


def render(pg, shape, radius_cap):
    """Sum of the truncated Gaussians (global centres; |dx|, |dy| <= radius) on an (H, W) image."""
    H, W = shape
    out = np.zeros(H * W)
    if len(pg) == 0:
        return out.reshape(shape)
    rad = render_radius(pg, radius_cap)
    for k in range(len(pg)):
        r = int(rad[k])
        ix = np.floor(pg[k, 1]).astype(np.int64) + np.arange(-r, r + 1)
        iy = np.floor(pg[k, 2]).astype(np.int64) + np.arange(-r, r + 1)
        X, Y = np.meshgrid(ix, iy)
        dx, dy = X.ravel() - pg[k, 1], Y.ravel() - pg[k, 2]
        v = gaussian(pg[k:k + 1], dx[None], dy[None])[0][0]
        ok = (np.abs(dx) <= rad[k]) & (np.abs(dy) <= rad[k]) & (X.ravel() >= 0) & (X.ravel() < W) \
            & (Y.ravel() >= 0) & (Y.ravel() < H)
        np.add.at(out, Y.ravel()[ok] * W + X.ravel()[ok], v[ok])
    return out.reshape(shape)


def windows(image, origins, shape):
    H, W = image.shape
    ys = origins[:, 1, None] + np.arange(shape[0])[None]
    xs = origins[:, 0, None] + np.arange(shape[1])[None]
    inside = ((ys >= 0) & (ys < H))[:, :, None] & ((xs >= 0) & (xs < W))[:, None, :]
    out = image[np.clip(ys, 0, H - 1)[:, :, None], np.clip(xs, 0, W - 1)[:, None, :]]
    return np.where(inside, out, 0.0)


def clean_cells(raw, background, model, pg, masks, origins, radius_cap):
    N, h, w = masks.shape
    yy, xx = np.indices((h, w), dtype=float)
    gx = (origins[:, 0, None, None] + xx[None]).reshape(N, -1)
    gy = (origins[:, 1, None, None] + yy[None]).reshape(N, -1)
    dx, dy = gx - pg[:, 1, None], gy - pg[:, 2, None]
    rad = render_radius(pg, radius_cap)[:, None]
    own = np.where((np.abs(dx) <= rad) & (np.abs(dy) <= rad), gaussian(pg, dx, dy)[0], 0.0).reshape(masks.shape)
    others = windows(model, origins, (h, w)) - own
    return np.where(masks, raw - windows(background, origins, (h, w)) - others, 0.0)


# ---------------------------------------------------------------------------------- fits
def sigma_caps(masks):
    return np.maximum(1.0, np.sqrt(masks.reshape(len(masks), -1).sum(axis=1) / np.pi))


def initial_params(data, masks, seeds_local):
    n = len(data)
    p = np.zeros((n, 6))
    if n == 0:
        return p
    peak = np.where(masks, data, -np.inf).reshape(n, -1).max(axis=1)
    core = masks & (data > 0.2 * peak[:, None, None])
    wgt = np.where(core, data, 0.0)
    yy, xx = np.indices(data.shape[1:], dtype=float)
    tot = np.maximum(wgt.sum(axis=(1, 2)), 1e-300)
    dx = xx[None] - seeds_local[:, 0, None, None]
    dy = yy[None] - seeds_local[:, 1, None, None]
    var = ((wgt * (dx * dx + dy * dy)).sum(axis=(1, 2)) / tot) / 2
    sig = np.sqrt(np.clip(var, 0.25, None))
    p[:, 0] = np.maximum(peak, 1e-12)
    p[:, 1:3] = seeds_local
    p[:, 3] = p[:, 5] = -np.log(sig)
    return p


def fit_gaussians(data, masks, init, max_iter=30):
    """One Gaussian per cell by LM, all cells in lockstep (each cell's arithmetic is its own):
    damping lam diag + 1e-12, amplitude clipped to [0, 3 max |cell data|], log-widths to
    [-log cap, -log 0.5], a trial taken when its cost does not rise (lam / 3, else lam x 4), a
    cell done after a relative decrease <= 1e-12, a step below 1e-7 or lam > 1e12."""
    n, h, w = data.shape
    params = np.asarray(init, dtype=np.float64).copy()
    if n == 0:
        return params
    yy, xx = np.indices((h, w), dtype=float)
    xx, yy = xx.ravel(), yy.ravel()
    caps = sigma_caps(masks)
    d = data.reshape(n, -1)
    wt = masks.reshape(n, -1).astype(float)
    lam = np.full(n, 1e-3)
    amp_cap = 3 * np.maximum(np.abs(np.where(wt > 0, d, 0)).max(axis=1), 1e-12)

    def cost_of(q, rows):
        r = d[rows] - gaussian(q, xx[None] - q[:, 1, None], yy[None] - q[:, 2, None])[0]
        return (wt[rows] * r * r).sum(axis=1)

    cost = cost_of(params, slice(None))
    active = np.isfinite(cost)
    for _ in range(max_iter):
        if not active.any():
            break
        idx = np.flatnonzero(active)
        q = params[idx]
        dx, dy = xx[None] - q[:, 1, None], yy[None] - q[:, 2, None]
        J = jacobian(q, dx, dy)
        r = d[idx] - gaussian(q, dx, dy)[0]
        Jw = J * wt[idx][:, :, None]
        A = np.einsum("nki,nkj->nij", Jw, J)
        g = np.einsum("nki,nk->ni", Jw, r)
        diag = np.einsum("nii->ni", A)
        damped = A + (lam[idx][:, None] * diag + 1e-12)[:, :, None] * np.eye(6)
        step = np.zeros_like(g)
        for k in range(len(idx)):
            try:
                step[k] = np.linalg.solve(damped[k], g[k])
            except np.linalg.LinAlgError:
                step[k] = np.linalg.lstsq(damped[k], g[k], rcond=None)[0]
        trial = q + step
        low = -np.log(caps[idx])
        trial[:, 0] = np.clip(trial[:, 0], 0.0, amp_cap[idx])
        trial[:, 3] = np.clip(trial[:, 3], low, -np.log(SIGMA_MIN))
        trial[:, 5] = np.clip(trial[:, 5], low, -np.log(SIGMA_MIN))
        new = cost_of(trial, idx)
        better = np.isfinite(new) & (new <= cost[idx])
        params[idx[better]] = trial[better]
        lam[idx[better]] /= 3
        lam[idx[~better]] *= 4
        small = np.abs(step).max(axis=1) < 1e-7
        done = (better & (cost[idx] - new <= 1e-12 * np.maximum(cost[idx], 1e-300))) | small
        cost[idx[better]] = new[better]
        active[idx[done | (lam[idx] > 1e12)]] = False
    return params


def fit_status(p, masks, caps):
    n, h, w = masks.shape
    st = np.full(n, STATUS_OK, dtype=np.int32)
    bad = ~np.isfinite(p).all(axis=1)
    sig = np.full(n, np.nan)
    if (~bad).any():
        sig[~bad] = sigma_major(p[~bad])
    st[bad] = STATUS_DIVERGED
    st[~bad & ((p[:, 0] <= 0) | (sig >= 0.999 * caps))] = STATUS_NON_GAUSSIAN
    c = np.where(np.isfinite(p[:, 1:3]), p[:, 1:3], 0.0)
    ci = np.round(c[:, 1]).astype(int).clip(0, h - 1), np.round(c[:, 0]).astype(int).clip(0, w - 1)
    inside = (p[:, 1] >= -0.5) & (p[:, 1] <= w - 0.5) & (p[:, 2] >= -0.5) & (p[:, 2] <= h - 0.5)
    inside &= masks[np.arange(n), ci[0], ci[1]]
    st[(st == STATUS_OK) & ~inside] = STATUS_OUTSIDE
    return st


# ----------------------------------------------------------------------------- background
def tile_centers(tile, shape):
    H, W = shape
    yc = np.array([(j * tile + min((j + 1) * tile, H) - 1) / 2 for j in range(-(-H // tile))])
    xc = np.array([(i * tile + min((i + 1) * tile, W) - 1) / 2 for i in range(-(-W // tile))])
    return yc, xc


def tile_medians(values, usable, tile):
    H, W = values.shape
    ny, nx = -(-H // tile), -(-W // tile)
    med = np.full((ny, nx), np.nan)
    cnt = np.zeros((ny, nx), dtype=np.int64)
    for j in range(ny):
        for i in range(nx):
            v = values[j * tile:(j + 1) * tile, i * tile:(i + 1) * tile][usable[j * tile:(j + 1) * tile,
                                                                                  i * tile:(i + 1) * tile]]
            cnt[j, i] = v.size
            if v.size:
                med[j, i] = np.median(v)
    return med, cnt


def fill_tiles(free_med, free_count, usable_med, tile, shape):
    """(medians or None, source): measured where enough free pixels, the others linearly
    interpolated over a Delaunay triangulation of the measured tile centres (nearest outside
    it); without any measured tile, the medians of all usable pixels."""
    measured = free_count >= max(16, int(0.1 * tile * tile))
    if measured.all():
        return free_med, "measured"
    if measured.any():
        from scipy.interpolate import LinearNDInterpolator
        from scipy.spatial import Delaunay, cKDTree

        yc, xc = tile_centers(tile, shape)
        grid = np.stack(np.meshgrid(yc, xc, indexing="ij"), axis=-1)
        known, vals = grid[measured], free_med[measured]
        try:
            tri = Delaunay(known) if len(vals) >= 3 else None
        except Exception:
            tri = None
        med = free_med.copy()
        miss = ~measured
        med[miss] = np.nan
        if tri is not None:
            med[miss] = LinearNDInterpolator(tri, vals)(grid[miss])
        still = miss & np.isnan(med)
        med[still] = vals[cKDTree(known).query(grid[still])[1]]
        return med, "interpolated"
    med = usable_med
    if np.isnan(med).all():
        return None, "inferred"
    return np.where(np.isnan(med), np.nanmedian(med), med), "inferred"


def upsample_tiles(med, tile, shape):
    H, W = shape
    ny = med.shape[0]
    yc, xc = tile_centers(tile, shape)
    rows = np.stack([np.interp(np.arange(W), xc, med[j]) for j in range(ny)])
    if ny == 1:
        return np.repeat(rows, H, axis=0)
    y = np.arange(H, dtype=float)
    j1 = np.clip(np.searchsorted(yc, y), 1, ny - 1)
    t = np.clip((y - yc[j1 - 1]) / (yc[j1] - yc[j1 - 1]), 0, 1)[:, None]
    return (1 - t) * rows[j1 - 1] + t * rows[j1]


def background_map(residual, free, tile):
    usable = np.ones(residual.shape, bool)
    fm, fc = tile_medians(residual, free, tile)
    um, _ = tile_medians(residual, usable, tile) if not (fc >= max(16, int(0.1 * tile * tile))).any() else (None, None)
    med, source = fill_tiles(fm, fc, um, tile, residual.shape)
    if med is None:
        return np.full(residual.shape, float(np.median(residual))), source
    return upsample_tiles(med, tile, residual.shape), source


def auto_tile(seeds, shape):
    from scipy.spatial import cKDTree

    if len(seeds) < 2:
        return int(max(shape)), float("nan")
    pitch = float(np.median(cKDTree(seeds).query(seeds, k=2)[0][:, 1]))
    return int(np.clip(round(4 * pitch), 16, max(shape))), pitch


# ------------------------------------------------------------------------------ deblend
def deblend(image, seeds, masks, origins, *, iterations=30, relaxation=0.6, tolerance=1e-3, max_iter=30,
            free_fraction=0.05, free_distance_fwhm=1.25):
    """dict(params, status, cleaned, background, model, iterations, converged, source)."""
    from scipy.ndimage import binary_erosion, distance_transform_edt

    image = np.asarray(image, dtype=np.float64)
    shape = image.shape
    seeds = np.asarray(seeds, dtype=np.float64).reshape(-1, 2)
    n = len(seeds)
    origins = np.asarray(origins, dtype=np.int64).reshape(-1, 2)
    tile = auto_tile(seeds, shape)[0]
    seed_px = np.ones(shape, bool)
    seed_px[np.clip(np.rint(seeds[:, 1]).astype(int), 0, shape[0] - 1),
            np.clip(np.rint(seeds[:, 0]).astype(int), 0, shape[1] - 1)] = False
    distance = distance_transform_edt(seed_px)
    roi_shape = masks.shape[1:]
    raw = windows(image, origins, roi_shape)
    cap = int(max(roi_shape))
    caps = sigma_caps(masks)
    border = masks & ~binary_erosion(masks, structure=np.ones((1, 3, 3)), border_value=0)
    bg0 = np.array([np.percentile(r[b], 10) if b.any() else 0.0 for r, b in zip(raw, border)])
    data = np.where(masks, raw - bg0[:, None, None], 0.0)
    params = fit_gaussians(data, masks, initial_params(data, masks, seeds - origins), max_iter)
    status = fit_status(params, masks, caps)

    def glob(p):
        g = p.copy()
        g[:, 1:3] += origins
        return g

    def background_of(model, p, st):
        good = (st == STATUS_OK) & (p[:, 0] > 0)
        amp = float(np.median(p[good, 0])) if good.any() else 1.0
        fwhm = float(np.median(2 * FWHM_K * sigma_major(p[good]))) if good.any() else 0.0
        free = (model < free_fraction * amp) & (distance >= free_distance_fwhm * fwhm)
        return background_map(image - model, free, tile)

    model = render(glob(params), shape, cap)
    used, converged = 0, False
    for used in range(1, int(iterations) + 1):
        background, _ = background_of(model, params, status)
        fitted = fit_gaussians(clean_cells(raw, background, model, glob(params), masks, origins, cap), masks, params,
                               max_iter)
        step = np.where(np.isfinite(fitted), fitted - params, 0.0)
        params = params + relaxation * step
        status = fit_status(params, masks, caps)
        model = render(glob(params), shape, cap)
        ok = (status == STATUS_OK) & np.isfinite(fitted).all(axis=1)
        if not ok.any():
            break
        if np.percentile(np.abs(step[ok][:, 1:]).max(axis=1), 99) < tolerance:
            converged = True
            break
    background, source = background_of(model, params, status)
    cleaned = clean_cells(raw, background, model, glob(params), masks, origins, cap)
    return dict(params=params, status=status, cleaned=cleaned, background=background, model=model,
                iterations=used, converged=converged, source=source, tile=tile)


# ------------------------------------------------------------------- refinement of cells
def weighted_batch(rois, masks):
    N, Hm, Wm = rois.shape
    out = np.full((N, 2), np.nan)
    yy, xx = np.indices((Hm, Wm), dtype=float)
    for i in range(N):
        bg = bg_from_border(rois[i], border=2)
        w = np.where(masks[i], np.maximum(rois[i] - bg, 0), 0)
        s = w.sum()
        cx, cy = (w * xx).sum() / max(s, 1e-12), (w * yy).sum() / max(s, 1e-12)
        if s <= 0 or not (np.isfinite(cx) and np.isfinite(cy)):
            ms = max(float(masks[i].sum()), 1e-12)
            cx, cy = (masks[i] * xx).sum() / ms, (masks[i] * yy).sum() / ms
        if 0 <= cx < Wm and 0 <= cy < Hm:
            out[i] = cx, cy
    return out


def logquad_batch(rois, masks):
    N, Hm, Wm = rois.shape
    out = np.full((N, 2), np.nan)
    Y, X = np.indices((Hm, Wm), dtype=float)
    A = np.stack([X * X, Y * Y, X * Y, X, Y, np.ones_like(X)], axis=-1)
    for i in range(N):
        bg = bg_from_border(rois[i], border=2)
        v = rois[i] - bg
        valid = masks[i] & (v > 0)
        z = np.where(valid, np.log(np.maximum(v, 1e-12)), 0.0)
        Av = A * valid[..., None]
        G = np.einsum("hwk,hwl->kl", Av, A) + np.eye(6) * 1e-9
        rhs = np.einsum("hwk,hw->k", Av, z)
        try:
            a, b, c, d, e, _ = np.linalg.solve(G, rhs)
        except np.linalg.LinAlgError:
            continue
        det = 4 * a * b - c * c
        if not abs(det) > 1e-12:
            continue
        cx, cy = (c * e - 2 * b * d) / det, (c * d - 2 * a * e) / det
        negdef = a < -1e-12 and b < -1e-12 and det > 1e-12
        if valid.sum() >= 6 and negdef and 0 <= cx < Wm and 0 <= cy < Hm and np.isfinite(cx) and np.isfinite(cy):
            out[i] = cx, cy
    return out


def deblended_centers(signal, seeds, refine, *, pad=3, iterations=30, upsample_factor=4):
    """(centres (N, 2), quality or None, state) of one band."""
    shape = signal.shape
    pitch = auto_tile(seeds, shape)[1]
    half = int(np.ceil(0.6 * pitch)) if np.isfinite(pitch) else 16
    labels = vr.voronoi_labels(seeds, shape)
    rows = np.array([[x - half, y - half, 2 * half + 1, 2 * half + 1, x, y] for x, y in seeds]).reshape(-1, 6)
    _, masks, origins = vr.extract_rois(signal, labels, rows, pad=pad)
    origins = np.asarray(origins, dtype=np.float64).reshape(-1, 2)
    st = deblend(signal, seeds, masks, origins, iterations=iterations)
    cleaned = st["cleaned"]
    quality = None
    if refine in ("radial_symmetry", "isophote_curvature"):
        p = st["params"]
        reach = np.where(np.isfinite(p).all(axis=1), 2 * FWHM_K * sigma_major(np.nan_to_num(p)), np.inf)
        yy, xx = np.indices(masks.shape[1:])
        loc = seeds - origins
        near = np.hypot(xx[None] - loc[:, 0, None, None], yy[None] - loc[:, 1, None, None]) \
            <= (reach + pad)[:, None, None]
        window = masks & near
        fn = vr.radial_symmetry if refine == "radial_symmetry" else vr.isophote_curvature
        local, quality = fn(np.where(window, cleaned, 0.0), window, upsample_factor=upsample_factor)
    else:
        p = st["params"]
        peak = np.where(np.isfinite(p[:, 0]) & (p[:, 0] > 0), p[:, 0],
                        np.where(masks, cleaned, -np.inf).reshape(len(seeds), -1).max(axis=1))
        support = masks & (cleaned > 0.2 * peak[:, None, None])
        local = logquad_batch(cleaned, support) if refine in ("logquad", "logquadratic") \
            else weighted_batch(cleaned, support)
    return local + origins, quality, st
