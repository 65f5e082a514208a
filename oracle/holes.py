"""Oracle (test infrastructure only): FP64 NumPy restatement of the outline hole stage.

The hole stage fits every hole's centre, brightness, level and (optionally) edge offset against
the complete page model, one colour class at a time (block Gauss-Seidel), each hole by one
Gauss-Newton step per sweep:

    page model  P(p) = sum_i a_i [T0 + delta_i T1 + delta_i^2 T2](p - c_i) + far light
    hole i      data(p) = P(p) + level_i   on its (2R+1)^2 window

Follows the GPU hole pass of the reference, subpx/_gpu/outline_onepass.py:
  Keys cubic sample .................. :48-79 (and subpx/_outline/page.py:42-74)
  hole_normal ........................ :252-307
  hole_update ........................ :309-363
  render_scatter ..................... :365-397
  far_splat / far_field .............. :399-427, 795-806
  CellList / colours ................. :504-519, 808-826
  nearest / spacing / window R ....... :225-250, 535-548
  set_template (near window, reach) .. :730-793
  fit_holes (the sweep driver) ....... :828-887
The template itself comes from an analytic erf-edged disk here (`disk_templates`): the hole
stage reads any template given as grids, and the reference's band-limited outline templates
(subpx/_outline/onepass.py Spec / Band) are a different subsystem.

Everything is FP64 (the reference samples FP32 templates); the product kernels are compared
against this at a tolerance stated in the tests.
"""
from __future__ import annotations

import math

import numpy as np


# ------------------------------------------------------------------------------------------ templates
def template_box(R):
    """Template box in px for a window half width R (outline_onepass.py:734)."""
    return max(32, int(8 * math.ceil(2 * (R + 6) / 8)))


def _erf_disk_profile(rho, radius, sigma):
    """T0, dT/d delta and (1/2) d^2T/d delta^2 of 1/2 erfc((rho - r - delta) / (sqrt2 sigma)) at delta = 0."""
    from scipy.special import erfc

    u = (rho - radius) / (math.sqrt(2.0) * sigma)
    t0 = 0.5 * erfc(u)
    g = np.exp(-u * u) / (math.sqrt(2.0 * math.pi) * sigma)
    return t0, g, g * u / (math.sqrt(2.0) * sigma)


def near_reach(radius, sigma, R, U=8, far_ramp=2.0):
    """(N, r0, Rr, r_light): the template grid size, the near part's radius, render half width and
    the light's reach, by set_template's rules (outline_onepass.py:745-757) on the disk's centre row."""
    from scipy.special import erfc

    N = template_box(R) * U
    xg = (np.arange(N) - N // 2) / U
    prof = np.abs(0.5 * erfc((np.abs(xg) - radius) / (math.sqrt(2.0) * sigma)))
    above = np.nonzero(prof > 1e-3 * prof.max())[0]
    reach_px = float(above[-1] - above[0] + 1) / (2 * U)
    Rr = min(int(math.ceil(reach_px)) + 1, N // (2 * U) - 2, R + 3)
    r0 = float(Rr)
    Rr = min(int(math.ceil(r0 + far_ramp)) + 1, N // (2 * U) - 2)
    return N, r0, Rr, r0 + far_ramp


def disk_templates(radius, sigma, R, U=8, far_ramp=2.0, far_box=128):
    """Near templates of an erf-edged disk on the fine grid (U bins per px, centre at N/2):
    g5 (N, N, 5) = [T0, T1, Tx, Ty, T2] and g3 (N, N, 3) = [T0, T1, T2], each times the near
    window w(rho) = (1 + cos(pi clip((rho - r0) / ramp, 0, 1))) / 2; the far kernel
    (far_box, far_box) = T0 (1 - w) at whole-pixel offsets. Tx, Ty = d(T0 w)/dx, d(T0 w)/dy.
    Returns dict(g5, g3, far, N, U, r0, Rr, r_light)."""
    N, r0, Rr, r_light = near_reach(radius, sigma, R, U, far_ramp)
    xg = (np.arange(N) - N // 2) / U
    X, Y = np.meshgrid(xg, xg)
    rho = np.hypot(X, Y)
    t0, t1, t2 = _erf_disk_profile(rho, radius, sigma)
    s = np.clip((rho - r0) / far_ramp, 0.0, 1.0)
    w = 0.5 * (1.0 + np.cos(np.pi * s))
    inside = (rho > r0) & (rho < r0 + far_ramp)
    dw = np.where(inside, -0.5 * np.pi * np.sin(np.pi * s) / far_ramp, 0.0)
    drho = -t1 * w + t0 * dw  # d(T0 w)/d rho (dT0/drho = -T1)
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(rho > 0, X / rho, 0.0)
        uy = np.where(rho > 0, Y / rho, 0.0)
    g5 = np.stack([t0 * w, t1 * w, drho * ux, drho * uy, t2 * w], axis=-1)
    g3 = g5[..., [0, 1, 4]]
    xb = np.arange(far_box) - far_box // 2
    XB, YB = np.meshgrid(xb, xb)
    rb = np.hypot(XB, YB)
    fb0, _, _ = _erf_disk_profile(rb, radius, sigma)
    wf = 1.0 - 0.5 * (1.0 + np.cos(np.pi * np.clip((rb - r0) / far_ramp, 0.0, 1.0)))
    return dict(g5=np.ascontiguousarray(g5), g3=np.ascontiguousarray(g3), far=fb0 * wf, N=N, U=U, r0=r0,
                Rr=Rr, r_light=r_light)


# ------------------------------------------------------------------------------------- cubic sampling
def keys_weights(t):
    """Keys (a = -0.5) weights of the taps at -1, 0, 1, 2 for fractions t in [0, 1) (page.py:42-47)."""
    x = np.stack((1 + t, t, 1 - t, 2 - t), axis=-1)
    return np.where(x <= 1, (1.5 * x - 2.5) * x * x + 1, ((-0.5 * x + 2.5) * x - 4) * x + 2)


def sample(grids, sx, sy):
    """Values of the interleaved fine grids (N, N, ng) at fine coordinates (sx, sy) (any equal
    shapes); taps outside the grid read 0 (outline_onepass.py:57-79). Returns (..., ng)."""
    N = grids.shape[0]
    fx, fy = np.floor(sx), np.floor(sy)
    wx, wy = keys_weights(sx - fx), keys_weights(sy - fy)
    ix, iy = fx.astype(np.int64), fy.astype(np.int64)
    out = np.zeros(sx.shape + (grids.shape[2],))
    for k in range(4):
        yy = iy + k - 1
        oky = (yy >= 0) & (yy < N)
        yc = np.clip(yy, 0, N - 1)
        row = np.zeros_like(out)
        for l in range(4):
            xx = ix + l - 1
            ok = oky & (xx >= 0) & (xx < N)
            v = grids[yc, np.clip(xx, 0, N - 1)]
            row += np.where(ok[..., None], wx[..., l, None] * v, 0.0)
        out += wy[..., k, None] * row
    return out


# ------------------------------------------------------------------------------------ spacing, colours
def nearest_distance(c):
    """Distance from every hole to its nearest other hole (outline_onepass.py:225-250), by brute force."""
    d = np.hypot(c[:, None, 0] - c[None, :, 0], c[:, None, 1] - c[None, :, 1])
    np.fill_diagonal(d, np.inf)
    return d.min(axis=1)


def window_half(spacing):
    """R = max(3, ceil(0.75 spacing)) (outline_onepass.py:548)."""
    return max(3, int(math.ceil(0.75 * spacing)))


def colours(c, h, w, spacing):
    """Colour classes: cells of 1.5 spacings; colour = (cell x parity) + 2 (cell y parity) +
    4 (rank of the hole inside its cell, by index). Returns the list of index arrays of the
    non-empty colours in colour order (outline_onepass.py:504-519, 808-826)."""
    cs = 1.5 * spacing
    ncx = int(math.ceil(w / cs)) + 1
    ncy = int(math.ceil(h / cs)) + 1
    cx = np.clip(np.floor(c[:, 0] / cs), 0, ncx - 1).astype(np.int64)
    cy = np.clip(np.floor(c[:, 1] / cs), 0, ncy - 1).astype(np.int64)
    cell = cy * ncx + cx
    order = np.argsort(cell, kind="stable")
    sc = cell[order]
    start = np.searchsorted(sc, np.arange(ncx * ncy + 1))
    rank = np.empty(len(c), dtype=np.int64)
    rank[order] = np.arange(len(c)) - start[sc]
    col = (cx % 2) + 2 * (cy % 2) + 4 * rank
    return [np.nonzero(col == k)[0] for k in range(int(col.max()) + 1) if np.any(col == k)]


# ------------------------------------------------------------------------------------------ kernels
def hole_normal(image, page, g5, idx, c, a, level, delta, U, R, nv0, nv1, rscale, size):
    """(n, 22): 15 upper-triangle sums of J^T W J, 5 of J^T W r, sum r^2, pixel count over each
    hole's window against the page model (outline_onepass.py:252-307)."""
    H, W = image.shape
    N = g5.shape[0]
    n = len(idx)
    off = np.arange(-R, R + 1)
    x, y = c[idx, 0], c[idx, 1]
    rx, ry = np.rint(x).astype(np.int64), np.rint(y).astype(np.int64)
    px = rx[:, None, None] + off[None, None, :]
    py = ry[:, None, None] + off[None, :, None]
    px, py = np.broadcast_to(px, (n, 2 * R + 1, 2 * R + 1)), np.broadcast_to(py, (n, 2 * R + 1, 2 * R + 1))
    valid = (px >= 0) & (px < W) & (py >= 0) & (py < H)
    sx = (px - x[:, None, None]) * U + N // 2
    sy = (py - y[:, None, None]) * U + N // 2
    v = sample(g5, sx, sy)
    f = np.clip(py, 0, H - 1) * W + np.clip(px, 0, W - 1)
    model = page.ravel()[f] + level[idx, None, None]
    resid = image.ravel()[f].astype(np.float64) - model
    wgt = 1.0 / (nv0 + nv1 * np.maximum(model, 0.0))
    aa = a[idx, None, None]
    de = delta[idx, None, None]
    out = np.zeros((n, 22))
    out[:, 20] = np.where(valid, resid * resid, 0.0).sum(axis=(1, 2))
    out[:, 21] = valid.sum(axis=(1, 2))
    if rscale > 0:
        q = resid / rscale
        wgt = wgt / (1.0 + q * q)
    J = [-aa * v[..., 2], -aa * v[..., 3], v[..., 0] + de * (v[..., 1] + de * v[..., 4]), np.ones_like(resid),
         aa * (v[..., 1] + 2 * de * v[..., 4]) if size else np.zeros_like(resid)]
    q = 0
    for r in range(5):
        for s in range(r, 5):
            out[:, q] = np.where(valid, wgt * J[r] * J[s], 0.0).sum(axis=(1, 2))
            q += 1
        out[:, 15 + r] = np.where(valid, J[r] * wgt * resid, 0.0).sum(axis=(1, 2))
    return out


def hole_update(acc, idx, c, a, level, delta, size, max_move, max_delta, a_med, relax):
    """Solve each hole's 5 x 5 (4 x 4 without size) system with a 1e-9 trace ridge, bound and
    apply the step (outline_onepass.py:309-363). Updates c, a, level, delta in place; returns
    (stat (n, 4) = [centre move, edge-offset move, leftover rms, edge-offset standard error],
    c_old (n, 2))."""
    n = len(idx)
    k = 5 if size else 4
    stat = np.zeros((n, 4))
    c_old = c[idx].copy()
    for i, h in enumerate(idx):
        o = acc[i]
        A = np.zeros((5, 5))
        q = 0
        for r in range(5):
            for s in range(r, 5):
                A[r, s] = A[s, r] = o[q]
                q += 1
        A = A[:k, :k]
        tr = np.trace(A)
        A = A + np.eye(k) * (1e-9 * tr / k)
        step = np.zeros(5)
        ok = True
        try:
            Ainv = np.linalg.inv(A)
            if not np.all(np.isfinite(Ainv)):
                ok = False
        except np.linalg.LinAlgError:
            ok = False
        if ok:
            step[:k] = Ainv @ o[15:15 + k]
            step[~np.isfinite(step) | (np.abs(step) > 1e30)] = 0.0
        step[2:5] *= relax
        stat[i, 2] = math.sqrt(o[20] / (o[21] if o[21] > 0 else 1.0))
        stat[i, 3] = math.sqrt(max(Ainv[4, 4], 0.0)) if (ok and size) else np.nan
        mx = min(max(step[0], -max_move), max_move)
        my = min(max(step[1], -max_move), max_move)
        c[h, 0] += mx
        c[h, 1] += my
        stat[i, 0] = max(abs(mx), abs(my))
        a0 = a[h]
        an = min(max(a0 + step[2], 0.5 * a0), 2.0 * a0)
        a[h] = min(max(an, 1e-3), 5.0 * a_med)
        level[h] += step[3]
        if size:
            dn = min(max(delta[h] + step[4], -max_delta), max_delta)
            stat[i, 1] = abs(dn - delta[h])
            delta[h] = dn
    return stat, c_old


def render_scatter(page, cache, g3, idx, c_old, c, a, delta, U, Rr, rl2):
    """Take each hole's cached light (FP32 values at offsets around rint(c_old)) out of the page
    model and put its new light in (outline_onepass.py:365-397). page (H, W) and cache
    (n, (2Rr+1)^2) float32 are updated in place."""
    H, W = page.shape
    N = g3.shape[0]
    off = np.arange(-Rr, Rr + 1)
    oy, ox = np.meshgrid(off, off, indexing="ij")
    oy, ox = oy.ravel(), ox.ravel()
    near = (ox * ox + oy * oy).astype(np.float64) <= rl2
    flat = page.ravel()
    for i, h in enumerate(idx):
        old = cache[i]
        m = near & (old != 0.0)
        pxo = int(np.rint(c_old[i, 0])) + ox[m]
        pyo = int(np.rint(c_old[i, 1])) + oy[m]
        np.add.at(flat, pyo * W + pxo, -old[m].astype(np.float64))
        x, y = c[h]
        px = int(np.rint(x)) + ox
        py = int(np.rint(y)) + oy
        ins = near & (px >= 0) & (px < W) & (py >= 0) & (py < H)
        v = sample(g3, (px[ins] - x) * U + N // 2, (py[ins] - y) * U + N // 2)
        de = delta[h]
        light = np.zeros(len(ox), dtype=np.float32)
        light[ins] = (a[h] * (v[:, 0] + de * (v[:, 1] + de * v[:, 2]))).astype(np.float32)
        np.add.at(flat, py[ins] * W + px[ins], light[ins].astype(np.float64))
        new = np.where(near, light, old)  # offsets beyond the light's reach keep what they held (zero)
        cache[i] = new


# ---------------------------------------------------------------------------------------- far light
def far_geometry(c, h, w, far_box):
    """Padding of the far-light grid: margins (my, mx) and padded size (hp, wp), even sizes with
    small prime factors (outline_onepass.py:453-461, 779-783)."""
    B = far_box
    out_x = max(0.0, -float(c[:, 0].min()), float(c[:, 0].max()) - (w - 1))
    out_y = max(0.0, -float(c[:, 1].min()), float(c[:, 1].max()) - (h - 1))
    my, mx = B // 2 + int(math.ceil(out_y)) + 4, B // 2 + int(math.ceil(out_x)) + 4
    return my, mx, even_fast(h + 2 * my), even_fast(w + 2 * mx)


def even_fast(n):
    """Smallest even m >= n whose only prime factors are 2, 3, 5 and 7."""
    m = n + (n & 1)
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 2


def far_field(c, a, delta, cd0, far, h, w, geom):
    """The holes' far light: brightness x (1 + 4 delta / cd0) splatted with Keys weights on the
    padded pixel grid (FP32), circularly convolved with the far kernel, cropped (FP64)
    (outline_onepass.py:399-427, 795-806)."""
    my, mx, hp, wp = geom
    B = far.shape[0]
    grid = np.zeros((hp, wp), dtype=np.float32)
    amp = a * (1.0 + 4.0 * delta / cd0)
    for i in range(len(c)):
        x, y = c[i, 0] + mx, c[i, 1] + my
        fx, fy = math.floor(x), math.floor(y)
        ix, iy = int(fx), int(fy)
        if ix < 1 or ix + 2 >= wp or iy < 1 or iy + 2 >= hp:
            continue
        wx = keys_weights(np.array(x - fx))
        wy = keys_weights(np.array(y - fy))
        for q in range(4):
            for p in range(4):
                grid[iy + q - 1, ix + p - 1] += np.float32(amp[i] * wy[q] * wx[p])
    pad = np.zeros((hp, wp))
    pad[:B, :B] = far.astype(np.float32)
    out = np.fft.irfft2(np.fft.rfft2(grid.astype(np.float64)) * np.fft.rfft2(pad), s=(hp, wp))
    oy, ox = my + B // 2, mx + B // 2
    return out[oy:oy + h, ox:ox + w]


# ------------------------------------------------------------------------------------------- driver
def fit_holes(image, c, a, level, delta, tmpl, *, R, noise, sweeps=2, size=False, max_move=0.25, robust_c=3.0,
              relax=1.0, far=True, cd0=None, chunks=None):
    """The hole pass (outline_onepass.py:828-887) on copies of the hole state. tmpl: the dict of
    `disk_templates`. Returns dict(c, a, level, delta, stat, log, page)."""
    image = np.asarray(image, dtype=np.float32)
    H, W = image.shape
    c, a, level, delta = (np.array(v, dtype=np.float64) for v in (c, a, level, delta))
    g5, g3, U, Rr = tmpl["g5"], tmpl["g3"], tmpl["U"], tmpl["Rr"]
    n = len(c)
    if chunks is None:
        chunks = colours(c, H, W, float(np.median(nearest_distance(c))))
    P = (2 * Rr + 1) ** 2
    max_delta = 0.3 * 0.5 * cd0 if cd0 is not None else 0.0
    page = np.zeros((H, W))
    rl2 = (tmpl["r_light"] + 0.75) ** 2
    caches = [np.zeros((len(idx), P), dtype=np.float32) for idx in chunks]
    c_olds = []
    for idx, cache in zip(chunks, caches):
        c_old = c[idx].copy()
        render_scatter(page, cache, g3, idx, c_old, c, a, delta, U, Rr, rl2)
        c_olds.append(c_old)
    geom = far_geometry(c, H, W, tmpl["far"].shape[0]) if far else None
    far_l = far_field(c, a, delta, cd0, tmpl["far"], H, W, geom) if far else None
    if far:
        page += far_l
    nv0, nv1 = noise ** 2, 0.0
    rms_typ = noise
    a_med = float(np.median(a))
    stat = np.zeros((n, 4))
    offs = np.cumsum([0] + [len(i) for i in chunks])
    log = []
    for it in range(sweeps):
        for j, idx in enumerate(chunks):
            acc = hole_normal(image, page, g5, idx, c, a, level, delta, U, R, nv0, nv1,
                              robust_c * max(rms_typ, 1e-6), size)
            st, c_olds[j] = hole_update(acc, idx, c, a, level, delta, size, max_move, max_delta, a_med, relax)
            stat[offs[j]:offs[j + 1]] = st
            render_scatter(page, caches[j], g3, idx, c_olds[j], c, a, delta, U, Rr, rl2)
        if far and it + 1 < sweeps:
            new_far = far_field(c, a, delta, cd0, tmpl["far"], H, W, geom)
            page += new_far - far_l
            far_l = new_far
        rms_typ = float(np.median(stat[:, 2]))
        q = np.percentile(stat[:, :2], 99, axis=0)
        log.append(dict(sweep=it + 1, move99=float(q[0]), delta99=float(q[1]), rms=rms_typ))
    return dict(c=c, a=a, level=level, delta=delta, stat=stat, log=log, page=page, chunks=chunks)


def render_page(c, a, level_bg, delta, tmpl, h, w, far=True, cd0=None):
    """A noise-free page of the model: every hole's near light (FP64, no FP32 cache rounding) +
    far light + a constant background (for known-answer tests)."""
    page = np.full((h, w), float(level_bg))
    g3, U, Rr, N = tmpl["g3"], tmpl["U"], tmpl["Rr"], tmpl["N"]
    rl2 = (tmpl["r_light"] + 0.75) ** 2
    off = np.arange(-Rr, Rr + 1)
    oy, ox = np.meshgrid(off, off, indexing="ij")
    oy, ox = oy.ravel(), ox.ravel()
    near = (ox * ox + oy * oy).astype(np.float64) <= rl2
    for i in range(len(c)):
        x, y = c[i]
        px, py = int(np.rint(x)) + ox, int(np.rint(y)) + oy
        ins = near & (px >= 0) & (px < w) & (py >= 0) & (py < h)
        v = sample(g3, (px[ins] - x) * U + N // 2, (py[ins] - y) * U + N // 2)
        np.add.at(page.ravel(), py[ins] * w + px[ins], a[i] * (v[:, 0] + delta[i] * (v[:, 1] + delta[i] * v[:, 2])))
    if far:
        page += far_field(c, a, delta, cd0, tmpl["far"], h, w, far_geometry(c, h, w, tmpl["far"].shape[0]))
    return page
