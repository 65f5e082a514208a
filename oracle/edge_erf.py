"""FP64 CPU oracle of the edge-erf centre refinement.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, step for step, in float64 NumPy:
* ``smooth_binom5``: reference subpx/_gpu/centers.py:518-523 (edge-replicated [1 4 6 4 1]/16);
* ``edge_gradmoment``: :526-544 (argmax of |diff|, window +-loc_rad clipped, |d|^8 moment at
  half-sample positions; ok when the weight sum > 1e-12);
* ``edge_polarity``: :648-662 (sum of the slope over +-loc_rad around the sample nearest
  the seed);
* ``edge_erf``: :665-779 (a + b (x - seed) + c (1 + pol erf((x - x0) / (sqrt2 s))) / 2,
  s = clip(exp(lsig), s_min, s_max); weights 1 within fit_rad plus 0.15 within fit_rad + 1;
  Marquardt damping damp * diag; LU solve; clamps; cond gate on the last damped normal
  matrix; rms gate);
* ``refine_edge_moment``: :547-604 (the same profiles and seeds, centre = midpoints of the
  grad-moment edges; ok needs the four seeds' weight sums > 1e-12, no ordering test);
* ``refine_edge_erf``: :782-879 (two passes of left/right profiles along x from a 9-row
  band mean and top/bottom along y, two binomial passes, grad-moment seeds, erf fits,
  centre = midpoints; ok needs all eight sub-fits and xR > xL, yB > yT).
The reference computes profiles and the grad-moment seed in float32 on the GPU; this
restatement and the HIP kernel use float64 throughout.

Pinned by the reference's own known answer (tests/test_gpu_edge_erf.py:24-33: the profile
15 + 70 (1 + erf((k - 10.37) / (sqrt2 1.3))) / 2, k = 0..23, at six page offsets, seed
offset + 10 -> edge 10.37 +- 1e-6) and by its diagnostic cases (tests/diag_erf_unit.py:
rising / falling, sigma 0.4..5, seeds off by up to 1.5 px) in tests/test_oracle_edge_erf.py.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

DEFAULTS = dict(band_rad=4, edge_rad=10, smooth_passes=2, loc_rad=3, grad_power=8.0, iters=2,
                erf_fit_rad=5, erf_gn_iters=5, erf_init_sigma=1.25, erf_sigma_min=0.35, erf_sigma_max=6.0,
                erf_max_shift=2.5, erf_damp=1e-4, erf_cond_max=1e8)


def smooth_binom5(x):
    x = np.asarray(x, dtype=np.float64)
    xm2 = np.concatenate([x[..., :1], x[..., :1], x[..., :-2]], axis=-1)
    xm1 = np.concatenate([x[..., :1], x[..., :-1]], axis=-1)
    xp1 = np.concatenate([x[..., 1:], x[..., -1:]], axis=-1)
    xp2 = np.concatenate([x[..., 2:], x[..., -1:], x[..., -1:]], axis=-1)
    return (xm2 + 4 * xm1 + 6 * x + 4 * xp1 + xp2) * (1.0 / 16.0)


def edge_gradmoment(prof, base_x, grad_power=8.0, loc_rad=3):
    d = np.abs(prof[:, 1:] - prof[:, :-1])
    B, Lm1 = d.shape
    ip = np.argmax(d, axis=1)
    ip0 = np.clip(ip, loc_rad, max(loc_rad, Lm1 - 1 - loc_rad))
    jj = np.clip(ip0[:, None] + np.arange(-loc_rad, loc_rad + 1)[None, :], 0, Lm1 - 1)
    dloc = d[np.arange(B)[:, None], jj]
    w = dloc ** float(grad_power)
    sw = w.sum(axis=1) + 1e-20
    xk = base_x[:, None] + (jj + 0.5)
    return (w * xk).sum(axis=1) / sw, sw > 1e-12


def edge_polarity(prof, base_x, x_seed, loc_rad=2):
    d = prof[:, 1:] - prof[:, :-1]
    B, Lm1 = d.shape
    xk = base_x[:, None] + (np.arange(Lm1)[None, :] + 0.5)
    ip = np.clip(np.argmin(np.abs(xk - x_seed[:, None]), axis=1), 0, max(0, Lm1 - 1))
    jj = np.clip(ip[:, None] + np.arange(-loc_rad, loc_rad + 1)[None, :], 0, max(0, Lm1 - 1))
    sgn = d[np.arange(B)[:, None], jj].sum(axis=1)
    return np.where(sgn >= 0, 1.0, -1.0)


def edge_erf(prof, base_x, x_seed, *, fit_rad=5, gn_iters=5, init_sigma=1.25, sigma_min=0.35,
             sigma_max=6.0, max_shift=2.5, damp=1e-4, cond_max=1e8, loc_rad=2):
    prof = np.asarray(prof, dtype=np.float64)
    base_x = np.asarray(base_x, dtype=np.float64)
    x_seed = np.asarray(x_seed, dtype=np.float64)
    B, L = prof.shape
    if L < 5:
        return x_seed.copy(), np.zeros(B, dtype=bool)
    xs = base_x[:, None] + np.arange(L)[None, :]
    xr = xs - x_seed[:, None]
    pol = edge_polarity(prof, base_x, x_seed, loc_rad=loc_rad)
    ymin, ymax = prof.min(axis=1), prof.max(axis=1)
    amp_floor = np.maximum((ymax - ymin) * 1e-6, 1e-9)
    a = ymin.copy()
    b = np.zeros(B)
    c = np.maximum(ymax - ymin, amp_floor)
    x0 = x_seed.copy()
    lsig = np.full(B, np.log(max(init_sigma, 1e-6)))
    frf = float(max(2, int(fit_rad)))
    msf = float(max(0.25, max_shift))
    ok = np.ones(B, dtype=bool)
    JTJ = np.zeros((B, 5, 5))
    for _ in range(gn_iters):
        sig = np.clip(np.exp(lsig), sigma_min, sigma_max)
        z = (xs - x0[:, None]) / (np.sqrt(2.0) * sig[:, None])
        s = 0.5 * (1.0 + pol[:, None] * erf(z))
        model = a[:, None] + b[:, None] * xr + c[:, None] * s
        w = (np.abs(xs - x0[:, None]) <= frf).astype(np.float64)
        if L >= 7:
            w += 0.15 * (np.abs(xs - x0[:, None]) <= frf + 1.0)
        sw = np.sqrt(w)
        r = (prof - model) * sw
        ez = np.exp(-(z * z))
        ds_dx0 = -pol[:, None] * ez / (np.sqrt(2 * np.pi) * sig[:, None])
        ds_dlsig = -pol[:, None] * z * ez / np.sqrt(np.pi)
        J = np.stack([sw, sw * xr, sw * s, sw * c[:, None] * ds_dx0, sw * c[:, None] * ds_dlsig], axis=2)
        JTJ = np.einsum("bik,bij->bkj", J, J)
        JTr = np.einsum("bik,bi->bk", J, r)
        diag = np.maximum(np.diagonal(JTJ, axis1=1, axis2=2), 1e-9)
        JTJ = JTJ + damp * diag[:, :, None] * np.eye(5)[None]
        delta = np.zeros((B, 5))
        for k in range(B):
            try:
                delta[k] = np.linalg.solve(JTJ[k], JTr[k])
            except np.linalg.LinAlgError:
                delta[k] = np.nan
        bad = ~np.all(np.isfinite(delta), axis=1)
        delta[bad] = 0.0
        ok &= ~bad
        delta[:, 3] = np.clip(delta[:, 3], -0.75, 0.75)
        delta[:, 4] = np.clip(delta[:, 4], -0.35, 0.35)
        a = a + delta[:, 0]
        b = b + delta[:, 1]
        c = np.maximum(c + delta[:, 2], amp_floor)
        x0 = np.clip(x0 + delta[:, 3], x_seed - msf, x_seed + msf)
        lsig = np.clip(lsig + delta[:, 4], np.log(sigma_min), np.log(sigma_max))
    cond = np.linalg.cond(JTJ)
    ok &= np.isfinite(cond) & (cond < cond_max)
    sig = np.clip(np.exp(lsig), sigma_min, sigma_max)
    z = (xs - x0[:, None]) / (np.sqrt(2.0) * sig[:, None])
    s = 0.5 * (1.0 + pol[:, None] * erf(z))
    model = a[:, None] + b[:, None] * xr + c[:, None] * s
    w = (np.abs(xs - x0[:, None]) <= frf).astype(np.float64)
    n_fit = np.maximum(w.sum(axis=1), 1.0)
    rms = np.sqrt((w * (prof - model) ** 2).sum(axis=1) / n_fit)
    dyn = np.maximum(ymax - ymin, 1e-9)
    ok &= np.isfinite(x0) & np.isfinite(sig) & np.isfinite(rms)
    ok &= np.abs(x0 - x_seed) <= msf + 1e-6
    ok &= (sig >= sigma_min) & (sig <= sigma_max)
    ok &= c > amp_floor
    ok &= rms <= 0.35 * dyn + 1e-6
    return np.where(ok, x0, np.nan), ok


def refine_edge_moment(gray, stats6, **kw):
    """(centers (K, 2), ok (K,)) of the gradient-moment edges for rows [x, y, w, h, cx, cy]."""
    return refine_edge_erf(gray, stats6, _erf=False, **kw)


def refine_edge_erf(gray, stats6, _erf=True, **kw):
    """(centers (K, 2), ok (K,)) for rows [x, y, w, h, cx, cy]."""
    o = dict(DEFAULTS)
    o.update(kw)
    g = np.asarray(gray, dtype=np.float64)
    st = np.asarray(stats6, dtype=np.float64)
    H, W = g.shape
    K = st.shape[0]
    w = st[:, 2].astype(np.float32).astype(np.float64)
    h = st[:, 3].astype(np.float32).astype(np.float64)
    xc = st[:, 4].astype(np.float32).astype(np.float64)
    yc = st[:, 5].astype(np.float32).astype(np.float64)
    br, er = int(o["band_rad"]), int(o["edge_rad"])
    dy = np.arange(-br, br + 1)
    dx = np.arange(-er, er + 1)
    ok_all = np.ones(K, dtype=bool)
    erf_kw = dict(fit_rad=o["erf_fit_rad"], gn_iters=o["erf_gn_iters"], init_sigma=o["erf_init_sigma"],
                  sigma_min=o["erf_sigma_min"], sigma_max=o["erf_sigma_max"], max_shift=o["erf_max_shift"],
                  damp=o["erf_damp"], cond_max=o["erf_cond_max"], loc_rad=max(1, int(o["loc_rad"]) - 1))

    def toint(v):
        v = np.rint(v)
        return np.where(np.isfinite(v), v, 0).astype(np.int64)

    for _ in range(int(o["iters"])):
        xL0 = toint(xc - 0.5 * (w - 1.0))
        xR0 = toint(xc + 0.5 * (w - 1.0))
        yT0 = toint(yc - 0.5 * (h - 1.0))
        yB0 = toint(yc + 0.5 * (h - 1.0))
        ys = np.clip(toint(yc)[:, None] + dy[None, :], 0, H - 1)
        xs_band = np.clip(toint(xc)[:, None] + dy[None, :], 0, W - 1)

        def prof_x(x0i):
            xs = np.clip(x0i[:, None] + dx[None, :], 0, W - 1)
            prof = g[ys[:, :, None], xs[:, None, :]].mean(axis=1)
            for _ in range(int(o["smooth_passes"])):
                prof = smooth_binom5(prof)
            return prof, x0i.astype(np.float64) - er

        def prof_y(y0i):
            yy = np.clip(y0i[:, None] + dx[None, :], 0, H - 1)
            prof = g[yy[:, None, :], xs_band[:, :, None]].mean(axis=1)
            for _ in range(int(o["smooth_passes"])):
                prof = smooth_binom5(prof)
            return prof, y0i.astype(np.float64) - er

        res = []
        oks = np.ones(K, dtype=bool)
        for prof, base in (prof_x(xL0), prof_x(xR0), prof_y(yT0), prof_y(yB0)):
            seed, ok0 = edge_gradmoment(prof, base, grad_power=o["grad_power"], loc_rad=int(o["loc_rad"]))
            if _erf:
                e, ok1 = edge_erf(prof, base, seed, **erf_kw)
                res.append(e)
                oks &= ok0 & ok1
            else:
                res.append(seed)
                oks &= ok0
        xL, xR, yT, yB = res
        if _erf:
            with np.errstate(invalid="ignore"):
                oks &= (xR > xL) & (yB > yT)
        ok_all &= oks
        xc = 0.5 * (xL + xR)
        yc = 0.5 * (yT + yB)
    return np.stack([xc, yc], axis=1), ok_all
