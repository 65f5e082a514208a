"""FP64 CPU oracle of the batched erf-disk / erf-ellipse Levenberg-Marquardt fit.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker.

The erf-disk model has no reference function (the task defines it); the batched LM it
follows is the reference's wave-per-cell Gaussian LM (reference
subpx/_gpu/deblend.py:55-152: normal equations :82-107, solve :113-130, accept/reject
:141-149) and `_fit_disks` (reference subpx/_cpu/invariant_cd.py:165-246). This file
states the algorithm the HIP kernel runs, step for step, in float64 with the exact
``scipy.special.erfc``:

  I(x, y) = B + A * erfc(u) / 2,  u = (rho - r) / (sqrt(2) s),
  rho = sqrt(dx^2 + dy^2)            (disk, parameters x0 y0 r s A B)
  rho = sqrt(dx^2 + (dy / q)^2)      (ellipse, 7th parameter q)

Patch: (2h+1)^2 pixels centred on rint(seed), pixels outside the image left out.
Moment start: B = mean of the patch's outer ring; the in-set is the pixels more than
half the peak above B; centre = weighted centroid of the in-set; r = sqrt(n_in / pi);
A = mean of (v - B) over the in-set; s = init_sigma; q = 1.
LM: Marquardt damping M_aa += lambda * max(M_aa, 1e-12 max_a M_aa); LDL^T solve; a
trial is refused (lambda *= up) when it leaves |x0|,|y0| <= h + 0.5,
0.25 < r < 2h + 1, 0.05 < s < h + 0.5 (and 0.25 < q < 4); a step with
max(|dx0|, |dy0|, |dr|, |ds|, r |dq|) < step_tol ends the fit with status "converged" and is
applied without a trial evaluation (the reported rms is that of the point before it);
otherwise the trial is accepted when its cost is at most cost * (1 + accept_rtol)
(lambda *= down, floor 1e-12) and refused otherwise; after max_iter attempts the status is
"max_iter".

Neighbour (Voronoi cell) mask, when neighbour lists are given: a patch pixel takes part
only when it is strictly nearer to the own seed than to each listed neighbour seed, ties
going to the lower feature index (after reference subpx/_cpu/cells.py:8-33), decided on the
FP64 linear form |q - s_j|^2 - |q - s_0|^2 = (a_j qx + b_j qy) + c_j in the kernel's operation
order; the start's background ring then also takes the cell's boundary pixels (in the cell
with a 4-neighbour outside it, reference subpx/_cpu/cells.py:72-80), found as
f_j(q) - max(|a_j|, |b_j|) < 0 (or = 0 towards a lower index), the minimum of f_j over the
4-neighbours.
``seed_neighbours`` lists the nb nearest other seeds within a radius by brute force.

Pinned by: recovery of noise-free renders to 1e-8 px, the Jacobian against central
differences, the lattice-sum identity sum_pixels erfc(u)/2 = pi (r^2 + s^2) for r >> s,
integer-shift and mirror equivariance (tests/test_oracle_erfdisk.py).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erfc

CONVERGED, MAX_ITER, SINGULAR, DIVERGED, TOO_FEW, OUTSIDE, BAD_START = range(7)
SQRT2 = np.sqrt(2.0)
INV_SQRT2PI = 1.0 / np.sqrt(2.0 * np.pi)


def seed_neighbours(seeds, radius, nb):
    """(K, nb) ids of the nb nearest other seeds within radius, by (squared distance, index); -1 pads."""
    s = np.asarray(seeds, dtype=np.float32).astype(np.float64).reshape(-1, 2)
    K = s.shape[0]
    out = np.full((K, nb), -1, dtype=np.int32)
    r2 = float(np.float32(radius)) ** 2
    for i in range(K):
        d = (s[:, 0] - s[i, 0]) ** 2 + (s[:, 1] - s[i, 1]) ** 2
        d[i] = np.inf
        cand = np.nonzero(d <= r2)[0]
        order = cand[np.lexsort((cand, d[cand]))][:nb]
        out[i, :order.size] = order
    return out


def _cell_forms(seeds, f, nbr_row):
    """(a, b, c, m, lower) per listed neighbour: f_j(q) = |q - s_j|^2 - |q - s_f|^2 =
    (a qx + b qy) + c, m = max(|a|, |b|), lower = neighbour index below f."""
    s = np.asarray(seeds, dtype=np.float32).astype(np.float64).reshape(-1, 2)
    x0, y0 = s[f]
    n0 = x0 * x0 + y0 * y0
    out = []
    for j in nbr_row:
        j = int(j)
        if j < 0 or j >= s.shape[0]:
            continue
        xj, yj = s[j]
        a, b = 2.0 * (x0 - xj), 2.0 * (y0 - yj)
        out.append((a, b, (xj * xj + yj * yj) - n0, max(abs(a), abs(b)), j < f))
    return out


def cell_mask(px, py, seeds, f, nbr_row):
    """True where pixel (px, py) lies in feature f's cell against the listed neighbours:
    every f_j(q) > 0, or = 0 towards a higher index."""
    inside = np.ones(np.shape(px), dtype=bool)
    for a, b, c, _, lower in _cell_forms(seeds, f, nbr_row):
        fj = (a * px + b * py) + c
        inside &= (fj > 0) | ((fj == 0) & (not lower))
    return inside


def cell_edge(px, py, seeds, f, nbr_row):
    """True where some 4-neighbour of (px, py) lies outside the cell: some
    g_j = f_j(q) - max(|a_j|, |b_j|) < 0, or = 0 towards a lower index."""
    edge = np.zeros(np.shape(px), dtype=bool)
    for a, b, c, m, lower in _cell_forms(seeds, f, nbr_row):
        gj = ((a * px + b * py) + c) - m
        edge |= (gj < 0) | ((gj == 0) & lower)
    return edge


def extract_patch(img: np.ndarray, seed_xy, half: int):
    """Values, validity, patch-local coordinates and the integer centre of one patch."""
    H, W = img.shape
    cx = float(np.rint(np.float32(seed_xy[0])))
    cy = float(np.rint(np.float32(seed_xy[1])))
    side = 2 * half + 1
    rows = np.arange(side)[:, None] + (int(cy) - half)
    cols = np.arange(side)[None, :] + (int(cx) - half)
    valid = (rows >= 0) & (rows < H) & (cols >= 0) & (cols < W)
    v = np.zeros((side, side), dtype=np.float64)
    rr = np.clip(rows, 0, H - 1)
    cc = np.clip(cols, 0, W - 1)
    v[:, :] = img[np.broadcast_to(rr, (side, side)), np.broadcast_to(cc, (side, side))]
    v[~valid] = 0.0
    fy, fx = np.meshgrid(np.arange(side) - half, np.arange(side) - half, indexing="ij")
    return v, valid, fx.astype(np.float64), fy.astype(np.float64), cx, cy


def moment_start(v, valid, half, init_sigma=1.0, ellipse=False, cell_edge=None):
    """Internal-order start [x0, y0, r, s, A, (q), B] or None for a flat patch. cell_edge:
    patch pixels with a 4-neighbour outside the feature's cell (added to the ring)."""
    side = 2 * half + 1
    ring = np.zeros((side, side), dtype=bool)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = True
    if cell_edge is not None:
        ring |= cell_edge
    ring &= valid
    B0 = float(v[ring].mean()) if ring.any() else 0.0
    w = v - B0
    vmax = float(w[valid].max())
    ins = valid & (w > 0.5 * vmax)
    s0 = float(w[ins].sum())
    nin = int(ins.sum())
    if not (vmax > 0) or not (s0 > 0) or nin < 1:
        return None
    fy, fx = np.meshgrid(np.arange(side) - half, np.arange(side) - half, indexing="ij")
    x0 = float((w[ins] * fx[ins]).sum() / s0)
    y0 = float((w[ins] * fy[ins]).sum() / s0)
    p = [x0, y0, np.sqrt(nin / np.pi), float(init_sigma), s0 / nin]
    if ellipse:
        p.append(1.0)
    p.append(B0)
    return np.asarray(p, dtype=np.float64)


def model_jac(p, fx, fy, ellipse):
    """Model values and Jacobian (n, P) in internal order at parameters ``p``."""
    x0, y0, r, s, A = p[:5]
    q = p[5] if ellipse else 1.0
    B = p[-1]
    dx = fx - x0
    dy = fy - y0
    rho = np.sqrt(dx * dx + (dy / q) ** 2)
    safe = np.where(rho > 0, rho, 1.0)
    u = (rho - r) / (SQRT2 * s)
    E = 0.5 * erfc(u)
    ez = np.exp(-u * u)
    G = A * INV_SQRT2PI / s * ez
    cols = [np.where(rho > 0, G * dx / safe, 0.0),
            np.where(rho > 0, G * dy / (q * q) / safe, 0.0),
            G,
            G * u * SQRT2,
            E]
    if ellipse:
        cols.append(np.where(rho > 0, G * dy * dy / (q ** 3) / safe, 0.0))
    cols.append(np.ones_like(E))
    return B + A * E, np.stack(cols, axis=-1)


def ldl_solve(M, g):
    """LDL^T solve with the kernel's pivot rule; None when a pivot is not positive."""
    P = M.shape[0]
    L = np.zeros((P, P))
    D = np.zeros(P)
    for j in range(P):
        d = M[j, j] - np.sum(L[j, :j] ** 2 * D[:j])
        if not (d > 1e-13 * M[j, j]) or not (d > 0):
            return None
        D[j] = d
        for i in range(j + 1, P):
            L[i, j] = (M[i, j] - np.sum(L[i, :j] * L[j, :j] * D[:j])) / d
    y = np.zeros(P)
    for i in range(P):
        y[i] = g[i] - np.sum(L[i, :i] * y[:i])
    x = np.zeros(P)
    for i in range(P - 1, -1, -1):
        x[i] = y[i] / D[i] - np.sum(L[i + 1:, i] * x[i + 1:])
    return x if np.all(np.isfinite(x)) else None


def lm_fit_patch(v, valid, half, ellipse=False, init=None, *, cell_edge=None, max_iter=30, step_tol=1e-6,
                 lambda0=1e-3, lambda_up=10.0, lambda_down=0.1, init_sigma=1.0,
                 accept_rtol=1e-6, budget=False, judge_rtol=1e-6):
    """Fit one patch. Returns (internal params or NaN, cost, iterations, status, npix).

    ``budget``: a budgeted (learned-start) pass of at most max_iter attempts with the residual
    judge at the start point and after every accepted step: there the undamped Gauss-Newton
    step d (J^T J d = J^T r) predicts the cost decrease d . J^T r; at most judge_rtol of the
    cost (and d keeps the point in bounds) -> d is applied and the fit is converged."""
    P = 7 if ellipse else 6
    npix = int(valid.sum())
    side = 2 * half + 1
    fy, fx = np.meshgrid(np.arange(side) - half, np.arange(side) - half, indexing="ij")
    fxv, fyv, vv = fx[valid], fy[valid], v[valid]
    nanp = np.full(P, np.nan)
    if npix < P + 2:
        return nanp, np.nan, 0, TOO_FEW, npix
    p = moment_start(v, valid, half, init_sigma, ellipse, cell_edge) if init is None else np.asarray(init, float)
    if p is None:
        return nanp, np.nan, 0, BAD_START, npix

    def evaluate(pp):
        mod, J = model_jac(pp, fxv, fyv, ellipse)
        res = vv - mod
        return J.T @ J, J.T @ res, float(res @ res)

    JTJ, JTr, cost = evaluate(p)
    lam = float(lambda0)
    lim = half + 0.5
    status = MAX_ITER if np.isfinite(cost) else DIVERGED
    it = 0
    any_solve = False

    def judge(p, JTJ, JTr, cost):
        d = ldl_solve(JTJ.copy(), JTr)
        if d is None:
            return None
        tr = p + d
        inside = (abs(tr[0]) <= lim and abs(tr[1]) <= lim and 0.25 < tr[2] < 2.0 * lim and 0.05 < tr[3] < lim)
        if ellipse:
            inside = inside and 0.25 < tr[5] < 4.0
        gain = float(d @ JTr)
        return tr if inside and 0.0 <= gain <= judge_rtol * cost else None

    fresh = budget
    while status == MAX_ITER and it < max_iter:
        if fresh:
            fresh = False
            tj = judge(p, JTJ, JTr, cost)
            if tj is not None:
                p = tj
                status = CONVERGED
                break
        it += 1
        M = JTJ.copy()
        dg = np.diag(M).copy()
        floor_d = 1e-12 * dg.max()
        M[np.diag_indices(P)] += lam * np.maximum(dg, floor_d)
        delta = ldl_solve(M, JTr)
        if delta is None:
            lam *= lambda_up
            continue
        any_solve = True
        tr = p + delta
        step = max(abs(delta[0]), abs(delta[1]), abs(delta[2]), abs(delta[3]))
        if ellipse:
            step = max(step, abs(delta[5]) * abs(p[2]))
        inside = (abs(tr[0]) <= lim and abs(tr[1]) <= lim and 0.25 < tr[2] < 2.0 * lim
                  and 0.05 < tr[3] < lim)
        if ellipse:
            inside = inside and 0.25 < tr[5] < 4.0
        if not inside:
            lam *= lambda_up
            continue
        if step < step_tol:
            # a step below the tolerance ends the fit: it is applied without a trial
            # evaluation (the cost, hence the rms, stays that of the point before it)
            p = tr
            status = CONVERGED
            break
        JTJ_t, JTr_t, cost_t = evaluate(tr)
        if cost_t <= cost * (1.0 + accept_rtol):
            p, JTJ, JTr, cost = tr, JTJ_t, JTr_t, cost_t
            lam = max(lam * lambda_down, 1e-12)
            fresh = budget
        else:
            lam *= lambda_up
    if fresh and status == MAX_ITER and np.isfinite(cost):
        tj = judge(p, JTJ, JTr, cost)
        if tj is not None:
            p = tj
            status = CONVERGED
    if not any_solve and status == MAX_ITER:
        status = SINGULAR
    if not np.isfinite(cost):
        status = DIVERGED
    if status <= MAX_ITER and (abs(p[0]) > half or abs(p[1]) > half):
        status = OUTSIDE
    return p, cost, it, status, npix


def to_output(p_internal, ellipse, cx, cy):
    """Internal order -> output row [x0, y0, r, s, A, B, q] in image coordinates."""
    p = p_internal
    q = p[5] if ellipse else 1.0
    return np.array([p[0] + cx, p[1] + cy, p[2], p[3], p[4], p[-1], q], dtype=np.float64)


def fit_batch(img, seeds, half, model, init=None, neighbours=None, only=None, learned_iters=0, **opts):
    """Oracle counterpart of spx_fit_batch (spx_fit_batch_cells with ``neighbours`` (K, nb)):
    (params (K,7), rms, iterations, status). ``only``: indices of the features to fit (the
    others stay NaN with status TOO_FEW); the cell masks still see every seed.
    ``learned_iters`` > 0: the budgeted fit of SpxFitOptions.learned_iters (from ``init``, or
    from the moment start without it) and, when it is not converged, the plain fit from the
    moment start; iterations count both."""
    img = np.asarray(img)
    seeds = np.asarray(seeds, dtype=np.float64).reshape(-1, 2)
    K = seeds.shape[0]
    half = np.broadcast_to(np.asarray(half, dtype=np.int64), (K,))
    model = np.broadcast_to(np.asarray(model, dtype=np.int64), (K,))
    params = np.full((K, 7), np.nan)
    rms = np.full(K, np.nan)
    iters = np.zeros(K, dtype=np.int64)
    status = np.zeros(K, dtype=np.int64)
    todo = range(K) if only is None else np.asarray(only, dtype=np.int64).ravel()
    if only is not None:
        status[:] = TOO_FEW
    for k in todo:
        k = int(k)
        h = int(half[k])
        ell = bool(model[k] == 1)
        if h < 2 or h > 31 or model[k] not in (0, 1):  # the kernel's accepted range
            status[k] = TOO_FEW
            continue
        v, valid, _, _, cx, cy = extract_patch(img, seeds[k], h)
        edge_px = None
        if neighbours is not None and (np.asarray(neighbours[k]) >= 0).any():
            side = 2 * h + 1
            gy, gx = np.meshgrid(np.arange(side) + (cy - h), np.arange(side) + (cx - h), indexing="ij")
            inside = cell_mask(gx, gy, seeds, k, neighbours[k])
            valid = valid & inside
            edge_px = inside & cell_edge(gx, gy, seeds, k, neighbours[k])
            v = np.where(valid, v, 0.0)
        p0 = None
        if init is not None:
            row = np.asarray(init[k], dtype=np.float64)
            p0 = [row[0] - cx, row[1] - cy, row[2], row[3], row[4]]
            if ell:
                p0.append(row[6])
            p0.append(row[5])
        if learned_iters > 0:
            kw = dict(opts, max_iter=int(learned_iters))
            p, cost, it, st, npix = lm_fit_patch(v, valid, h, ell, p0, cell_edge=edge_px, budget=True, **kw)
            if st not in (CONVERGED, TOO_FEW):
                p, cost, it2, st, npix = lm_fit_patch(v, valid, h, ell, None, cell_edge=edge_px, **opts)
                it += it2
        else:
            p, cost, it, st, npix = lm_fit_patch(v, valid, h, ell, p0, cell_edge=edge_px, **opts)
        params[k] = to_output(p, ell, cx, cy)
        P = 7 if ell else 6
        rms[k] = np.sqrt(cost / max(npix - P, 1)) if st not in (TOO_FEW, BAD_START) else np.nan
        iters[k] = it
        status[k] = st
    return params, rms, iters, status
