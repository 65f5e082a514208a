"""Oracle (test infrastructure only): FP64 NumPy restatement of the one-pass outline fit
(``fit_outline_fast``): one outline and one blur per page from the page's plain stack, with the
neighbours' light in the stack's model, then every hole centred against the complete page model
and the page fit repeated. Follows the reference:

  analytic template T^ = O^ x OTF x pixel box ......... subpx/_outline/onepass.py:38-96
  Band (half-plane frequencies, to_grid) .............. :101-126
  ShapeFitter: frequencies, classes, tables ........... :165-218
               to_classes ............................. :221-231
               outline points and derivatives ......... :246-279
               outline_hat / otf / template_hat ....... :281-355
               synth / evaluate ....................... :357-410
               stages, grid of starts, priors, LM ..... :413-623
  second_is_narrower / exchanged ...................... :627-644
  find_seeds .......................................... subpx/_gpu/outline_onepass.py:463-501
  CellList / nearest / window geometry ................ :504-519, 227-250, 528-560
  plain stack, hole_vs_stack, build_stack ............. :81-171, 573-625
  pair map (B-spline splat + FFT) ..................... :184-224, 663-698
  stack_data / start_theta ............................ :700-730
  set_template (Spec / Band grids, far kernel) ........ :733-793
  hole pass ........................................... :828-887 (via oracle.holes)
  good_mask / restack ................................. :889-903, 627-640
  page_fits / in_order / fit_images ................... :940-1119

Everything is FP64; the hole pass samples templates rounded to FP32 as the reference (and the
product) store them. Only tests/, the smoke check and the benchmark's CPU baseline use this module.
"""
from __future__ import annotations

import math

import numpy as np

from . import holes as H
from .outline import Template, render_model_page, stage_map

STACK_BINS = 4
PAIR_BINS = 8
F_MAX = 0.6
N_POLY = 256
PAIR_REACH_MARGIN_PX = 13.0
FIRST_STAGES, FIRST_ITERATIONS, FIRST_FTOL = ("round", "super"), (10, 12), 1e-6
FINAL_ITERATIONS = 30
CENTRING_SWEEPS = 2
GRID_SIZES = np.linspace(0.4, 1.8, 15)
GRID_WIDTHS = np.linspace(0.45, 2.2, 12)


def lib_template(blur):
    """The parameter bookkeeping template of the one-pass fit (outline_onepass.py:440-445)."""
    return Template(upsample=8, size_px=24, blur=blur)


# ------------------------------------------------------------------------------- analytic template
def outline_polygon(t, theta, n_poly, delta=0.0):
    """Vertices (x, y) of the outline at n_poly equal angles, moved delta px along its normal by
    the radial rule r -> r + delta hypot(r, r') / r (onepass.py:44-50)."""
    phi = 2.0 * np.pi * np.arange(n_poly) / n_poly
    r, dr = t.radius_at(np.asarray(theta, float), phi)
    if delta:
        r = r + delta * np.hypot(r, dr) / r
    return r * np.cos(phi), r * np.sin(phi)


def polygon_hat(x, y, fx, fy):
    """Fourier transform of the polygon's inside at frequencies (fx, fy) (any equal shapes):
    i / (2 pi |f|^2) sum_edges (f x e) sinc(f . e) exp(-2 pi i f . m), the area at f = 0
    (onepass.py:52-70; divergence theorem, exact for a polygon)."""
    x2, y2 = np.roll(x, -1), np.roll(y, -1)
    ex, ey = x2 - x, y2 - y
    mx, my = 0.5 * (x + x2), 0.5 * (y + y2)
    area = 0.5 * float(np.sum(x * y2 - x2 * y))
    shape = np.shape(fx)
    f_x = np.asarray(fx, float).reshape(-1, 1)
    f_y = np.asarray(fy, float).reshape(-1, 1)
    cross = f_x * ey[None] - f_y * ex[None]
    dot = f_x * ex[None] + f_y * ey[None]
    s = (cross * np.sinc(dot) * np.exp(-2j * np.pi * (f_x * mx[None] + f_y * my[None]))).sum(axis=1)
    f2 = (f_x * f_x + f_y * f_y)[:, 0]
    out = np.where(f2 == 0, area, 1j * s / (2.0 * np.pi * np.where(f2 == 0, 1.0, f2)))
    return out.reshape(shape)


def blur_otf(t, theta, fx, fy):
    """Blur transfer function of theta's blur part (onepass.py:73-87)."""
    bp = np.asarray(theta, float)[t.n_shape:]
    c = 2.0 * np.pi ** 2
    if t.blur == "gauss2":
        s1x, s1y, s2x, s2y = np.exp(bp[:4])
        w2 = 1.0 / (1.0 + math.exp(-bp[4]))
        g1 = np.exp(-c * (s1x ** 2 * fx ** 2 + s1y ** 2 * fy ** 2))
        g2 = np.exp(-c * (s2x ** 2 * fx ** 2 + s2y ** 2 * fy ** 2))
        return (1.0 - w2) * g1 + w2 * g2
    sx, sy = np.exp(bp[:2])
    g = np.exp(-c * (sx ** 2 * fx ** 2 + sy ** 2 * fy ** 2))
    if t.blur == "lens":
        s = np.clip(np.hypot(fx, fy) / math.exp(bp[2]), 0.0, 1.0)
        g = g * (2.0 / np.pi) * (np.arccos(s) - s * np.sqrt(1.0 - s * s))
    return g


def analytic_hat(t, theta, fx, fy, n_poly=512, delta=0.0, box_px=0.0):
    """T^(f) = O^(outline moved by delta) x OTF x sinc(fx) sinc(fy) [x the box of box_px]
    (onepass.py:89-96)."""
    x, y = outline_polygon(t, theta, n_poly, delta)
    out = polygon_hat(x, y, fx, fy) * blur_otf(t, theta, fx, fy) * np.sinc(fx) * np.sinc(fy)
    if box_px > 0:
        out = out * np.sinc(fx * box_px) * np.sinc(fy * box_px)
    return out


class Band:
    """Frequencies k / L with |kx|, |ky| <= f_max L in the half plane kx >= 0 (ky in FFT order)
    and their synthesis on a grid of U samples per px over the box L, origin at sample N // 2
    (onepass.py:101-126)."""

    def __init__(self, box_px=48, f_max=0.75, upsample=8):
        self.L, self.U = int(box_px), int(upsample)
        self.N = self.L * self.U
        self.kmax = int(math.floor(f_max * self.L))
        ky = np.concatenate([np.arange(0, self.kmax + 1), np.arange(-self.kmax, 0)])
        kx = np.arange(0, self.kmax + 1)
        self.fy, self.fx = np.meshgrid(ky / self.L, kx / self.L, indexing="ij")
        self.rows = np.where(ky >= 0, ky, self.N + ky)
        self.centre = np.exp(-2j * np.pi * (self.fx + self.fy) * ((self.N // 2) / self.U))

    def to_grid(self, coeff):
        """Real (..., N, N) samples of sum_k coeff_k exp(2 pi i f_k . u) / L^2 at u = (n - N // 2) / U."""
        full = np.zeros(coeff.shape[:-2] + (self.N, self.N // 2 + 1), dtype=complex)
        full[..., self.rows, :self.kmax + 1] = coeff * self.centre
        return np.fft.irfft2(full, s=(self.N, self.N)) * self.U ** 2


# ---------------------------------------------------------------------------------- page fitter
class ShapeFitter:
    """Outline + blur of a batch of pages from their plain stacks by Levenberg-Marquardt, scale
    and constant solved exactly for every theta (onepass.py:165-623). Model of the stack's bin at
    offset u_b of sub-pixel class q: S(b) = scale Re sum_f T^(f) Gamma_q(f) exp(2 pi i f . u_b) / L^2
    + const over the frequencies |f| <= f_max of the half plane (kx > 0 counted twice)."""

    def __init__(self, t, *, box_px=48, f_max=0.6, U=4, R=7, n_poly=256):
        self.t, self.blur, self.ks = t, t.blur, t.ks
        self.n_shape, self.n_blur, self.P = t.n_shape, t.n_blur, t.n_params
        self.L, self.U, self.R, self.n_poly = int(box_px), int(U), int(R), int(n_poly)
        self.phi = 2.0 * np.pi * np.arange(self.n_poly) / self.n_poly
        kmax = int(math.floor(f_max * self.L))
        KY, KX = np.meshgrid(np.arange(-kmax, kmax + 1), np.arange(0, kmax + 1), indexing="ij")
        keep = (KY * KY + KX * KX) <= (f_max * self.L) ** 2
        self.kmax, self.nky, self.nkx = kmax, 2 * kmax + 1, kmax + 1
        self.sel = np.flatnonzero(keep.ravel())
        self.F = len(self.sel)
        self.fx = (KX / self.L).ravel()[self.sel]
        self.fy = (KY / self.L).ravel()[self.sel]
        self.pix = np.sinc(self.fx) * np.sinc(self.fy)
        self.dbl = np.where(KX.ravel()[self.sel] == 0, 1.0, 2.0)
        self.nmax = int(math.floor((self.R - 0.5) * self.U))
        n = np.arange(-self.nmax, self.nmax + 1)
        cls = np.mod(n, self.U)
        self.npts = max(int(np.sum(cls == q)) for q in range(self.U))
        self.idx = np.full((self.U, self.npts), -1, dtype=np.int64)
        for q in range(self.U):
            where = np.flatnonzero(cls == q)
            self.idx[q, :len(where)] = where
        self.pad = self.idx >= 0
        u = np.where(self.pad, n[np.clip(self.idx, 0, None)], 0) / self.U
        kyv = np.arange(-kmax, kmax + 1) / self.L
        kxv = np.arange(0, kmax + 1) / self.L
        self.Ey = np.exp(2j * np.pi * u[:, :, None] * kyv[None, None]) * self.pad[:, :, None]
        self.ExT = (np.exp(2j * np.pi * u[:, :, None] * kxv[None, None]) * self.pad[:, :, None]).transpose(0, 2, 1) \
            / self.L ** 2
        self.lo, self.hi = t.bounds()

    # bins by class
    def to_classes(self, grid):
        """(..., Ms, Ms) bins (hole centre in the middle) -> (..., U, U, npts, npts) by sub-pixel
        class, 0 where padded; and the mask of real bins."""
        pos = grid.shape[-1] // 2 - self.nmax + np.clip(self.idx, 0, None)
        out = grid[..., pos[:, None, :, None], pos[None, :, None, :]]
        mask = self.pad[:, None, :, None] & self.pad[None, :, None, :]
        return np.where(mask, out, 0.0), mask

    # model
    def outline_geometry(self, th):
        """Outline points, tangents, r dr/d(shape params) (B, n, n_shape), areas (B,), dphi."""
        phi = self.phi[None, :]
        a, b = np.exp(th[:, 0:1]), np.exp(th[:, 1:2])
        p = 1.0 + np.exp(th[:, 2:3])
        tt = phi - th[:, 3:4]
        ct, st = np.cos(tt), np.sin(tt)
        ua, ub = np.abs(ct / a), np.abs(st / b)
        A, Bq = ua ** p, ub ** p
        S = A + Bq
        base = S ** (-1.0 / p)
        dbase = -(1.0 / p) * base ** (1.0 + p) * (p * ua ** (p - 1.0) * np.sign(ct) * (-st / a)
                                                   + p * ub ** (p - 1.0) * np.sign(st) * (ct / b))
        mod, dmod = np.ones_like(tt), np.zeros_like(tt)
        for i, k in enumerate(self.ks):
            ck, sk = th[:, 4 + 2 * i:5 + 2 * i], th[:, 5 + 2 * i:6 + 2 * i]
            mod = mod + ck * np.cos(k * phi) + sk * np.sin(k * phi)
            dmod = dmod - ck * k * np.sin(k * phi) + sk * k * np.cos(k * phi)
        r = base * mod
        dr = dbase * mod + base * dmod
        with np.errstate(divide="ignore", invalid="ignore"):
            la = np.where(ua > 0, A * np.log(np.maximum(ua, 1e-300)), 0.0)
            lb = np.where(ub > 0, Bq * np.log(np.maximum(ub, 1e-300)), 0.0)
        g = [base * (A / S) * mod, base * (Bq / S) * mod,
             (p - 1.0) * base * (np.log(S) / (p * p) - (la + lb) / (p * S)) * mod, -dbase * mod]
        for k in self.ks:
            g += [base * np.cos(k * phi) + 0.0 * tt, base * np.sin(k * phi) + 0.0 * tt]
        g = np.stack(g, axis=-1) * r[..., None]
        x, y = r * np.cos(phi), r * np.sin(phi)
        xd, yd = dr * np.cos(phi) - r * np.sin(phi), dr * np.sin(phi) + r * np.cos(phi)
        dphi = 2.0 * np.pi / self.n_poly
        return x, y, xd, yd, g, 0.5 * (r * r).sum(axis=1) * dphi, dphi

    def outline_hat(self, th, grads=True):
        """O^ (trapezoid rule over the outline) and its shape derivatives: (B, F, 1 + n_shape)."""
        x, y, xd, yd, g, area, dphi = self.outline_geometry(np.asarray(th, float))
        E = np.exp(-2j * np.pi * (self.fx[None, :, None] * x[:, None, :] + self.fy[None, :, None] * y[:, None, :]))
        f2 = self.fx ** 2 + self.fy ** 2
        s0 = self.fx[None] * (E @ yd[:, :, None])[..., 0] - self.fy[None] * (E @ xd[:, :, None])[..., 0]
        o0 = np.where(f2[None] == 0, area[:, None], 1j * s0 * dphi / (2.0 * np.pi * np.where(f2 == 0, 1.0, f2))[None])
        if not grads:
            return o0[..., None]
        return np.concatenate([o0[..., None], (E @ g) * dphi], axis=-1)

    def otf(self, th, grads=True):
        """OTF and its blur derivatives: (B, F, 1 + n_blur)."""
        bp = np.asarray(th, float)[:, self.n_shape:]
        fx2, fy2 = (self.fx ** 2)[None], (self.fy ** 2)[None]
        c = 2.0 * np.pi ** 2
        if self.blur == "gauss2":
            s1x, s1y, s2x, s2y = (np.exp(bp[:, i:i + 1]) for i in range(4))
            w2 = 1.0 / (1.0 + np.exp(-bp[:, 4:5]))
            g1 = np.exp(-c * (s1x ** 2 * fx2 + s1y ** 2 * fy2))
            g2 = np.exp(-c * (s2x ** 2 * fx2 + s2y ** 2 * fy2))
            out = [(1 - w2) * g1 + w2 * g2]
            if grads:
                out += [(1 - w2) * g1 * (-2 * c * s1x ** 2 * fx2), (1 - w2) * g1 * (-2 * c * s1y ** 2 * fy2),
                        w2 * g2 * (-2 * c * s2x ** 2 * fx2), w2 * g2 * (-2 * c * s2y ** 2 * fy2),
                        (g2 - g1) * w2 * (1 - w2)]
            return np.stack(out, axis=-1)
        sx, sy = np.exp(bp[:, 0:1]), np.exp(bp[:, 1:2])
        g = np.exp(-c * (sx ** 2 * fx2 + sy ** 2 * fy2))
        if self.blur == "lens":
            s = np.clip(np.sqrt(fx2 + fy2) / np.exp(bp[:, 2:3]), 0.0, 1.0)
            root = np.sqrt(1.0 - s * s)
            lens = (2 / np.pi) * (np.arccos(s) - s * root)
            out = [g * lens]
            if grads:
                out += [g * lens * (-2 * c * sx ** 2 * fx2), g * lens * (-2 * c * sy ** 2 * fy2),
                        g * (4 / np.pi) * s * root]
            return np.stack(out, axis=-1)
        out = [g]
        if grads:
            out += [g * (-2 * c * sx ** 2 * fx2), g * (-2 * c * sy ** 2 * fy2)]
        return np.stack(out, axis=-1)

    def template_hat(self, th, tm=None):
        """T^ (B, F, 1); with tm (P, nf) also dT^ along the stage's free directions (B, F, 1 + nf)."""
        grads = tm is not None
        o = self.outline_hat(th, grads)
        b = self.otf(th, grads)
        pix = self.pix[None, :, None]
        if not grads:
            return o * b * pix
        full = np.concatenate([o[..., 1:] * b[..., :1], o[..., :1] * b[..., 1:]], axis=-1)
        return np.concatenate([o[..., :1] * b[..., :1], full @ tm], axis=-1) * pix

    def spectrum(self, coeff):
        """Band coefficients (B, F, C) on the (nky, nkx) grid, kx > 0 doubled: (B, C, nky, nkx)."""
        B, F, C = coeff.shape
        full = np.zeros((B, C, self.nky * self.nkx), dtype=complex)
        full[:, :, self.sel] = (coeff * self.dbl[None, :, None]).transpose(0, 2, 1)
        return full.reshape(B, C, self.nky, self.nkx)

    def synth(self, coeff, gamma):
        """Stack model by class: (B, C, U, U, npts, npts)."""
        full = self.spectrum(coeff)[:, :, None, None] * gamma[:, None]
        left = self.Ey[None, None, :, None] @ full
        return (left @ self.ExT[None, None, None, :]).real

    def evaluate(self, th, y, w, gamma, tm=None):
        """cost, scale, const (B,) and with tm also g = J^T r (B, nf) and A = J^T J (B, nf, nf),
        scale and constant projected out (variable projection, Kaufman's Jacobian)."""
        B = y.shape[0]
        M = self.synth(self.template_hat(th, tm), gamma)
        M = M.reshape(B, M.shape[1], -1)
        y = y.reshape(B, -1)
        w = w.reshape(B, -1)
        m = M[:, 0]
        ws = w.sum(axis=1)
        mw = (w * m).sum(axis=1) / ws
        yw = (w * y).sum(axis=1) / ws
        dm = m - mw[:, None]
        dy = y - yw[:, None]
        var = (w * dm * dm).sum(axis=1)
        scale = (w * dm * dy).sum(axis=1) / np.where(var > 0, var, 1.0)
        const = yw - scale * mw
        res = scale[:, None] * dm - dy
        cost = (w * res * res).sum(axis=1)
        if tm is None:
            return cost, scale, const
        sw = np.sqrt(w)
        r = sw * res
        D = M[:, 1:] * (scale[:, None] * sw)[:, None, :]
        u1 = sw / np.sqrt(ws)[:, None]
        u2 = sw * dm / np.sqrt(np.maximum(var, 1e-300))[:, None]
        D = D - u1[:, None, :] * (D @ u1[:, :, None])
        D = D - u2[:, None, :] * (D @ u2[:, :, None])
        return cost, scale, const, (D @ r[:, :, None])[:, :, 0], D @ D.transpose(0, 2, 1)

    # fitting
    def stage(self, name):
        free, ties = stage_map(self.t, name)
        tm = np.zeros((self.P, len(free)))
        for i, k in enumerate(free):
            tm[k, i] = 1.0
        for k, src in ties.items():
            tm[k, free.index(src)] = 1.0
        return free, tm

    def grid_start(self, base, y, w, gamma):
        """The round stage's grid of starts: 15 sizes x 12 blur widths compared on each page's most
        populated sub-pixel class; the cheapest per page (onepass.py:422-486)."""
        B = base.shape[0]
        j = self.n_shape
        sizes = np.exp(base[:, 0])[:, None] * GRID_SIZES[None, :]
        tha = np.repeat(base[:, None, :], 15, axis=1)
        tha[:, :, 0] = tha[:, :, 1] = np.log(sizes)
        ths = np.repeat(base[:1], 12, axis=0)
        ths[:, j] = ths[:, j + 1] = np.log(GRID_WIDTHS)
        if self.blur == "gauss2":
            ths[:, j + 2] = ths[:, j + 3] = np.log(2 * GRID_WIDTHS)
            ths[:, j + 4] = math.log(0.05 / 0.95)
        elif self.blur == "lens":
            ths[:, j + 2] = math.log(0.5)
        tha = np.clip(tha, self.lo + 1e-6, self.hi - 1e-6)
        ths = np.clip(ths, self.lo + 1e-6, self.hi - 1e-6)
        bl = self.otf(ths, False)[..., 0]
        U, npts = self.U, self.npts
        yc = y.reshape(B, U * U, npts * npts)
        wc = w.reshape(B, U * U, npts * npts)
        top = wc.sum(axis=2).argmax(axis=1)
        out = base.copy()
        tables = []
        for b in range(B):
            qy, qx = divmod(int(top[b]), U)
            o = self.outline_hat(tha[b], False)[..., 0]
            coeff = (o[:, None, :] * bl[None] * self.pix[None, None]).reshape(180, self.F)
            spec = self.spectrum(coeff.T[None])[0] * gamma[b, qy, qx][None]
            M = ((self.Ey[qy][None] @ spec) @ self.ExT[qx][None]).real.reshape(180, -1)
            yb, wb = yc[b, top[b]][None], wc[b, top[b]][None]
            ws = wb.sum()
            dm = M - (wb * M).sum(axis=1, keepdims=True) / ws
            dy = yb - (wb * yb).sum() / ws
            var = (wb * dm * dm).sum(axis=1, keepdims=True)
            scale = (wb * dm * dy).sum(axis=1, keepdims=True) / np.where(var > 0, var, 1.0)
            table = (wb * (scale * dm - dy) ** 2).sum(axis=1)
            ia, isg = divmod(int(np.argmin(table)), 12)
            out[b, 0] = out[b, 1] = tha[b, ia, 0]
            out[b, j:] = ths[isg, j:]
            tables.append(table.reshape(15, 12))
        self.last_grid = tables
        return out

    def default_prior(self):
        """Weak priors (chi-square units) on the second Gaussian's widths and weight (onepass.py:488-502)."""
        j = self.n_shape
        rows = []
        if self.blur == "gauss2":
            for k in (0, 1):
                L = np.zeros(self.P)
                L[j + 2 + k], L[j + k] = 1.0, -1.0
                rows.append((L, math.log(2.0), 0.7))
            L = np.zeros(self.P)
            L[j + 4] = 1.0
            rows.append((L, math.log(0.05 / 0.95), 1.5))
        return rows

    def fit(self, theta0, gamma, y, w, stages=("round", "super", "full"), grid=True, max_iter=(12, 20, 30),
            ftol=1e-7, xtol=1e-7):
        """Staged Levenberg-Marquardt for the batch (onepass.py:504-623): bounds as an active set,
        Nielsen's gain-ratio damping, the weak priors. Returns theta (B, P), scale, const, cost,
        [(stage, iterations, stopped (B,))]."""
        theta = np.clip(np.asarray(theta0, float).copy(), self.lo + 1e-6, self.hi - 1e-6)
        B = theta.shape[0]
        rows = self.default_prior()
        PL = np.stack([r[0] for r in rows]) if rows else np.zeros((0, self.P))
        pm = np.array([r[1] for r in rows])
        pw = np.array([1.0 / r[2] ** 2 for r in rows])
        if isinstance(max_iter, int):
            max_iter = (max_iter,) * len(stages)
        info = []
        scale = const = cost = None
        for stage, cap in zip(stages, max_iter):
            free, tm = self.stage(stage)
            base = theta.copy()
            if stage == "round":
                base[:, 1] = base[:, 0]
                base[:, 2] = 0.0
                base[:, 3] = 0.0
                base[:, 4:self.n_shape] = 0.0
                if grid:
                    base = self.grid_start(base, y, w, gamma)
            lo, hi = self.lo[free] + 1e-6, self.hi[free] - 1e-6
            fixed = base * (1.0 - tm.sum(axis=1))[None, :]
            nf = len(free)
            pj = PL @ tm
            pa = (pj.T * pw[None, :]) @ pj

            def expand(q):
                return fixed + q @ tm.T

            def evaluate(q):
                th = expand(q)
                c, s, k, g, A = self.evaluate(th, y, w, gamma, tm)
                dev = th @ PL.T - pm[None, :]
                return c + (pw[None, :] * dev * dev).sum(axis=1), s, k, g + (dev * pw[None, :]) @ pj, A + pa[None]

            q = np.clip(base[:, free], lo, hi)
            cost, scale, const, g, A = evaluate(q)
            lam = np.full(B, 1e-3)
            nu = np.full(B, 2.0)
            done = np.zeros(B, dtype=bool)
            its = 0
            for its in range(1, cap + 1):
                q_new = q.copy()
                for b in range(B):
                    if not done[b]:
                        q_new[b] = np.clip(q[b] + bounded_step(q[b], g[b], A[b], lam[b], lo, hi), lo, hi)
                dq = q_new - q
                cost2, scale2, const2, g2, A2 = evaluate(q_new)
                pred = -(2.0 * (g * dq).sum(axis=1) + np.einsum("bi,bij,bj->b", dq, A, dq))
                rho = (cost - cost2) / np.maximum(pred, 1e-300)
                better = (cost2 < cost) & ~done
                small = (((cost - cost2) <= ftol * cost)
                         | (np.abs(dq).max(axis=1) <= xtol * (xtol + np.abs(q).max(axis=1))))
                q = np.where(better[:, None], q_new, q)
                A = np.where(better[:, None, None], A2, A)
                g = np.where(better[:, None], g2, g)
                scale, const = np.where(better, scale2, scale), np.where(better, const2, const)
                done = done | (better & small) | (~better & (lam > 1e7)) | (pred <= ftol * cost)
                cost = np.where(better, cost2, cost)
                grow = np.maximum(1.0 / 3.0, 1.0 - (2.0 * rho - 1.0) ** 3)
                lam = np.where(better, lam * grow, lam * nu)
                nu = np.where(better, 2.0, nu * 2.0)
                if done.all():
                    break
            theta = expand(q)
            info.append((stage, its, done.copy()))
        return theta, scale, const, cost, info


def bounded_step(q, g, A, lam, lo, hi):
    """The damped Gauss-Newton step of one page with the bounds as an active set: a parameter on a
    bound whose downhill direction points outside takes no step; one the step carries past a bound
    is put on it and the rest solved again (onepass.py:559-585)."""
    nf = len(q)
    frozen = ((q <= lo) & (g > 0)) | ((q >= hi) & (g < 0))
    held = np.zeros(nf)
    step = held
    for _ in range(nf):
        k = ~frozen
        if not k.any():
            step = held
            break
        Ak = A[np.ix_(k, k)]
        d = np.maximum(np.diag(Ak), 1e-10 * np.diag(Ak).max())
        rhs = -(g[k] + A[np.ix_(k, ~k)] @ held[~k])
        try:
            sk = np.linalg.solve(Ak + np.diag(lam * d), rhs)
        except np.linalg.LinAlgError:
            sk = np.zeros(int(k.sum()))
        step = held.copy()
        step[k] = np.where(np.isfinite(sk), sk, 0.0)
        over = (q + step > hi) & k
        under = (q + step < lo) & k
        if not (over | under).any():
            break
        held = np.where(over, hi - q, np.where(under, lo - q, held))
        frozen = frozen | over | under
    return step


def second_is_narrower(t, theta):
    """gauss2: the second Gaussian is the narrower (sum of its log widths) (onepass.py:627-630)."""
    j = t.n_shape
    return bool(theta[j + 2] + theta[j + 3] < theta[j] + theta[j + 1])


def exchanged(t, theta):
    """gauss2 with its two Gaussians exchanged (s1 <-> s2, w2 -> 1 - w2), clipped into the bounds
    (onepass.py:633-644)."""
    j = t.n_shape
    lo, hi = t.bounds()
    theta = np.asarray(theta, dtype=float)
    out = theta.copy()
    out[j:j + 2], out[j + 2:j + 4] = theta[j + 2:j + 4], theta[j:j + 2]
    out[j + 4] = -theta[j + 4]
    return np.clip(out, lo + 1e-6, hi - 1e-6)


# ---------------------------------------------------------------------------------------- seeds
def gaussian_taps(sigma, truncate=2.5):
    """Normalised Gaussian taps w[0..r], r = int(truncate sigma + 0.5) (scipy's gaussian_filter1d)."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (float(sigma) ** 2) * x * x)
    return (phi / phi.sum())[r:]


def smooth(img, sigma):
    """Separable Gaussian smoothing, 'nearest' edges, rows (axis 0) first, FP64; each output is
    w0 x + sum_{k=r..1} w_k (x[-k] + x[+k]) in that order."""
    w = gaussian_taps(sigma)
    r = len(w) - 1
    out = np.asarray(img, dtype=np.float64)
    for axis in (0, 1):
        n = out.shape[axis]
        idx = np.arange(n)

        def take(k):
            return np.take(out, np.clip(idx + k, 0, n - 1), axis=axis)

        acc = w[0] * out
        for k in range(r, 0, -1):
            acc = acc + w[k] * (take(-k) + take(k))
        out = acc
    return out


def window_max(a, nms):
    """Maximum over the (2 nms + 1)^2 window, 'nearest' edges."""
    n0, n1 = a.shape
    out = np.full_like(a, -np.inf)
    i0, i1 = np.arange(n0), np.arange(n1)
    for dy in range(-nms, nms + 1):
        rows = a[np.clip(i0 + dy, 0, n0 - 1)]
        for dx in range(-nms, nms + 1):
            out = np.maximum(out, rows[:, np.clip(i1 + dx, 0, n1 - 1)])
    return out


def find_seeds(img, smooth_sigma=1.0, nms=2, rel_threshold=0.3):
    """Bright blobs' centres (N, 2) [x, y] (outline_onepass.py:463-501): local maxima (+-nms px)
    of the smoothed page above lo + rel_threshold (hi - lo) (lo, hi = 5th / 99.8th percentiles of
    every 4th pixel), not on the frame; the window and the smoothing widen (up to 3 tries) while
    the blobs are wider than the window; each maximum refined by a parabola through the log of
    the height above lo - 0.02 (hi - lo) along x and along y, clipped to +-0.5 px."""
    img = np.asarray(img, dtype=np.float64)
    sub = img[::4, ::4].ravel()
    lo, hi = (float(v) for v in np.percentile(sub, [5.0, 99.8]))
    thr = lo + rel_threshold * (hi - lo)
    above = float((img[::2, ::2] > thr).mean()) * img.size
    sm = None
    ys = xs = np.zeros(0, dtype=np.int64)
    for _ in range(3):
        sm = smooth(img, smooth_sigma)
        peak = (sm == window_max(sm, nms)) & (sm > thr)
        peak[0, :] = peak[-1, :] = False
        peak[:, 0] = peak[:, -1] = False
        ys, xs = np.nonzero(peak)
        n = len(xs)
        if n == 0:
            break
        width = math.sqrt(4.0 * above / (math.pi * n))
        if width <= 2 * nms + 1:
            break
        nms = int(math.ceil(0.6 * width))
        smooth_sigma = max(smooth_sigma, 0.2 * width)
    if len(xs) == 0:
        return np.zeros((0, 2))
    base = lo - 0.02 * (hi - lo)

    def lg(v):
        return np.log(np.maximum(v - base, 1e-6))

    c0 = lg(sm[ys, xs])

    def offset(lft, rgt):
        den = lft - 2 * c0 + rgt
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.clip(np.where(den < 0, 0.5 * (lft - rgt) / den, 0.0), -0.5, 0.5)

    dx = offset(lg(sm[ys, xs - 1]), lg(sm[ys, xs + 1]))
    dy = offset(lg(sm[ys - 1, xs]), lg(sm[ys + 1, xs]))
    return np.stack([xs + dx, ys + dy], axis=1).astype(np.float64)


# ------------------------------------------------------------------------------------ cell list
class CellList:
    """Holes sorted by square cells of side cs (outline_onepass.py:504-519)."""

    def __init__(self, c, h, w, cs):
        self.cs = float(cs)
        self.ncx = int(math.ceil(w / cs)) + 1
        self.ncy = int(math.ceil(h / cs)) + 1
        cx = np.clip(np.floor(c[:, 0] / cs), 0, self.ncx - 1).astype(np.int64)
        cy = np.clip(np.floor(c[:, 1] / cs), 0, self.ncy - 1).astype(np.int64)
        self.cell = cy * self.ncx + cx
        self.order = np.argsort(self.cell, kind="stable")
        self.start = np.searchsorted(self.cell[self.order], np.arange(self.ncx * self.ncy + 1))

    def members(self, cx, cy, r):
        """Holes of the cells within r cells of (cx, cy), in cell order."""
        out = []
        for yy in range(max(cy - r, 0), min(cy + r, self.ncy - 1) + 1):
            x0, x1 = max(cx - r, 0), min(cx + r, self.ncx - 1)
            if x0 <= x1:
                out.append(self.order[self.start[yy * self.ncx + x0]:self.start[yy * self.ncx + x1 + 1]])
        return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


def nearest(c, cells):
    """Distance to the nearest other hole within the 3 x 3 cells round the hole's cell; 1e15 when
    none is there (outline_onepass.py:227-250)."""
    out = np.empty(len(c))
    for i in range(len(c)):
        cx, cy = int(math.floor(c[i, 0] / cells.cs)), int(math.floor(c[i, 1] / cells.cs))
        js = cells.members(cx, cy, 1)
        js = js[js != i]
        d = np.min((c[js, 0] - c[i, 0]) ** 2 + (c[js, 1] - c[i, 1]) ** 2) if len(js) else 1e30
        out[i] = math.sqrt(d)
    return out


# ----------------------------------------------------------------------------------------- page
class Page:
    """One image through steps A to C (outline_onepass.py:520-935)."""

    def __init__(self, image, seeds, blur="gauss2"):
        self.img = np.asarray(image, dtype=np.float32)
        self.h, self.w = self.img.shape
        self.U, self.pair_U, self.f_max = STACK_BINS, PAIR_BINS, F_MAX
        self.blur = blur
        self.t = lib_template(blur)
        self.c = np.array(seeds, dtype=np.float64).reshape(-1, 2)
        self.n = len(self.c)
        if self.n < 2:
            raise ValueError(f"{self.n} holes found: no page fit")
        self.cells = CellList(self.c, self.h, self.w, 16.0)
        nn = nearest(self.c, self.cells)
        found = nn < 1e20
        self.spacing = float(np.median(nn[found])) if found.any() else 16.0
        if self.spacing > 15.0:
            self.cells = CellList(self.c, self.h, self.w, 2.2 * self.spacing)
            nn = nearest(self.c, self.cells)
            self.spacing = float(np.median(nn[nn < 1e20]))
        self.nn = nn
        self.R = max(3, int(math.ceil(0.75 * self.spacing)))
        self.Ms = 2 * int(math.ceil((self.R + 0.5) * self.U)) + 1
        self.reach = float(self.R) + PAIR_REACH_MARGIN_PX
        self.L = int(8 * math.ceil((2 * self.reach + 8) / 8))
        self.a = np.ones(self.n)
        self.level = np.zeros(self.n)

    # stack
    def interior(self):
        m = self.R + 1
        c = self.c
        return (c[:, 0] >= m) & (c[:, 0] <= self.w - 1 - m) & (c[:, 1] >= m) & (c[:, 1] <= self.h - 1 - m)

    def _window(self, i):
        """Pixels of hole i's window inside the frame whose bin lies in the stack: (py, px, by, bx, uy, ux)."""
        x, y = self.c[i]
        R, U, Ms = self.R, self.U, self.Ms
        off = np.arange(-R, R + 1)
        py = int(np.rint(y)) + off[:, None] + 0 * off[None, :]
        px = int(np.rint(x)) + off[None, :] + 0 * off[:, None]
        uy, ux = py - y, px - x
        by = np.rint(uy * U).astype(np.int64) + Ms // 2
        bx = np.rint(ux * U).astype(np.int64) + Ms // 2
        ok = (py >= 0) & (py < self.h) & (px >= 0) & (px < self.w) & (by >= 0) & (by < Ms) & (bx >= 0) & (bx < Ms)
        return py[ok], px[ok], by[ok], bx[ok], uy[ok], ux[ok]

    def stack_of(self, use):
        """The plain mean of the windows of the holes `use` on the quarter-pixel bins: stack (NaN
        where no pixel fell) and pixels per bin (outline_onepass.py:81-107, 582-592)."""
        Ms = self.Ms
        sums, wsum = np.zeros(Ms * Ms), np.zeros(Ms * Ms)
        for i in np.flatnonzero(use):
            py, px, by, bx, _, _ = self._window(i)
            np.add.at(sums, by * Ms + bx, self.img[py, px].astype(np.float64))
            np.add.at(wsum, by * Ms + bx, 1.0)
        with np.errstate(invalid="ignore", divide="ignore"):
            stack = np.where(wsum > 0, sums / np.maximum(wsum, 1e-300), np.nan)
        return stack.reshape(Ms, Ms), wsum.reshape(Ms, Ms)

    def versus(self, stack, core_px):
        """Per hole: brightness and level by 2 x 2 least squares against the stack (pixels within
        core_px), leftover rms over the window, one Gauss-Newton step of the centre (core pixels):
        (n, 6) = a, level, rms, dx, dy, pixels (outline_onepass.py:112-171)."""
        filled = np.where(np.isfinite(stack), stack, 0.0)
        gy, gx = np.gradient(filled, 1.0 / self.U)
        S, Sx, Sy = stack.ravel(), gx.ravel(), gy.ravel()
        core2 = core_px * core_px
        out = np.zeros((self.n, 6))
        for i in range(self.n):
            py, px, by, bx, uy, ux = self._window(i)
            b = by * self.Ms + bx
            s = S[b]
            fin = np.isfinite(s)
            v = self.img[py, px].astype(np.float64)
            core = fin & (ux * ux + uy * uy <= core2)
            cn, s1, s2 = float(core.sum()), s[core].sum(), (s[core] ** 2).sum()
            sy, y1 = (s[core] * v[core]).sum(), v[core].sum()
            det = cn * s2 - s1 * s1
            a, lev = (1.0, 0.0) if not det > 0 else ((cn * sy - s1 * y1) / det, (s2 * y1 - s1 * sy) / det)
            res = v[fin] - a * s[fin] - lev
            nf = float(fin.sum())
            rms = math.sqrt((res * res).sum() / (nf if nf > 0 else 1.0))
            rc = (v - a * np.where(fin, s, 0.0) - lev)[core]
            jx, jy = -a * Sx[b][core], -a * Sy[b][core]
            sxx, sxy, syy = (jx * jx).sum(), (jx * jy).sum(), (jy * jy).sum()
            gxs, gys = (jx * rc).sum(), (jy * rc).sum()
            d2 = sxx * syy - sxy * sxy
            dx, dy = (0.0, 0.0) if not d2 > 0 else ((syy * gxs - sxy * gys) / d2, (sxx * gys - sxy * gxs) / d2)
            out[i] = (a, lev, rms, dx, dy, nf)
        return out

    def build_stack(self, core_frac=0.5):
        """Stack of the interior holes, every hole against it, the stack again without the unusual
        holes, every hole against that (outline_onepass.py:594-625)."""
        inside = self.interior()
        if not inside.any():
            raise ValueError(f"no hole lies {self.R + 1} px or more inside the image (its window): nothing to stack")
        self.core = max(2.0, core_frac * self.spacing)
        stack, _ = self.stack_of(inside)
        o = self.versus(stack, self.core)
        med = float(np.median(o[inside, 0]))
        a = np.maximum(o[:, 0] / med, 1e-3)
        typical = float(np.median(o[inside, 2]))
        use = inside & (a > 0.4) & (a < 2.5) & (o[:, 2] < 2.0 * typical) & (o[:, 5] > 0)
        if not use.any():
            raise ValueError("no hole of usual brightness and leftover to stack")
        self.a = a
        self.stack, self.wsum = self.stack_of(use)
        o = self.versus(self.stack, self.core)
        self.a_rel = np.maximum(o[:, 0], 1e-3)
        self.lev_rel = o[:, 1]
        self.rms = o[:, 2]
        self.noise = float(np.median(self.rms[use]))
        self.use = use
        self.first_use, self.first_a = use.copy(), a.copy()

    # pair map
    def pair_hist(self, sf):
        """The 16 sub-pixel classes' neighbour maps (U^2, Ng, Ng) before the FFT: every stacked hole
        i adds g_i a_j at c_j - c_i - s_i of every hole j within the square of half side `reach`
        (j = i included) as cubic B-spline taps, Ug per px, wrapped (outline_onepass.py:184-224)."""
        U, Ug = self.U, self.pair_U
        Ng = self.L * Ug
        v = -self.c * U
        nb = np.rint(v)
        q = np.mod(nb, U).astype(np.int64)
        s = (v - nb) / U
        cls = q[:, 1] * U + q[:, 0]
        sw = self.use.astype(np.float64)
        wq = np.bincount(cls, weights=sw, minlength=U * U)
        g = np.where(sw > 0, sw / np.maximum(wq[cls], 1e-300), 0.0)
        amp = self.a / float(self.a[self.use].mean())
        hist = np.zeros((U * U, Ng * Ng))
        cs = self.cells.cs
        r = int(math.ceil(self.reach / cs))
        for i in np.flatnonzero(g != 0):
            x, y = self.c[i]
            js = self.cells.members(int(math.floor(x / cs)), int(math.floor(y / cs)), r)
            dx, dy = self.c[js, 0] - x, self.c[js, 1] - y
            near = (np.abs(dx) <= self.reach) & (np.abs(dy) <= self.reach)
            js, dx, dy = js[near], dx[near], dy[near]
            wgt = g[i] * amp[js]
            pxf = (dx - s[i, 0]) * Ug + Ng // 2
            pyf = (dy - s[i, 1]) * Ug + Ng // 2
            fx, fy = np.floor(pxf), np.floor(pyf)
            wx, wy = bspline_weights(pxf - fx), bspline_weights(pyf - fy)
            x0, y0 = fx.astype(np.int64) - 1, fy.astype(np.int64) - 1
            for qq in range(4):
                ry = np.mod(y0 + qq, Ng)
                for pp in range(4):
                    rx = np.mod(x0 + pp, Ng)
                    np.add.at(hist[cls[i]], ry * Ng + rx, wgt * wy[:, qq] * wx[:, pp])
        self.holes_per_class = wq
        return hist.reshape(U * U, Ng, Ng)

    def pair_map(self, sf):
        """Gamma_q (U, U, nky, nkx) on the fitter's band (outline_onepass.py:663-698)."""
        Ug = self.pair_U
        Ng = self.L * Ug
        spec = np.fft.rfft2(self.pair_hist(sf))
        ky = np.arange(-sf.kmax, sf.kmax + 1)
        spec = spec[:, np.mod(ky, Ng), :sf.kmax + 1]
        fy, fx = np.meshgrid(ky / self.L, np.arange(sf.kmax + 1) / self.L, indexing="ij")
        centre = np.exp(-2j * np.pi * (fx + fy) * ((Ng // 2) / Ug))
        fix = np.sinc(fx / Ug) ** 4 * np.sinc(fy / Ug) ** 4
        return (spec / (centre * fix)[None]).reshape(self.U, self.U, sf.nky, sf.nkx)

    def stack_data(self, sf):
        """Stack and weights (1 / variance of a bin's mean) by class (outline_onepass.py:700-706)."""
        y, mask = sf.to_classes(np.where(np.isfinite(self.stack), self.stack, 0.0))
        w, _ = sf.to_classes(self.wsum)
        return y, np.where(mask, w, 0.0) / self.noise ** 2

    def start_theta(self):
        """Round first guess from the stack's second moment (outline_onepass.py:708-728)."""
        st = self.stack
        n = np.arange(self.Ms) - self.Ms // 2
        rr = np.hypot(n[None, :], n[:, None]) / self.U
        ok = np.isfinite(st)
        if not ok.any():
            return self.t.start(0.6, 1.0)
        base = np.percentile(st[ok], 10)
        core = ok & (rr <= 0.6 * self.spacing)
        light = np.clip(np.where(ok, st, 0.0) - base, 0.0, None)
        total = float(light[core].sum())
        if total <= 0:
            return self.t.start(0.6, 1.0)
        sigma2 = float((light[core] * rr[core] ** 2).sum()) / total / 2.0
        fw = 2.355 * math.sqrt(max(sigma2, 0.0))
        return self.t.start(max(0.5 * math.sqrt(max(fw ** 2 - 2.355 ** 2, 0.25)), 0.6), 1.0)

    # hole pass
    def set_template(self, theta, scale, const, far_box=128, far_ramp=2.0, tU=8):
        """Near templates (T0, dT/d delta, Tx, Ty, (1/2) d2T/d delta2) of theta on a fine grid,
        their reach and render half width, the far kernel, every hole's brightness and level in
        the model's units (outline_onepass.py:733-793)."""
        self.theta = np.asarray(theta, float)
        tbox = max(32, int(8 * math.ceil(2 * (self.R + 6) / 8)))
        band = Band(box_px=tbox, f_max=0.75, upsample=tU)
        h = 0.1
        hat = np.stack([analytic_hat(self.t, self.theta, band.fx, band.fy, 256, d) for d in (0.0, h, -h)])
        t0g, tp, tm = band.to_grid(hat)
        td = (tp - tm) / (2 * h)
        t2 = (tp + tm - 2 * t0g) / (2 * h * h)
        N = band.N
        prof = np.abs(t0g[N // 2])
        above = np.nonzero(prof > 1e-3 * prof.max())[0]
        reach_px = float(above[-1] - above[0] + 1) / (2 * tU)
        Rr = min(int(math.ceil(reach_px)) + 1, N // (2 * tU) - 2, self.R + 3)
        xg = (np.arange(N) - N // 2) / tU
        rr = np.hypot(xg[None, :], xg[:, None])
        r0 = float(Rr)
        wn = 0.5 * (1 + np.cos(np.pi * np.clip((rr - r0) / far_ramp, 0.0, 1.0)))
        t0n, tdn, t2n = t0g * wn, td * wn, t2 * wn
        Rr = min(int(math.ceil(r0 + far_ramp)) + 1, N // (2 * tU) - 2)
        f1 = np.fft.fftfreq(N, d=1.0 / tU)
        sp = np.fft.fft2(np.fft.ifftshift(t0n))
        tx = np.fft.fftshift(np.fft.ifft2(sp * (2j * np.pi * f1[None, :])).real)
        ty = np.fft.fftshift(np.fft.ifft2(sp * (2j * np.pi * f1[:, None])).real)
        f32 = lambda v: v.astype(np.float32).astype(np.float64)  # noqa: E731  (the stored FP32 grids)
        g5 = np.ascontiguousarray(np.stack([f32(t0n), f32(tdn), f32(tx), f32(ty), f32(t2n)], axis=-1))
        fb = Band(box_px=far_box, f_max=0.49, upsample=1)
        b0 = fb.to_grid(analytic_hat(self.t, self.theta, fb.fx, fb.fy, 128, 0.0))
        xb = np.arange(far_box) - far_box // 2
        rb = np.hypot(xb[None, :], xb[:, None])
        wf = 1.0 - 0.5 * (1 + np.cos(np.pi * np.clip((rb - r0) / far_ramp, 0.0, 1.0)))
        self.tmpl = dict(g5=g5, g3=np.ascontiguousarray(g5[..., [0, 1, 4]]), far=(b0 * wf).astype(np.float32),
                         N=N, U=tU, r0=r0, Rr=Rr, r_light=r0 + far_ramp)
        self.a = np.maximum(self.a_rel * scale, 1e-3)
        self.level = self.a_rel * const + self.lev_rel
        self.delta = np.zeros(self.n)
        self.cd0 = self.t.cd_px(self.theta)

    def fit_holes(self, sweeps=CENTRING_SWEEPS):
        """The hole pass with the size held (outline_onepass.py:828-887)."""
        chunks = H.colours(self.c, self.h, self.w, self.spacing)
        out = H.fit_holes(self.img, self.c, self.a, self.level, self.delta, self.tmpl, R=self.R, noise=self.noise,
                          sweeps=sweeps, size=False, cd0=self.cd0, chunks=chunks)
        self.c, self.a, self.level, self.delta = out["c"], out["a"], out["level"], out["delta"]
        order = np.concatenate(out["chunks"])
        self.hole_rms = np.empty(self.n)
        self.hole_rms[order] = out["stat"][:, 2]
        self.passes = out["log"]
        return out

    def good_mask(self):
        """Holes fit for the stack after the hole pass (outline_onepass.py:889-903)."""
        inside = self.interior()
        med = np.median(self.a[inside])
        typical = np.median(self.hole_rms[inside])
        return inside & (self.a > 0.4 * med) & (self.a < 2.5 * med) & (self.hole_rms < 2.0 * typical)

    def restack(self, good, scale):
        """The plain stack again from the hole pass's centres (outline_onepass.py:627-640)."""
        if not good.any():
            raise ValueError("no hole of usual brightness and leftover to stack after the hole pass")
        self.model_a = self.a
        self.a = np.maximum(self.a / scale, 1e-3)
        self.use = good
        self.stack, self.wsum = self.stack_of(good)
        self.a_rel = self.a
        self.noise = float(np.median(self.hole_rms[good]))


def fitter_key(page):
    return (page.blur, page.R, page.U, page.f_max, page.L)


def fitter_for(page, cache):
    """The page fitter of a window size, one per (blur, R, U, f_max, L) in `cache`."""
    key = fitter_key(page)
    if key not in cache:
        cache[key] = ShapeFitter(page.t, box_px=page.L, f_max=page.f_max, U=page.U, R=page.R, n_poly=N_POLY)
    return cache[key]


def page_fits(pages, theta0=None, stages=("round", "super", "full"), grid=True, max_iter=(12, 20, 30), ftol=1e-7,
              cache=None):
    """Page fits of pages grouped by fitter (window size), each group in lockstep
    (outline_onepass.py:940-990). Per page (theta, scale, const, record)."""
    cache = {} if cache is None else cache
    out = [None] * len(pages)
    groups = {}
    for i, pg in enumerate(pages):
        groups.setdefault(fitter_key(pg), []).append(i)
    for members in groups.values():
        sf = fitter_for(pages[members[0]], cache)
        gam, ys, ws, th = [], [], [], []
        for i in members:
            pg = pages[i]
            gam.append(pg.pair_map(sf))
            y, w = pg.stack_data(sf)
            ys.append(y)
            ws.append(w)
            th.append(pg.start_theta() if theta0 is None else theta0[i])
        ys, ws, gam = np.stack(ys), np.stack(ws), np.stack(gam)
        theta, scale, const, cost, info = sf.fit(np.stack(th), gam, ys, ws, stages=stages, grid=grid,
                                                 max_iter=max_iter, ftol=ftol)
        bins = (ws > 0).reshape(len(members), -1).sum(axis=1)
        for k, i in enumerate(members):
            out[i] = (theta[k], float(scale[k]), float(const[k]),
                      dict(chi2=float(cost[k]), bins=int(bins[k]), iterations=int(info[-1][1]),
                           converged=bool(info[-1][2][k])))
    return out


def in_order(pages, blur, fits, cache=None):
    """gauss2: a fit that ends with its second Gaussian the narrower is run again from the
    exchanged labels; the lower cost is kept (outline_onepass.py:993-1022)."""
    for f in fits:
        f[3]["exchanged"] = False
    if blur != "gauss2":
        return fits
    idx = [i for i, f in enumerate(fits) if second_is_narrower(pages[i].t, f[0])]
    if not idx:
        return fits
    again = page_fits([pages[i] for i in idx], theta0=[exchanged(pages[i].t, fits[i][0]) for i in idx],
                      stages=("full",), grid=False, max_iter=(FINAL_ITERATIONS,), cache=cache)
    out = list(fits)
    for i, f in zip(idx, again):
        f[3]["exchanged"] = bool(f[3]["chi2"] < fits[i][3]["chi2"])
        if f[3]["exchanged"]:
            out[i] = f
    return out


def fit_images(images, seeds, blur="gauss2", exchange=True):
    """Steps A to C for a list of images (bright holes) with their seeds (one (N, 2) array each,
    or None for find_seeds) (outline_onepass.py:1038-1119). One dict per image."""
    pages = []
    for k, image in enumerate(images):
        s = find_seeds(np.asarray(image, dtype=np.float32)) if seeds is None or seeds[k] is None else seeds[k]
        pg = Page(image, s, blur=blur)
        pg.build_stack()
        pages.append(pg)
    cache = {}
    first = page_fits(pages, stages=FIRST_STAGES, max_iter=FIRST_ITERATIONS, ftol=FIRST_FTOL, cache=cache)
    for pg, (theta, scale, const, _) in zip(pages, first):
        pg.set_template(theta, scale, const)
        pg.fit_holes()
        pg.restack(pg.good_mask(), scale)
    fits = page_fits(pages, theta0=[f[0] for f in first], stages=("full",), grid=False,
                     max_iter=(FINAL_ITERATIONS,), cache=cache)
    fits = in_order(pages, blur if exchange else None, fits, cache=cache)
    out = []
    for pg, f0, (theta, scale, const, record) in zip(pages, first, fits):
        out.append(dict(centers_xy=pg.c.copy(), brightness=pg.model_a.copy(), level=pg.level.copy(),
                        rms=pg.hole_rms.copy(), good=pg.use.copy(), theta=np.asarray(theta, float), scale=scale,
                        const=const, theta_first=np.asarray(f0[0], float), stack=pg.stack.copy(),
                        stack_weight=pg.wsum.copy(), upsample=pg.U, spacing_px=pg.spacing, window_px=pg.R,
                        pair_reach_px=pg.reach, noise_sd=pg.noise, **record))
    return out


def bspline_weights(t):
    """Cubic B-spline weights of the taps at -1, 0, 1, 2 for fractions t (outline_onepass.py:174-181)."""
    t = np.asarray(t, float)
    u = 1.0 - t
    return np.stack([u * u * u / 6, (3 * t ** 3 - 6 * t * t + 4) / 6, (-3 * t ** 3 + 3 * t * t + 3 * t + 1) / 6,
                     t ** 3 / 6], axis=-1)


def drawn_page(blur="gauss2", seed=5, noise=3.0, shape=(120, 170), pitch=10.0, seed_error=0.15):
    """A test page of the outline model (known answer): holes of one outline and blur on a square
    lattice (+-0.3 px jitter), brightness 900-1100, level 100, N(0, noise^2) noise, drawn with the
    fit_outline template of the blur's box (24 px gauss2, 48 px lens); seeds within +-seed_error
    px. Returns (template, theta, centres, page, seeds)."""
    rng = np.random.default_rng(seed)
    t = Template(upsample=8, size_px=24 if blur == "gauss2" else 48, blur=blur)
    theta = t.pack(2.3, 2.2, 2.4, 0.1, [[0.02, -0.01], [0.01, 0.0]],
                   t.blur_pack(s1x=1.2, s1y=1.1, s2x=2.5, s2y=2.5, w2=0.06, sx=1.1, sy=1.0, cutoff=0.45))
    X, Y = np.meshgrid(np.arange(6, shape[1], pitch), np.arange(6, shape[0], pitch))
    c = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.3, 0.3, (X.size, 2))
    a = rng.uniform(900, 1100, len(c))
    page = render_model_page(t, theta, c, a, 100.0, np.zeros(len(c)), shape, noise=noise, seed=seed)
    seeds = c + rng.uniform(-seed_error, seed_error, c.shape)
    return t, theta, c, page, seeds
