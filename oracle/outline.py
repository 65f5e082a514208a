"""Oracle (test infrastructure only): FP64 NumPy restatement of the page outline fit.

One sharp outline (a superellipse with harmonics) convolved with one blur is shared by every hole
of a page; every hole has its own centre, brightness, level and edge offset. The fit alternates a
hole stage (Gauss-Newton per hole against the complete page model, block Gauss-Seidel over colour
classes) and a shape stage (least squares of the outline and blur on the neighbour-cleaned stacked
average hole). Follows the reference:

  Template (outline x blur x pixel, fine grid) ........ subpx/_outline/model.py:25-267
  stage_map / fit_stages ............................... subpx/_outline/model.py:270-368
  PageFit: colours, windows, set_theta, far light ...... subpx/_outline/page.py:78-414
           initialise, stacks, noise line ............. :420-540
           delta_information, fit_holes ............... :553-729
           cleaned_stack, fit_shape, good_holes ....... :732-803
           run, results ............................... :806-937
  fit_outline, ellipse_widths, outline_rays ............ subpx/outline.py:34-232

The hole stage reuses the hole-pass functions of `oracle.holes` (normal equations, bounded step,
light render; the same rules as the library's fit_holes) with this fit's driver: greedy colours
in chunks of 2000 holes, weights from the noise-variance line, the first sweep's robust scale
from each chunk's median leftover, square render windows, far light from three kernels
(brightness, brightness x offset, brightness x offset^2).

Only tests/, the smoke check and the benchmark's CPU baseline use this module.
"""
from __future__ import annotations

import math

import numpy as np

from . import holes as H

BLUR_FAMILIES = {"gauss2": 24, "lens": 48}
MIN_HOLES = 10
UPSAMPLE = 8


# ----------------------------------------------------------------------------------------- template
def lens_otf(nu, cutoff):
    """Transfer function of a circular pupil with cut-off `cutoff` (cycles / px) (model.py:22-24)."""
    s = np.clip(nu / cutoff, 0.0, 1.0)
    return (2.0 / np.pi) * (np.arccos(s) - s * np.sqrt(1.0 - s * s))


class Template:
    """T = pixel box * blur * coverage(outline moved by delta) on an N x N grid of 1/U px
    (model.py:25-267). theta = [log a, log b, log(p - 1), phi, (c_k, s_k) per harmonic, blur...]."""

    N_TABLE = 8192

    def __init__(self, upsample=8, size_px=24, harmonics=(1, 3), blur="gauss2", aa=4):
        if blur not in ("gauss2", "lens", "gauss1"):
            raise ValueError("blur must be 'gauss2', 'lens' or 'gauss1'")
        self.U, self.N, self.ks, self.blur, self.aa = int(upsample), int(size_px * upsample), tuple(
            int(k) for k in harmonics), str(blur), int(aa)
        n, u, a = self.N, self.U, self.aa
        x = (np.arange(n) - n // 2) / u
        self.xx, self.yy = np.meshgrid(x, x)
        m = n * a
        xa = (np.arange(m) - m // 2 + 0.5) / (u * a) - 0.5 / u
        X, Y = np.meshgrid(xa, xa)
        self.rho_aa = np.hypot(X, Y)
        self.table_angles = -np.pi + 2.0 * np.pi * np.arange(self.N_TABLE) / self.N_TABLE
        self.ang_idx_aa = np.mod(np.rint((np.arctan2(Y, X) + np.pi) / (2.0 * np.pi) * self.N_TABLE).astype(np.int64),
                                 self.N_TABLE)
        f = np.fft.fftfreq(n, d=1.0 / u)
        self.fx, self.fy = f[None, :], f[:, None]
        self.nu = np.hypot(self.fx, self.fy)
        self.pixel_otf = np.sinc(self.fx) * np.sinc(self.fy)
        self.angles = np.linspace(-np.pi, np.pi, 4096, endpoint=False)
        self.edge_fix = 1.0 / (np.sinc(self.fx / u) * np.sinc(self.fy / u) * np.sinc(self.nu / (u * a)))

    @property
    def n_shape(self):
        return 4 + 2 * len(self.ks)

    @property
    def n_blur(self):
        return {"gauss2": 5, "lens": 3, "gauss1": 2}[self.blur]

    @property
    def n_params(self):
        return self.n_shape + self.n_blur

    def pack(self, a, b, p, phi, coeffs, blur_params):
        return np.concatenate([[math.log(a), math.log(b), math.log(p - 1.0), phi],
                               np.asarray(coeffs, float).ravel(), np.asarray(blur_params, float)])

    def blur_pack(self, **kw):
        if self.blur == "gauss2":
            w2 = kw.get("w2", 0.05)
            return [math.log(kw.get("s1x", 1.0)), math.log(kw.get("s1y", 1.0)), math.log(kw.get("s2x", 2.0)),
                    math.log(kw.get("s2y", 2.0)), math.log(w2 / (1.0 - w2))]
        head = [math.log(kw.get("sx", 1.0)), math.log(kw.get("sy", 1.0))]
        return head + [math.log(kw.get("cutoff", 0.45))] if self.blur == "lens" else head

    def unpack(self, theta):
        th = np.asarray(theta, float)
        return (math.exp(th[0]), math.exp(th[1]), 1.0 + math.exp(th[2]), th[3],
                th[4:self.n_shape].reshape(-1, 2), th[self.n_shape:])

    def start(self, half_width_px, blur_sigma_px=1.0):
        s = blur_sigma_px
        return self.pack(half_width_px, half_width_px, 2.0, 0.0, np.zeros((len(self.ks), 2)),
                         self.blur_pack(s1x=s, s1y=s, s2x=2 * s, s2y=2 * s, sx=s, sy=s))

    def bounds(self):
        lo, hi = np.full(self.n_params, -np.inf), np.full(self.n_params, np.inf)
        lo[0:2], hi[0:2] = math.log(0.2), math.log(8.0)
        lo[2], hi[2] = math.log(0.3), math.log(30.0)
        lo[3], hi[3] = -np.pi / 4, np.pi / 4
        lo[4:self.n_shape], hi[4:self.n_shape] = -0.3, 0.3
        j = self.n_shape
        if self.blur == "gauss2":
            lo[j:j + 4], hi[j:j + 4] = math.log(0.15), math.log(4.0)
            lo[j + 4], hi[j + 4] = math.log(0.005 / 0.995), math.log(0.6 / 0.4)
        elif self.blur == "lens":
            lo[j:j + 2], hi[j:j + 2] = math.log(0.1), math.log(4.0)
            lo[j + 2], hi[j + 2] = math.log(0.2), math.log(8.0)
        else:
            lo[j:j + 2], hi[j:j + 2] = math.log(0.15), math.log(4.0)
        return lo, hi

    # outline
    def radius_at(self, theta, angles):
        """r(angle) and dr / d angle in px (model.py:130-147)."""
        a, b, p, phi, coeffs, _ = self.unpack(theta)
        t = np.asarray(angles, float) - phi
        c, s = np.cos(t), np.sin(t)
        ua, ub = np.abs(c / a), np.abs(s / b)
        base = (ua ** p + ub ** p) ** (-1.0 / p)
        dS = p * ua ** (p - 1) * np.sign(c) * (-s / a) + p * ub ** (p - 1) * np.sign(s) * (c / b)
        dbase = -(1.0 / p) * base ** (1 + p) * dS
        ang = np.asarray(angles, float)
        mod, dmod = np.ones_like(ang), np.zeros_like(ang)
        for i, k in enumerate(self.ks):
            mod = mod + coeffs[i, 0] * np.cos(k * ang) + coeffs[i, 1] * np.sin(k * ang)
            dmod = dmod - coeffs[i, 0] * k * np.sin(k * ang) + coeffs[i, 1] * k * np.cos(k * ang)
        return base * mod, dbase * mod + base * dmod

    def indicator(self, theta, delta=0.0):
        """Coverage of each fine pixel by the outline moved delta px along its normal: a one-sample
        linear ramp across the edge on the aa x aa sub-grid, averaged (model.py:149-159)."""
        r, dr = self.radius_at(theta, self.table_angles)
        edge = r + delta * np.hypot(r, dr) / r if delta else r
        inside = np.clip(0.5 + (edge[self.ang_idx_aa] - self.rho_aa) * (self.U * self.aa), 0.0, 1.0)
        return inside.reshape(self.N, self.aa, self.N, self.aa).mean(axis=(1, 3))

    # blur
    def otf(self, theta):
        bp = self.unpack(theta)[5]
        if self.blur == "gauss2":
            s1x, s1y, s2x, s2y = np.exp(bp[:4])
            w2 = 1.0 / (1.0 + math.exp(-bp[4]))
            g1 = np.exp(-2 * np.pi ** 2 * (s1x ** 2 * self.fx ** 2 + s1y ** 2 * self.fy ** 2))
            g2 = np.exp(-2 * np.pi ** 2 * (s2x ** 2 * self.fx ** 2 + s2y ** 2 * self.fy ** 2))
            return (1 - w2) * g1 + w2 * g2
        sx, sy = np.exp(bp[:2])
        g = np.exp(-2 * np.pi ** 2 * (sx ** 2 * self.fx ** 2 + sy ** 2 * self.fy ** 2))
        return g * lens_otf(self.nu, math.exp(bp[2])) if self.blur == "lens" else g

    def blur_kernel(self, theta):
        k = np.fft.fftshift(np.fft.ifft2(self.otf(theta)).real)
        return k / k.sum()

    def blur_fwhm_px(self, theta):
        k = self.blur_kernel(theta)
        c = self.N // 2
        out = []
        for prof in (k[c], k[:, c]):
            half = prof[c] / 2
            j = c
            while j + 1 < len(prof) and prof[j + 1] > half:
                j += 1
            out.append(2 * (j - c + (prof[j] - half) / max(prof[j] - prof[j + 1], 1e-12)) / self.U)
        return tuple(out)

    def blur_sigma_px(self, theta):
        k = self.blur_kernel(theta)
        return math.sqrt((k * self.xx ** 2).sum()), math.sqrt((k * self.yy ** 2).sum())

    # template
    def render(self, theta, delta=0.0, box_px=0.0):
        spec = np.fft.fft2(np.fft.ifftshift(self.indicator(theta, delta))) * self.edge_fix * self.otf(theta) \
            * self.pixel_otf
        if box_px > 0:
            spec = spec * np.sinc(self.fx * box_px) * np.sinc(self.fy * box_px)
        return np.fft.fftshift(np.fft.ifft2(spec).real)

    def render_quad(self, theta, h=0.1):
        t0, tp, tm = self.render(theta), self.render(theta, h), self.render(theta, -h)
        return t0, (tp - tm) / (2 * h), (tp + tm - 2 * t0) / (2 * h * h)

    def derivatives_xy(self, t):
        spec = np.fft.fft2(np.fft.ifftshift(t))
        return (np.fft.fftshift(np.fft.ifft2(spec * (2j * np.pi * self.fx)).real),
                np.fft.fftshift(np.fft.ifft2(spec * (2j * np.pi * self.fy)).real))

    # readings
    def outline_area(self, theta, delta=0.0):
        r, dr = self.radius_at(theta, self.angles)
        if delta:
            r = r + delta * np.hypot(r, dr) / r
        return float(np.pi * np.mean(r ** 2))

    def cd_px(self, theta, delta=0.0):
        return 2.0 * math.sqrt(self.outline_area(theta, delta) / math.pi)

    def extents_px(self, theta):
        r, _ = self.radius_at(theta, np.array([0.0, math.pi, math.pi / 2, -math.pi / 2]))
        return float(r[0] + r[1]), float(r[2] + r[3])

    def centroid_px(self, theta):
        ind = self.indicator(theta)
        s = ind.sum()
        return float((ind * self.xx).sum() / s), float((ind * self.yy).sum() / s)

    def describe(self, theta, pixel_nm=1.0):
        a, b, p, phi, coeffs, _ = self.unpack(theta)
        fw, (ex, ey) = self.blur_fwhm_px(theta), self.extents_px(theta)
        return dict(cd_nm=self.cd_px(theta) * pixel_nm, extent_x_nm=ex * pixel_nm, extent_y_nm=ey * pixel_nm,
                    superellipse=dict(a_nm=a * pixel_nm, b_nm=b * pixel_nm, p=p, phi_deg=math.degrees(phi)),
                    harmonics={f"k{k}": [float(coeffs[i, 0]), float(coeffs[i, 1])] for i, k in enumerate(self.ks)},
                    blur_fwhm_nm=[fw[0] * pixel_nm, fw[1] * pixel_nm],
                    blur_rms_nm=[v * pixel_nm for v in self.blur_sigma_px(theta)], blur_family=self.blur)


# ------------------------------------------------------------------------------------- shape stage
def stage_map(t, stage):
    """Free parameter indices and ties of a stage (model.py:270-290)."""
    j = t.n_shape
    if stage == "round":
        free, ties = [0, j], {1: 0}
        if t.blur == "gauss2":
            free += [j + 2, j + 4]
            ties.update({j + 1: j, j + 3: j + 2})
        elif t.blur == "lens":
            free += [j + 2]
            ties.update({j + 1: j})
        else:
            ties.update({j + 1: j})
        return free, ties
    if stage == "super":
        return [0, 1, 2, 3] + list(range(j, t.n_params)), {}
    return list(range(t.n_params)), {}


def fit_stages(t, render_for, y, sw, theta0, stages=("round", "super", "full"), max_nfev=80, grid=True):
    """Staged bounded least squares of scale * render_for(theta) + const on y with sqrt-weights sw,
    scale and const solved linearly per theta (model.py:293-368). Returns (theta, scale, const, chi2)."""
    from scipy.optimize import least_squares

    lo_f, hi_f = t.bounds()
    theta = np.clip(np.asarray(theta0, float), lo_f + 1e-6, hi_f - 1e-6)
    w = sw * sw
    wsum = float(w.sum())

    def resid(th):
        m = render_for(th)
        mw, yw = float((w * m).sum()) / wsum, float((w * y).sum()) / wsum
        var = float((w * (m - mw) ** 2).sum())
        sc = float((w * (m - mw) * (y - yw)).sum()) / var if var > 0 else 1.0
        co = yw - sc * mw
        return sw * (sc * m + co - y), sc, co

    chi2, scale, const = float("nan"), 1.0, 0.0
    for stage in stages:
        free, ties = stage_map(t, stage)
        base = theta.copy()
        if stage == "round":
            base[1], base[2], base[3] = base[0], 0.0, 0.0
            base[4:t.n_shape] = 0.0
            if grid:
                best, a0, j = None, math.exp(base[0]), t.n_shape
                for a in a0 * np.linspace(0.4, 1.8, 15):
                    for s in np.linspace(0.45, 2.2, 12):
                        q = base.copy()
                        q[0] = q[1] = math.log(a)
                        q[j] = q[j + 1] = math.log(s)
                        if t.blur == "gauss2":
                            q[j + 2] = q[j + 3] = math.log(2 * s)
                            q[j + 4] = math.log(0.05 / 0.95)
                        elif t.blur == "lens":
                            q[j + 2] = math.log(0.5)
                        q = np.clip(q, lo_f + 1e-6, hi_f - 1e-6)
                        chi = float(np.mean(resid(q)[0] ** 2))
                        if best is None or chi < best[0]:
                            best = (chi, q)
                base = best[1]

        def expand(q, base=base, free=free, ties=ties):
            th = base.copy()
            th[free] = q
            for k, src in ties.items():
                th[k] = th[src]
            return th

        lo, hi = lo_f[free], hi_f[free]
        r = least_squares(lambda q: resid(expand(q))[0], np.clip(base[free], lo + 1e-6, hi - 1e-6), bounds=(lo, hi),
                          method="trf", x_scale="jac", xtol=1e-9, ftol=1e-10, max_nfev=max_nfev)
        theta = expand(r.x)
        _, scale, const = resid(theta)
        chi2 = float(np.mean(r.fun ** 2))
    return theta, scale, const, chi2


# ---------------------------------------------------------------------------------------- readings
def ellipse_widths(t, theta, delta):
    """x / y widths of the equal-area ellipse with the offset outline's x / y spread (outline.py:196-217)."""
    delta = np.asarray(delta, float)
    r, dr = t.radius_at(theta, t.angles)
    g = np.hypot(r, dr) / r
    c, s = np.cos(t.angles), np.sin(t.angles)

    def integral(k, weight):
        return sum(math.comb(k, m) * 2 * np.pi * np.mean(r ** (k - m) * g ** m * weight) * delta ** m
                   for m in range(k + 1))

    area = integral(2, 1.0) / 2
    mx, my = integral(3, c) / 3, integral(3, s) / 3
    vx = integral(4, c * c) / 4 / area - (mx / area) ** 2
    vy = integral(4, s * s) / 4 / area - (my / area) ** 2
    cd = 2 * np.sqrt(np.maximum(area, 1e-12) / np.pi)
    ratio = np.sqrt(np.sqrt(np.maximum(vx, 1e-12) / np.maximum(vy, 1e-12)))
    return cd * ratio, cd / ratio


def outline_rays(t, theta, delta, n_rays=128):
    """Distance from the outline's centroid to the outline at angles 2 pi j / n, offset delta along
    the normal to first order (outline.py:220-251)."""
    cx, cy = t.centroid_px(theta)
    phi = 2 * np.pi * np.arange(int(n_rays)) / int(n_rays)
    ex, ey = np.cos(phi), np.sin(phi)
    ang = phi.copy()
    for _ in range(30):
        r, dr = t.radius_at(theta, ang)
        u, v = np.cos(ang), np.sin(ang)
        f = ex * (r * v - cy) - ey * (r * u - cx)
        fp = ex * (dr * v + r * u) - ey * (dr * u - r * v)
        step = f / np.where(np.abs(fp) > 1e-12, fp, 1e-12)
        ang = ang - np.clip(step, -0.5, 0.5)
        if np.max(np.abs(step)) < 1e-13:
            break
    r, dr = t.radius_at(theta, ang)
    u, v = np.cos(ang), np.sin(ang)
    rho0 = ex * (r * u - cx) + ey * (r * v - cy)
    nx, ny = r * u + dr * v, r * v - dr * u
    stretch = np.hypot(nx, ny) / (ex * nx + ey * ny)
    return rho0[None, :] + np.asarray(delta, float)[:, None] * stretch[None, :]


# ------------------------------------------------------------------------------------------- page
class Page:
    """The page fit (page.py:78-937) with a free level per hole, far light on, no plane terms."""

    def __init__(self, image, seeds, template, *, chunk=2000, far_ramp=4.0, far_box=128, far_upsample=2,
                 start_sigma=1.0):
        self.image = np.asarray(image, dtype=np.float64)
        self.h, self.w = self.image.shape
        self.c = np.array(seeds, dtype=np.float64, copy=True).reshape(-1, 2)
        self.n = len(self.c)
        self.template = template
        self.chunk = int(chunk)
        self.spacing = float(np.median(H.nearest_distance(self.c)))
        self.R = max(3, int(math.ceil(0.75 * self.spacing)))
        self.colour = greedy_colours(self.c, 1.5 * self.spacing)
        self.chunks = []
        for k in range(int(self.colour.max()) + 1):
            idx = np.flatnonzero(self.colour == k)
            self.chunks += [idx[s:s + self.chunk] for s in range(0, len(idx), self.chunk)]
        self.far_ramp, self.far_box, self.far_upsample = float(far_ramp), int(far_box), int(far_upsample)
        B = self.far_box
        out_y = max(0.0, -self.c[:, 1].min(), self.c[:, 1].max() - (self.h - 1))
        out_x = max(0.0, -self.c[:, 0].min(), self.c[:, 0].max() - (self.w - 1))
        self.margin = (B // 2 + int(math.ceil(out_y)) + 4, B // 2 + int(math.ceil(out_x)) + 4)
        self.a, self.level, self.delta = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self.delta_sd = np.full(self.n, np.nan)
        self.size_on, self.clip_px, self.var_line, self.rms_typ = True, None, (1.0, 0.0), None
        self.theta, self.start_sigma, self.history, self.converged = None, float(start_sigma), [], False
        self._far_template = None

    def _slices(self):
        return [np.arange(s, min(s + self.chunk, self.n)) for s in range(0, self.n, self.chunk)]

    def _windows(self, idx, R):
        off = np.arange(-R, R + 1)
        oy, ox = np.meshgrid(off, off, indexing="ij")
        r = np.rint(self.c[idx])
        px = r[:, 0, None].astype(np.int64) + ox.ravel()[None, :]
        py = r[:, 1, None].astype(np.int64) + oy.ravel()[None, :]
        valid = (px >= 0) & (px < self.w) & (py >= 0) & (py < self.h)
        flat = np.clip(py, 0, self.h - 1) * self.w + np.clip(px, 0, self.w - 1)
        return flat, valid, px, py

    def _fine(self, px, py, idx):
        t = self.template
        return (px - self.c[idx, 0, None]) * t.U + t.N // 2, (py - self.c[idx, 1, None]) * t.U + t.N // 2

    def _ramp(self, r, r0):
        s = np.clip((r - r0) / self.far_ramp, 0.0, 1.0)
        return 0.5 * (1 + np.cos(np.pi * s))

    # template
    def set_theta(self, theta):
        self.theta = np.asarray(theta, float)
        t = self.template
        t0, td, t2 = t.render_quad(self.theta)
        self.rms_typ = None
        prof = np.abs(t0[t0.shape[0] // 2])
        above = np.flatnonzero(prof > 1e-3 * prof.max())
        reach = (above[-1] - above[0] + 1) / (2 * t.U)
        cap = t.N // (2 * t.U) - 2
        if reach > cap and t.N // t.U < 96:
            self.template = Template(upsample=t.U, size_px=2 * (t.N // t.U), harmonics=t.ks, blur=t.blur, aa=t.aa)
            return self.set_theta(theta)
        Rr = min(int(math.ceil(reach)) + 1, cap, self.R + 3)
        self.t0_full = t0
        r0 = float(Rr)
        self.far_r0 = r0
        wn = self._ramp(np.hypot(t.xx, t.yy), r0)
        t0n, tdn, t2n = t0 * wn, td * wn, t2 * wn
        tx, ty = t.derivatives_xy(t0n)
        self.g5 = np.stack([t0n, tdn, tx, ty, t2n], axis=-1)
        self.g3 = np.stack([t0n, tdn, t2n], axis=-1)
        self.Rr = min(int(math.ceil(r0 + self.far_ramp)) + 1, cap)
        if self._far_template is None:
            self._far_template = Template(upsample=self.far_upsample, size_px=self.far_box, harmonics=t.ks,
                                          blur=t.blur, aa=t.aa)
        ft, B = self._far_template, self.far_box
        k = np.arange(B) * ft.U
        rr = np.hypot(*np.meshgrid(np.arange(B) - B // 2, np.arange(B) - B // 2))
        wf = 1.0 - self._ramp(rr, r0)
        self.far_kernels = [g[np.ix_(k, k)] * wf for g in ft.render_quad(self.theta)]

    def far_field(self):
        """Keys splat of a, a delta, a delta^2 on the padded grid, convolved with the far kernels."""
        B = self.far_box
        my, mx = self.margin
        hp, wp = self.h + 2 * my, self.w + 2 * mx
        x, y = self.c[:, 0] + mx, self.c[:, 1] + my
        ix, iy = np.floor(x), np.floor(y)
        wx, wy = H.keys_weights(x - ix), H.keys_weights(y - iy)
        taps = np.arange(-1, 3)
        cx = ix.astype(np.int64)[:, None] + taps
        cy = iy.astype(np.int64)[:, None] + taps
        ok = (cx[:, 0] >= 0) & (cx[:, -1] < wp) & (cy[:, 0] >= 0) & (cy[:, -1] < hp)
        cx, cy = np.clip(cx, 0, wp - 1), np.clip(cy, 0, hp - 1)
        flat = (cy[:, :, None] * wp + cx[:, None, :]).reshape(self.n, -1)
        wgt = (wy[:, :, None] * wx[:, None, :]).reshape(self.n, -1) * ok[:, None]
        spec = 0.0
        for g, val in zip(self.far_kernels, (self.a, self.a * self.delta, self.a * self.delta ** 2)):
            s = np.bincount(flat.ravel(), weights=(wgt * val[:, None]).ravel(), minlength=hp * wp).reshape(hp, wp)
            pad = np.zeros((hp, wp))
            pad[:B, :B] = g
            spec = spec + np.fft.rfft2(s) * np.fft.rfft2(pad)
        far = np.fft.irfft2(spec, s=(hp, wp))
        oy, ox = my + B // 2, mx + B // 2
        return far[oy:oy + self.h, ox:ox + self.w]

    def render_page(self):
        t = self.template
        page = np.zeros(self.h * self.w)
        for idx in self._slices():
            flat, valid, px, py = self._windows(idx, self.Rr)
            sx, sy = self._fine(px, py, idx)
            v = H.sample(self.g3, sx, sy)
            de = self.delta[idx, None]
            light = self.a[idx, None] * (v[..., 0] + de * (v[..., 1] + de * v[..., 2]))
            page += np.bincount(flat.ravel(), weights=np.where(valid, light, 0.0).ravel(), minlength=page.size)
        return page.reshape(self.h, self.w) + self.far_field()

    # start
    def plain_stack(self):
        t = self.template
        sums, cnt = np.zeros(t.N * t.N), np.zeros(t.N * t.N)
        for idx in self._slices():
            flat, valid, px, py = self._windows(idx, self.R)
            sx, sy = self._fine(px, py, idx)
            bx, by = np.rint(sx).astype(np.int64), np.rint(sy).astype(np.int64)
            ok = valid & (bx >= 0) & (bx < t.N) & (by >= 0) & (by < t.N)
            b = (by * t.N + bx)[ok]
            sums += np.bincount(b, weights=self.image.ravel()[flat][ok], minlength=t.N * t.N)
            cnt += np.bincount(b, minlength=t.N * t.N)
        with np.errstate(invalid="ignore", divide="ignore"):
            return (sums / cnt).reshape(t.N, t.N)

    def initialise(self):
        if self.theta is None:
            stack = self.plain_stack()
            mid = stack.shape[0] // 2
            prof = stack[mid]
            with np.errstate(invalid="ignore"):
                base = np.nanmin(prof)
            half = base + 0.5 * (prof[mid] - base)
            lft = rgt = mid
            while lft > 0 and np.isfinite(prof[lft - 1]) and prof[lft - 1] >= half:
                lft -= 1
            while rgt + 1 < len(prof) and np.isfinite(prof[rgt + 1]) and prof[rgt + 1] >= half:
                rgt += 1
            fw = (rgt - lft + 1) / self.template.U
            hw = max(0.5 * math.sqrt(max(fw ** 2 - (2.355 * self.start_sigma) ** 2, 0.25)), 0.6)
            self.set_theta(self.template.start(hw, self.start_sigma))
        page = np.zeros((self.h, self.w))
        for _ in range(2):
            new_a, new_l = np.zeros(self.n), np.zeros(self.n)
            for idx in self._slices():
                flat, valid, px, py = self._windows(idx, self.R)
                sx, sy = self._fine(px, py, idx)
                v0 = H.sample(self.g3, sx, sy)[..., 0]
                y = np.where(valid, self.image.ravel()[flat] - (page.ravel()[flat] - self.a[idx, None] * v0), 0.0)
                wv = valid.astype(np.float64)
                s11, s12, s22 = (wv * v0 * v0).sum(1), (wv * v0).sum(1), wv.sum(1)
                r1, r2 = (wv * v0 * y).sum(1), (wv * y).sum(1)
                det = s11 * s22 - s12 * s12
                new_a[idx] = (r1 * s22 - r2 * s12) / det
                new_l[idx] = (s11 * r2 - s12 * r1) / det
            self.a, self.level = np.maximum(new_a, 1e-3), new_l
            page = self.render_page()

    # noise and weights
    def fit_noise(self, page):
        res_all, lev_all = [], []
        for idx in self._slices()[:3]:
            flat, valid, _, _ = self._windows(idx, self.R)
            model = page.ravel()[flat] + self.level[idx, None]
            res_all.append((self.image.ravel()[flat] - model)[valid])
            lev_all.append(model[valid])
        res, lev = np.concatenate(res_all), np.concatenate(lev_all)
        edges = np.percentile(lev, np.linspace(0, 100, 13))
        xs, ys = [], []
        for lo, hi in zip(edges[:-1], edges[1:]):
            s = (lev >= lo) & (lev < hi)
            if s.sum() > 200:
                r = res[s]
                xs.append(lev[s].mean())
                ys.append((1.4826 * np.median(np.abs(r - np.median(r)))) ** 2)
        xs, ys = np.asarray(xs), np.asarray(ys)
        if len(xs) >= 3:
            v1, v0 = np.polyfit(xs, ys, 1)
            v1, v0 = max(v1, 0.0), max(v0, 0.05 * ys.mean())
        else:
            v0, v1 = max(float(ys.mean()), 1e-3), 0.0
        self.var_line = (float(v0), float(v1))

    def _normal(self, page, idx, rscale, size):
        return H.hole_normal(self.image, page, self.g5, idx, self.c, self.a, self.level, self.delta,
                             self.template.U, self.R, self.var_line[0], self.var_line[1], rscale, size)

    def delta_information(self):
        page = self.render_page()
        for idx in self.chunks:
            acc = self._normal(page, idx, 0.0, True)
            for i, h in enumerate(idx):
                A = np.zeros((5, 5))
                q = 0
                for r in range(5):
                    for s in range(r, 5):
                        A[r, s] = A[s, r] = acc[i, q]
                        q += 1
                A += np.eye(5) * (1e-9 * np.trace(A) / 5)
                try:
                    self.delta_sd[h] = math.sqrt(max(np.linalg.inv(A)[4, 4], 0.0))
                except np.linalg.LinAlgError:
                    self.delta_sd[h] = np.nan
        return float(np.nanmedian(self.delta_sd)), 0.5 * self.template.cd_px(self.theta)

    # hole stage
    def fit_holes(self, iterations=5, size=True, max_move=0.25, stop_px=0.002, robust_c=3.0):
        t = self.template
        max_delta = 0.3 * 0.5 * t.cd_px(self.theta)
        self.clip_px = float(max_delta)
        U, Rr = t.U, self.Rr
        rl2 = float(4 * (Rr + 1) ** 2)  # the whole square window
        page = np.zeros((self.h, self.w))
        caches = [np.zeros((len(idx), (2 * Rr + 1) ** 2), dtype=np.float32) for idx in self.chunks]
        c_olds = []
        for idx, cache in zip(self.chunks, caches):
            c_olds.append(self.c[idx].copy())
            H.render_scatter(page, cache, self.g3, idx, c_olds[-1], self.c, self.a, self.delta, U, Rr, rl2)
        far = self.far_field()
        page += far
        if not self.history:
            self.fit_noise(page)
        a_med = float(np.median(self.a))
        for _ in range(iterations):
            moves, dmoves, rms_all = [], [], []
            for j, idx in enumerate(self.chunks):
                s = self.rms_typ
                if s is None:
                    o = self._normal(page, idx, 0.0, False)
                    s = float(np.median(np.sqrt(o[:, 20] / np.maximum(o[:, 21], 1.0))))
                acc = self._normal(page, idx, robust_c * max(s, 1e-6) if robust_c else 0.0, size)
                st, c_olds[j] = H.hole_update(acc, idx, self.c, self.a, self.level, self.delta, size, max_move,
                                              max_delta, a_med, 1.0)
                if size:
                    self.delta_sd[idx] = st[:, 3]
                    dmoves.append(st[:, 1])
                moves.append(st[:, 0])
                rms_all.append(st[:, 2])
                H.render_scatter(page, caches[j], self.g3, idx, c_olds[j], self.c, self.a, self.delta, U, Rr, rl2)
            new_far = self.far_field()
            page += new_far - far
            far = new_far
            self.rms_typ = float(np.median(np.concatenate(rms_all)))
            m99 = float(np.percentile(np.concatenate(moves), 99))
            d99 = float(np.percentile(np.concatenate(dmoves), 99)) if dmoves else 0.0
            if m99 < stop_px and d99 < stop_px:
                break
        return page

    # shape stage
    def window_rms(self, page):
        o = self._normal(page, np.arange(self.n), 0.0, False)
        return np.sqrt(o[:, 20] / np.maximum(o[:, 21], 1.0))

    def inside(self):
        c = self.c
        return ((c[:, 0] >= self.R + 1) & (c[:, 0] <= self.w - 2 - self.R) & (c[:, 1] >= self.R + 1)
                & (c[:, 1] <= self.h - 2 - self.R))

    def good_holes(self, page):
        inside = self.inside()
        med = np.median(self.a[inside]) if inside.any() else np.median(self.a)
        dmed = np.median(self.delta[inside] if inside.any() else self.delta)
        usual = (self.a > 0.4 * med) & (self.a < 2.5 * med) & (np.abs(self.delta - dmed) < 0.6)
        rms = self.window_rms(page)
        typical = np.median(rms[inside]) if inside.any() else np.median(rms)
        return inside & usual & (rms < 2.0 * typical), rms

    def cleaned_stack(self, page, good, radius_px=None):
        t = self.template
        sums, wsum = np.zeros(t.N * t.N), np.zeros(t.N * t.N)
        v0_, v1_ = self.var_line
        for idx in self._slices():
            g = good[idx]
            if not g.any():
                continue
            flat, valid, px, py = self._windows(idx, self.R)
            sx, sy = self._fine(px, py, idx)
            v = H.sample(self.g3, sx, sy)
            a, de = self.a[idx, None], self.delta[idx, None]
            own = a * (v[..., 0] + de * (v[..., 1] + de * v[..., 2]))
            model = page.ravel()[flat] + self.level[idx, None]
            cleaned = (self.image.ravel()[flat] - model + own) / a
            wgt = a * a / (v0_ + v1_ * np.maximum(model, 0.0))
            bx, by = np.rint(sx).astype(np.int64), np.rint(sy).astype(np.int64)
            ok = valid & g[:, None] & (bx >= 0) & (bx < t.N) & (by >= 0) & (by < t.N)
            if radius_px is not None:
                ok &= np.hypot(sx - t.N // 2, sy - t.N // 2) <= radius_px * t.U
            b = (by * t.N + bx)[ok]
            sums += np.bincount(b, weights=(cleaned * wgt)[ok], minlength=t.N * t.N)
            wsum += np.bincount(b, weights=wgt[ok], minlength=t.N * t.N)
        with np.errstate(invalid="ignore", divide="ignore"):
            stack = np.where(wsum > 0, sums / wsum, np.nan)
        return stack.reshape(t.N, t.N), wsum.reshape(t.N, t.N)

    def fit_shape(self, stack, weight, stages):
        t = self.template
        ok = np.isfinite(stack) & (weight > 0)
        return fit_stages(t, lambda th: t.render(th, box_px=1.0 / t.U)[ok], stack[ok], np.sqrt(weight[ok]),
                          self.theta, stages=stages)

    def run(self, rounds=14, iterations=5, relax=0.5, tol_px=0.08 / 21.0, release_px=0.3 / 21.0,
            per_hole_size="auto"):
        self.initialise()
        all_stages = ("round", "super", "full")
        level, chi2_prev = 0, None
        self.size_on = per_hole_size is True
        for rnd in range(rounds):
            first = rnd == 0
            if rnd == 1 and per_hole_size == "auto":
                sd, radius = self.delta_information()
                self.size_on = sd < 0.3 * radius
            page = self.fit_holes(iterations=iterations + 3 if rnd < 3 else max(3, iterations // 2 + 1),
                                  size=self.size_on and not first)
            self.fit_noise(page)
            good, rms = self.good_holes(page)
            stack, weight = self.cleaned_stack(page, good, radius_px=0.55 * self.spacing if first else None)
            stages = all_stages[:level + 1]
            proposal, scale, const, chi2 = self.fit_shape(stack, weight, stages)
            t = self.template
            change_px = t.cd_px(proposal) - t.cd_px(self.theta)
            kappa = 1.0 if first else relax
            self.a = self.a * scale ** kappa
            self.set_theta(self.theta + kappa * (proposal - self.theta))
            self.history.append(dict(round=rnd + 1, stages=list(stages), chi2=chi2, change_px=change_px,
                                     scale=scale, const=const, good=int(good.sum())))
            settled = chi2_prev is not None and abs(chi2 - chi2_prev) < 0.05 * chi2_prev
            chi2_prev = chi2
            if not first and settled and abs(change_px) < release_px and level < len(all_stages) - 1:
                level += 1
                chi2_prev = None
                continue
            if level == len(all_stages) - 1 and settled and abs(change_px) < tol_px:
                self.converged = True
                break
        page = self.fit_holes(iterations=2 * iterations, size=self.size_on)
        if not self.size_on:
            self.delta_information()
        good, rms = self.good_holes(page)
        stack, weight = self.cleaned_stack(page, good)
        out = self.results(good, rms)
        out.update(stack=stack, stack_weight=weight, page=page)
        return out

    def results(self, good, rms):
        t = self.template
        cx0, cy0 = t.centroid_px(self.theta)
        dmean = float(np.mean(self.delta[good])) if good.any() else 0.0
        delta = self.delta - dmean
        r, dr = t.radius_at(self.theta, t.angles)
        gr = np.hypot(r, dr) / r
        area0, perim, quad = math.pi * np.mean(r * r), 2 * math.pi * np.mean(r * gr), math.pi * np.mean(gr * gr)
        area = area0 + perim * delta + quad * delta ** 2
        return dict(centers_xy=self.c + np.array([cx0, cy0]), brightness=self.a.copy(), level=self.level.copy(),
                    edge_offset_px=delta, edge_offset_raw=self.delta.copy(),
                    cd_px=2.0 * np.sqrt(np.maximum(area, 1e-6) / math.pi), cd_outline_px=t.cd_px(self.theta),
                    cd_sd_px=self.delta_sd * (perim + 2 * quad * delta) / np.sqrt(math.pi * np.maximum(area, 1e-6)),
                    theta=self.theta.copy(), per_hole_size=self.size_on, clip_px=self.clip_px,
                    at_clip=np.abs(self.delta) >= 0.99 * (self.clip_px or np.inf), good=good, rms=rms,
                    var_line=self.var_line, history=self.history, converged=self.converged,
                    spacing_px=self.spacing, window_px=self.R)


def greedy_colours(c, reach):
    """Holes within `reach` of each other get different colours; each hole in index order takes the
    smallest colour none of its neighbours holds (page.py:185-199)."""
    d2 = (c[:, None, 0] - c[None, :, 0]) ** 2 + (c[:, None, 1] - c[None, :, 1]) ** 2
    near = d2 <= reach * reach
    colour = np.full(len(c), -1, dtype=np.int64)
    for i in range(len(c)):
        used = set(colour[near[i]].tolist())
        k = 0
        while k in used:
            k += 1
        colour[i] = k
    return colour


def fit_outline(image, seeds, *, blur="gauss2", rounds=14, invert=False, n_rays=128, per_hole_size="auto",
                template=None):
    """The page outline fit (outline.py:34-170) in pixels. Returns a dict of the result's fields."""
    img = np.asarray(image)
    seeds = np.asarray(seeds, dtype=np.float64).reshape(-1, 2)
    signal = img.astype(np.float64)
    top = float(signal.max())
    if invert:
        signal = top - signal
    t = template if template is not None else Template(upsample=UPSAMPLE, size_px=BLUR_FAMILIES[blur], blur=blur)
    pg = Page(signal, seeds, t)
    res = pg.run(rounds=rounds, per_hole_size=per_hole_size)
    t = pg.template
    theta = res["theta"]
    res["cd_x_px"], res["cd_y_px"] = ellipse_widths(t, theta, res["edge_offset_px"])
    res["rays_px"] = outline_rays(t, theta, res["edge_offset_px"], n_rays)
    if invert:
        res["level"] = top - res["level"]
    res["rounds"] = len(pg.history)
    res["template_obj"] = t
    return res


def render_model_page(t, theta, centers, brightness, level, delta, shape, noise=0.0, seed=0):
    """A page of the outline model for known-answer tests: every hole's full template light
    a (T0 + delta T1 + delta^2 T2) sampled at its pixel offsets out to the template box, plus
    the level and N(0, noise^2) pixel noise (FP64)."""
    h, w = shape
    t0, td, t2 = t.render_quad(np.asarray(theta, float))
    g3 = np.stack([t0, td, t2], axis=-1)
    page = np.full((h, w), float(level))
    Rb = t.N // (2 * t.U) - 2
    off = np.arange(-Rb, Rb + 1)
    oy, ox = np.meshgrid(off, off, indexing="ij")
    for (x, y), a, de in zip(np.asarray(centers, float), np.asarray(brightness, float), np.asarray(delta, float)):
        px, py = int(np.rint(x)) + ox.ravel(), int(np.rint(y)) + oy.ravel()
        ok = (px >= 0) & (px < w) & (py >= 0) & (py < h)
        v = H.sample(g3, (px[ok] - x) * t.U + t.N // 2, (py[ok] - y) * t.U + t.N // 2)
        np.add.at(page.ravel(), py[ok] * w + px[ok], a * (v[:, 0] + de * (v[:, 1] + de * v[:, 2])))
    if noise:
        page += np.random.default_rng(seed).normal(0.0, noise, page.shape)
    return page
