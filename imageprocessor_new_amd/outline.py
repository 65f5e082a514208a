"""Page outline fit on the GPU: ``fit_outline`` (reference subpx/outline.py:34-170).

One sharp outline (a superellipse turned by phi, times 1 + harmonics) convolved with one blur
(two centred anisotropic Gaussians, or a Gaussian times a circular pupil's transfer function)
is shared by every hole of a page; each hole has its own centre, brightness, level and edge
offset (a move of the outline along its normal). The fit alternates

- the hole stage: Gauss-Newton of (x, y, a, level[, delta]) per hole against the complete page
  model, one colour class (chunk) at a time: the HIP kernels of the hole pass
  (``spx_hole_step`` / ``spx_hole_render`` / ``spx_hole_sweep``) with this fit's templates,
  a far light of three kernels (``spx_outline_far_splat`` + FFT) and square render windows;
- the shape stage: the neighbour-cleaned stacked average hole (``spx_outline_windows`` gathers,
  one FP64 scatter-add) fitted by bounded least squares over the outline and blur parameters
  (SciPy's trust-region solver on the host, as the reference; every model evaluation renders
  the template on the device: ``spx_outline_coverage`` + rocFFT).

Templates (subpx/_outline/model.py:25-267): T = pixel box * blur * coverage(outline) on an
N x N grid of 1/U px, rendered spectrally; T1, T2 (first and second edge-offset terms) by central
differences of 0.1 px; Tx, Ty by spectral derivatives. The driver follows PageFit
(subpx/_outline/page.py:78-937): greedy colours in chunks of 2000 holes, a noise-variance line
from the leftovers, Cauchy weights at 3 typical leftovers, the staged release round -> super ->
full, relaxation 0.5, box doubling when the light outgrows it, results from the exact area of
the offset outline.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, stream_handle
from .holes import _even_fast
from .types import OutlineFitResult

BLUR_FAMILIES = {"gauss2": 24, "lens": 48}  # blur family -> first template box (px)
MIN_HOLES = 10
UPSAMPLE = 8


def _ptr(t):
    return None if t is None else int(t.data_ptr())


def _off(t, elements: int) -> int:
    """Device address of element `elements` of a contiguous tensor."""
    return int(t.data_ptr()) + int(elements) * t.element_size()


def _np_median(x: torch.Tensor) -> float:
    """NumPy's median (mean of the middle pair) of a 1-D device tensor."""
    v = torch.sort(x.reshape(-1)).values
    n = v.numel()
    return float(0.5 * (v[(n - 1) // 2] + v[n // 2]))


# ------------------------------------------------------------------------------------- template
# The shape and blur parameters as the solver moves them: (name, transform, lower, upper) with
# the bounds in the solver's coordinate. "log": a length or width (log of px); "log1": the
# superellipse exponent p as log(p - 1); "logit": a light share as log(w / (1 - w)); "lin": the
# turn phi (rad) and the harmonic amplitudes. Shape rows come first, then the blur's.
def _shape_rows(ks):
    rows = [("a", "log", math.log(0.2), math.log(8.0)), ("b", "log", math.log(0.2), math.log(8.0)),
            ("p", "log1", math.log(0.3), math.log(30.0)), ("phi", "lin", -math.pi / 4, math.pi / 4)]
    for k in ks:
        rows += [(f"c{k}", "lin", -0.3, 0.3), (f"s{k}", "lin", -0.3, 0.3)]
    return rows


_WIDTH = (math.log(0.15), math.log(4.0))
_BLUR_ROWS = {
    "gauss2": [("s1x", "log") + _WIDTH, ("s1y", "log") + _WIDTH, ("s2x", "log") + _WIDTH, ("s2y", "log") + _WIDTH,
               ("w2", "logit", math.log(0.005 / 0.995), math.log(0.6 / 0.4))],
    "lens": [("sx", "log", math.log(0.1), math.log(4.0)), ("sy", "log", math.log(0.1), math.log(4.0)),
             ("cutoff", "log", math.log(0.2), math.log(8.0))],
    "gauss1": [("sx", "log") + _WIDTH, ("sy", "log") + _WIDTH],
}
# physical defaults of the blur parameters (px, light share, cycles / px)
_BLUR_DEFAULTS = {"s1x": 1.0, "s1y": 1.0, "s2x": 2.0, "s2y": 2.0, "w2": 0.05, "sx": 1.0, "sy": 1.0, "cutoff": 0.45}


def _to_solver(kind, value):
    if kind == "log":
        return math.log(value)
    if kind == "log1":
        return math.log(value - 1.0)
    if kind == "logit":
        return math.log(value / (1.0 - value))
    return float(value)


def _from_solver(kind, u):
    if kind == "log":
        return math.exp(u)
    if kind == "log1":
        return 1.0 + math.exp(u)
    if kind == "logit":
        return 1.0 / (1.0 + math.exp(-u))
    return float(u)


# ------------------------------------------------------------------------------------- template
class OutlineTemplate:
    """The page's hole template on the device (model.py:25-267). theta = [log a, log b,
    log(p - 1), phi, (c_k, s_k) per harmonic k, blur parameters]; lengths in px."""

    N_TABLE = 8192

    def __init__(self, upsample=8, size_px=24, harmonics=(1, 3), blur="gauss2", aa=4, device=None):
        if blur not in ("gauss2", "lens", "gauss1"):
            raise ValueError("blur must be 'gauss2', 'lens' or 'gauss1'")
        self.U, self.N, self.aa = int(upsample), int(size_px * upsample), int(aa)
        self.ks, self.blur = tuple(int(k) for k in harmonics), str(blur)
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        n, u, a = self.N, self.U, self.aa
        # the aa x aa sub-samples of every fine pixel: radius and angle-table index (fixed per grid)
        m = n * a
        xa = (np.arange(m) - m // 2 + 0.5) / (u * a) - 0.5 / u
        X, Y = np.meshgrid(xa, xa)
        self.rho_d = torch.from_numpy(np.hypot(X, Y)).to(dev)
        ang = np.mod(np.rint((np.arctan2(Y, X) + np.pi) / (2.0 * np.pi) * self.N_TABLE), self.N_TABLE)
        self.ang_d = torch.from_numpy(ang.astype(np.int32)).to(dev)
        self.table_angles = -np.pi + 2.0 * np.pi * np.arange(self.N_TABLE) / self.N_TABLE
        self.angles = np.linspace(-np.pi, np.pi, 4096, endpoint=False)
        x = (np.arange(n) - n // 2) / u
        self.xx, self.yy = np.meshgrid(x, x)
        self.xx_d, self.yy_d = torch.from_numpy(self.xx).to(dev), torch.from_numpy(self.yy).to(dev)
        f = torch.from_numpy(np.fft.fftfreq(n, d=1.0 / u)).to(dev)
        self.fx, self.fy = f[None, :], f[:, None]
        self.nu = torch.hypot(self.fx, self.fy)
        self.pixel_otf = torch.sinc(self.fx) * torch.sinc(self.fy)
        self.edge_fix = 1.0 / (torch.sinc(self.fx / u) * torch.sinc(self.fy / u) * torch.sinc(self.nu / (u * a)))

    # parameters: one row per entry of theta (_shape_rows + _BLUR_ROWS)
    @property
    def rows(self):
        return _shape_rows(self.ks) + _BLUR_ROWS[self.blur]

    def index(self, name):
        return [r[0] for r in self.rows].index(name)

    @property
    def n_shape(self):
        return len(_shape_rows(self.ks))

    @property
    def n_blur(self):
        return len(_BLUR_ROWS[self.blur])

    @property
    def n_params(self):
        return self.n_shape + self.n_blur

    def pack(self, a, b, p, phi, coeffs, blur_params):
        """theta from the outline's physical values and a blur vector from blur_pack."""
        shape = [_to_solver(kind, v) for (_, kind, _, _), v in zip(_shape_rows(()), (a, b, p, phi))]
        return np.concatenate([shape, np.asarray(coeffs, float).ravel(), np.asarray(blur_params, float)])

    def blur_pack(self, **physical):
        """The blur part of theta from physical values (missing ones take _BLUR_DEFAULTS)."""
        return [_to_solver(kind, physical.get(name, _BLUR_DEFAULTS[name])) for name, kind, _, _ in _BLUR_ROWS[self.blur]]

    def unpack(self, theta):
        """(a, b, p, phi, harmonic pairs (k, 2), blur part of theta) of a theta."""
        th = np.asarray(theta, float)
        a, b, p, phi = (_from_solver(kind, u) for (_, kind, _, _), u in zip(_shape_rows(()), th[:4]))
        return a, b, p, phi, th[4:self.n_shape].reshape(-1, 2), th[self.n_shape:]

    def start(self, half_width_px, blur_sigma_px=1.0):
        """A round outline of the given half width under a blur of the given width (the second
        Gaussian twice as wide)."""
        s = float(blur_sigma_px)
        return self.pack(half_width_px, half_width_px, 2.0, 0.0, np.zeros((len(self.ks), 2)),
                         self.blur_pack(s1x=s, s1y=s, s2x=2 * s, s2y=2 * s, sx=s, sy=s))

    def bounds(self):
        rows = self.rows
        return np.array([r[2] for r in rows], float), np.array([r[3] for r in rows], float)

    # outline (host: a few thousand angles)
    def radius_at(self, theta, angles):
        """r(angle) and dr / d angle (px) of the outline (model.py:130-147)."""
        a, b, p, phi, coeffs, _ = self.unpack(theta)
        ang = np.asarray(angles, float)
        ct, st = np.cos(ang - phi), np.sin(ang - phi)
        qa, qb = np.abs(ct / a), np.abs(st / b)
        base = (qa ** p + qb ** p) ** (-1.0 / p)
        dbase = -base ** (1 + p) * (qa ** (p - 1) * np.sign(ct) * (-st / a) + qb ** (p - 1) * np.sign(st) * (ct / b))
        mod, dmod = np.ones_like(ang), np.zeros_like(ang)
        for i, k in enumerate(self.ks):
            ck, sk = np.cos(k * ang), np.sin(k * ang)
            mod = mod + coeffs[i, 0] * ck + coeffs[i, 1] * sk
            dmod = dmod + k * (coeffs[i, 1] * ck - coeffs[i, 0] * sk)
        return base * mod, dbase * mod + base * dmod

    def _columns(self, thetas):
        """Per-outline parameters of a batch of thetas (B, n_params) as (B, 1) device columns (one
        host-to-device copy): a, b, p, phi, the harmonic pairs, and the blur's variances / weight."""
        th = np.atleast_2d(np.asarray(thetas, float))
        nk = len(self.ks)
        coeffs = th[:, 4:self.n_shape].reshape(th.shape[0], -1, 2)
        bp = th[:, self.n_shape:]
        cols = [np.exp(th[:, 0]), np.exp(th[:, 1]), 1.0 + np.exp(th[:, 2]), th[:, 3]]
        cols += [coeffs[:, i, 0] for i in range(nk)] + [coeffs[:, i, 1] for i in range(nk)]
        if self.blur == "gauss2":
            s = np.exp(bp[:, :4])
            cols += [s[:, 0] ** 2, s[:, 1] ** 2, s[:, 2] ** 2, s[:, 3] ** 2, 1.0 / (1.0 + np.exp(-bp[:, 4]))]
        else:
            s = np.exp(bp[:, :2])
            cols += [s[:, 0] ** 2, s[:, 1] ** 2] + ([np.exp(bp[:, 2])] if self.blur == "lens" else [])
        m = torch.from_numpy(np.ascontiguousarray(np.stack(cols, axis=1), dtype=np.float64)).to(self.device)
        c = [m[:, j:j + 1] for j in range(m.shape[1])]
        return dict(a=c[0], b=c[1], p=c[2], phi=c[3], c=c[4:4 + nk], s=c[4 + nk:4 + 2 * nk], blur=c[4 + 2 * nk:])

    def edge_tables(self, thetas, delta=0.0, columns=None) -> torch.Tensor:
        """The outlines' radii, moved `delta` px along the normal, on the 8192-entry angle table
        (B, 8192) FP64 on the device: the formulas of radius_at, evaluated where the coverage
        reads them."""
        q = self._columns(thetas) if columns is None else columns
        if not hasattr(self, "_table_d"):
            self._table_d = torch.from_numpy(self.table_angles).to(self.device)[None, :]
        ang = self._table_d
        a, b, p = q["a"], q["b"], q["p"]
        ct, st = torch.cos(ang - q["phi"]), torch.sin(ang - q["phi"])
        qa, qb = torch.abs(ct / a), torch.abs(st / b)
        base = (qa ** p + qb ** p) ** (-1.0 / p)
        mod = torch.ones_like(base)
        for i, k in enumerate(self.ks):
            mod = mod + q["c"][i] * torch.cos(k * ang) + q["s"][i] * torch.sin(k * ang)
        r = base * mod
        if not delta:
            return r
        dbase = -base ** (1 + p) * (qa ** (p - 1) * torch.sign(ct) * (-st / a) + qb ** (p - 1) * torch.sign(st) * (ct / b))
        dmod = torch.zeros_like(base)
        for i, k in enumerate(self.ks):
            dmod = dmod + k * (q["s"][i] * torch.cos(k * ang) - q["c"][i] * torch.sin(k * ang))
        dr = dbase * mod + base * dmod
        return r + delta * torch.hypot(r, dr) / r

    def edge_table(self, theta, delta=0.0) -> torch.Tensor:
        return self.edge_tables(theta, delta)[0]

    def indicator_batch(self, thetas, delta=0.0, columns=None) -> torch.Tensor:
        """Fine-pixel coverage of each outline moved `delta` px along its normal (B, N, N)."""
        edge = self.edge_tables(thetas, delta, columns).contiguous()
        nb = int(edge.shape[0])
        out = torch.empty((nb, self.N, self.N), dtype=torch.float64, device=self.device)
        _lib.call("spx_outline_coverage", _ptr(edge), _ptr(self.rho_d), _ptr(self.ang_d), self.N, self.aa,
                  float(self.U * self.aa), nb, _ptr(out), stream_handle())
        return out

    def indicator(self, theta, delta=0.0):
        """Fine-pixel coverage of the outline moved `delta` px along its normal (device, N x N)."""
        return self.indicator_batch(theta, delta)[0]

    # blur
    def otf_batch(self, thetas, columns=None) -> torch.Tensor:
        """Blur transfer functions (B, N, N) of a batch of thetas."""
        q = self._columns(thetas) if columns is None else columns
        bl = [v[:, :, None] for v in q["blur"]]
        fx2, fy2 = (self.fx * self.fx)[None], (self.fy * self.fy)[None]
        k = -2.0 * math.pi ** 2
        if self.blur == "gauss2":
            g1 = torch.exp(k * (bl[0] * fx2 + bl[1] * fy2))
            g2 = torch.exp(k * (bl[2] * fx2 + bl[3] * fy2))
            return (1.0 - bl[4]) * g1 + bl[4] * g2
        g = torch.exp(k * (bl[0] * fx2 + bl[1] * fy2))
        if self.blur == "lens":
            sc = torch.clamp(self.nu[None] / bl[2], 0.0, 1.0)
            g = g * (2.0 / math.pi) * (torch.acos(sc) - sc * torch.sqrt(1.0 - sc * sc))
        return g

    def otf(self, theta):
        return self.otf_batch(theta)[0]

    def blur_kernel(self, theta) -> np.ndarray:
        k = torch.fft.fftshift(torch.fft.ifft2(self.otf(theta)).real).cpu().numpy()
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
    def render_batch(self, thetas, delta=0.0, box_px=0.0) -> torch.Tensor:
        """T of each theta on the device (B, N, N) FP64: coverage -> spectrum x sharp-edge
        correction x blur x pixel box [x an extra box of box_px] -> back (model.py:200-205)."""
        dims = (-2, -1)
        q = self._columns(thetas)
        spec = torch.fft.fft2(torch.fft.ifftshift(self.indicator_batch(thetas, delta, q), dim=dims))
        factor = (self.edge_fix * self.pixel_otf)[None] * self.otf_batch(thetas, q)
        if box_px > 0:
            factor = factor * (torch.sinc(self.fx * box_px) * torch.sinc(self.fy * box_px))[None]
        return torch.fft.fftshift(torch.fft.ifft2(spec * factor).real, dim=dims)

    def render(self, theta, delta=0.0, box_px=0.0):
        return self.render_batch(theta, delta, box_px)[0]

    def render_quad(self, theta, h=0.1):
        t0, tp, tm = self.render(theta), self.render(theta, h), self.render(theta, -h)
        return t0, (tp - tm) / (2 * h), (tp + tm - 2 * t0) / (2 * h * h)

    def derivatives_xy(self, t):
        spec = torch.fft.fft2(torch.fft.ifftshift(t))
        w = 2j * math.pi
        return (torch.fft.fftshift(torch.fft.ifft2(spec * (w * self.fx)).real),
                torch.fft.fftshift(torch.fft.ifft2(spec * (w * self.fy)).real))

    # readings
    def outline_area(self, theta, delta=0.0):
        r, dr = self.radius_at(theta, self.angles)
        if delta:
            r = r + delta * np.hypot(r, dr) / r
        return float(np.pi * np.mean(r * r))

    def cd_px(self, theta, delta=0.0):
        return 2.0 * math.sqrt(self.outline_area(theta, delta) / math.pi)

    def extents_px(self, theta):
        r, _ = self.radius_at(theta, np.array([0.0, math.pi, math.pi / 2, -math.pi / 2]))
        return float(r[0] + r[1]), float(r[2] + r[3])

    def centroid_px(self, theta):
        ind = self.indicator(theta)
        s = ind.sum()
        return float((ind * self.xx_d).sum() / s), float((ind * self.yy_d).sum() / s)

    def describe(self, theta, pixel_size=1.0):
        a, b, p, phi, coeffs, _ = self.unpack(theta)
        fw, (ex, ey) = self.blur_fwhm_px(theta), self.extents_px(theta)
        return dict(cd_nm=self.cd_px(theta) * pixel_size, extent_x_nm=ex * pixel_size, extent_y_nm=ey * pixel_size,
                    superellipse=dict(a_nm=a * pixel_size, b_nm=b * pixel_size, p=p, phi_deg=math.degrees(phi)),
                    harmonics={f"k{k}": [float(coeffs[i, 0]), float(coeffs[i, 1])] for i, k in enumerate(self.ks)},
                    blur_fwhm_nm=[v * pixel_size for v in fw],
                    blur_rms_nm=[v * pixel_size for v in self.blur_sigma_px(theta)], blur_family=self.blur)


def _level_bins(res: torch.Tensor, lev: torch.Tensor):
    """Twelve level bins between the 0, 8.3, ..., 100th percentiles of lev (lo <= lev < hi);
    per bin with more than 200 pixels the mean level and the squared robust spread
    (1.4826 x median absolute deviation) of res, by NumPy's interpolation rules (page.py:515-524)."""
    m = lev.numel()
    if m == 0:
        return [], []
    lev_sorted, perm = torch.sort(lev)
    res_by_lev = res[perm]
    pos = np.linspace(0, 100, 13) / 100.0 * (m - 1)
    lo_i, hi_i = np.floor(pos).astype(np.int64), np.ceil(pos).astype(np.int64)
    frac = torch.from_numpy(pos - lo_i).to(lev.device)
    vlo, vhi = lev_sorted[torch.from_numpy(lo_i).to(lev.device)], lev_sorted[torch.from_numpy(hi_i).to(lev.device)]
    d = vhi - vlo
    edges = torch.where(frac >= 0.5, vhi - d * (1.0 - frac), vlo + d * frac)
    # bin j is a contiguous run of the sorted levels
    bounds = torch.searchsorted(lev_sorted, edges).cpu().numpy()
    xs, ys = [], []
    for j in range(12):
        b0, b1 = int(bounds[j]), int(bounds[j + 1])
        if b1 - b0 > 200:
            r = res_by_lev[b0:b1]
            xs.append(float(lev_sorted[b0:b1].mean()))
            mad = torch.quantile(torch.abs(r - torch.quantile(r, 0.5)), 0.5)
            ys.append(float((1.4826 * mad) ** 2))
    return xs, ys


# ---------------------------------------------------------------------------------- shape stage
# The shape fit releases the model in three stages. "round": a circle (b tied to a) under an
# isotropic blur (every y width tied to its x width); "super": the superellipse and its turn
# with the whole blur; "full": the harmonics as well. Parameters by name, ties as {copy: source}.
_ROUND_BLUR = {"gauss2": (("s1x", "s2x", "w2"), {"s1y": "s1x", "s2y": "s2x"}),
               "lens": (("sx", "cutoff"), {"sy": "sx"}),
               "gauss1": (("sx",), {"sy": "sx"})}


def _stage_parameters(t, stage):
    """(free theta indices in ascending order, ties {index: source index}) of a stage."""
    names = [r[0] for r in t.rows]
    blur_names = names[t.n_shape:]
    if stage == "round":
        free_blur, blur_ties = _ROUND_BLUR[t.blur]
        free = ("a",) + free_blur
        ties = dict(blur_ties, b="a")
    elif stage == "super":
        free, ties = ("a", "b", "p", "phi") + tuple(blur_names), {}
    else:
        free, ties = tuple(names), {}
    return sorted(names.index(n) for n in free), {names.index(k): names.index(v) for k, v in ties.items()}


def _round_start_grid(t, base):
    """The round stage's start candidates: the base theta at 15 outline sizes (0.4 - 1.8 of its
    half width) times 12 blur widths (0.45 - 2.2 px; the second Gaussian twice as wide with 5 %
    of the light, the pupil cut-off at 0.5 cycles / px), sizes outermost."""
    sizes = math.exp(base[0]) * np.linspace(0.4, 1.8, 15)
    widths = np.linspace(0.45, 2.2, 12)
    pairs = [(math.log(a), math.log(w), math.log(2.0 * w)) for a in sizes for w in widths]
    A, S, S2 = (np.array(v) for v in zip(*pairs))
    cand = np.repeat(np.asarray(base, float)[None], A.size, axis=0)
    idx = t.index
    cand[:, idx("a")] = cand[:, idx("b")] = A
    first = "s1" if t.blur == "gauss2" else "s"
    cand[:, idx(first + "x")] = cand[:, idx(first + "y")] = S
    if t.blur == "gauss2":
        cand[:, idx("s2x")] = cand[:, idx("s2y")] = S2
        cand[:, idx("w2")] = math.log(0.05 / 0.95)
    elif t.blur == "lens":
        cand[:, idx("cutoff")] = math.log(0.5)
    lo, hi = t.bounds()
    return np.clip(cand, lo + 1e-6, hi - 1e-6)


def _fd_steps(x0, f0, lo, hi):
    """Forward-difference steps of SciPy's '2-point' Jacobian (its own step and bound rules,
    scipy.optimize._numdiff), so that a batched Jacobian has the columns SciPy would form."""
    from scipy.optimize import _numdiff

    h = _numdiff._compute_absolute_step(None, x0, f0, "2-point")
    h, _ = _numdiff._adjust_scheme_to_bounds(x0, h, 1, "1-sided", lo, hi)
    return h


def fit_stages(t, render_for, y, sw, theta0, stages=("round", "super", "full"), max_nfev=80, grid=True, log=None,
               batch=48):
    """Bounded least squares of scale * render_for(theta) + const on y (device, sqrt-weights sw)
    in stages, scale and const solved linearly per theta (model.py:293-368). render_for maps a
    batch of thetas (B, n_params) to the device model values (B, K). The solver (SciPy 'trf')
    runs on the host; the start grid and every finite-difference Jacobian are rendered as one
    batch each (forward steps by SciPy's own rules), a residual vector is one copy to the host."""
    from scipy.optimize import least_squares

    lo_all, hi_all = t.bounds()
    theta = np.clip(np.asarray(theta0, float), lo_all + 1e-6, hi_all - 1e-6)
    w = sw * sw
    wsum = w.sum()
    yw = (w * y).sum() / wsum

    def resid_batch(ths):
        ths = np.atleast_2d(ths)
        rows = []
        for s0 in range(0, len(ths), batch):
            m = render_for(ths[s0:s0 + batch])
            mw = (w * m).sum(1, keepdim=True) / wsum
            dm = m - mw
            var = (w * dm * dm).sum(1, keepdim=True)
            cov = (w * dm * (y - yw)).sum(1, keepdim=True)
            sc = torch.where(var > 0, cov / torch.where(var > 0, var, torch.ones_like(var)), torch.ones_like(var))
            co = yw - sc * mw
            rows.append(torch.cat([sw * (sc * m + co - y), sc, co], dim=1))
        host = torch.cat(rows).cpu().numpy()
        return host[:, :-2], host[:, -2], host[:, -1]

    chi2, scale, const = float("nan"), 1.0, 0.0
    for stage in stages:
        free, ties = _stage_parameters(t, stage)
        base = theta.copy()
        if stage == "round":
            # a circle without harmonics, the best of the start grid when asked
            base[t.index("b")] = base[t.index("a")]
            base[t.index("p")] = 0.0
            base[t.index("phi")] = 0.0
            base[4:t.n_shape] = 0.0
            if grid:
                cands = _round_start_grid(t, base)
                chi = np.mean(resid_batch(cands)[0] ** 2, axis=1)
                base = cands[int(np.argmin(chi))]

        def expand(q, base=base, free=free, ties=ties):
            th = np.array(base, copy=True) if q.ndim == 1 else np.repeat(base[None], len(q), axis=0)
            th[..., free] = q
            for k, src in ties.items():
                th[..., k] = th[..., src]
            return th

        lo, hi = lo_all[free], hi_all[free]
        last = {}

        def fun(q):
            r = resid_batch(expand(q))[0][0]
            last["q"], last["f"] = q.copy(), r
            return r

        def jac(q):
            f0 = last["f"] if "q" in last and np.array_equal(last["q"], q) else fun(q)
            h = _fd_steps(q, f0, lo, hi)
            xs = q[None, :] + np.diag(h)
            dx = np.diag(xs) - q
            fx = resid_batch(expand(xs))[0]
            return ((fx - f0[None, :]) / dx[:, None]).T

        r = least_squares(fun, np.clip(base[free], lo + 1e-6, hi - 1e-6), jac=jac, bounds=(lo, hi), method="trf",
                          x_scale="jac", xtol=1e-9, ftol=1e-10, max_nfev=max_nfev)
        theta = expand(r.x)
        _, sc, co = resid_batch(theta)
        scale, const = float(sc[0]), float(co[0])
        chi2 = float(np.mean(r.fun ** 2))
        if log is not None:
            log(f"    shape stage {stage}: chi2 {chi2:.3f}, CD {t.cd_px(theta):.3f} px, scale {scale:.4f}, "
                f"evaluations {r.nfev}")
    return theta, scale, const, chi2


# -------------------------------------------------------------------------------------- readings
def ellipse_widths(t, theta, delta):
    """x / y diameters of the equal-area ellipse with the x / y spread of the outline moved
    delta along its normal (outline.py:196-217): area and moments are polynomials in delta."""
    delta = np.asarray(delta, float)
    r, dr = t.radius_at(theta, t.angles)
    g = np.hypot(r, dr) / r
    cs, sn = np.cos(t.angles), np.sin(t.angles)

    def moment(k, weight):
        # integral over the angle of (r + delta g)^k weight
        return sum(math.comb(k, m) * 2.0 * np.pi * np.mean(r ** (k - m) * g ** m * weight) * delta ** m
                   for m in range(k + 1))

    area = moment(2, 1.0) / 2
    mx, my = moment(3, cs) / 3, moment(3, sn) / 3
    var_x = moment(4, cs * cs) / 4 / area - (mx / area) ** 2
    var_y = moment(4, sn * sn) / 4 / area - (my / area) ** 2
    cd = 2.0 * np.sqrt(np.maximum(area, 1e-12) / np.pi)
    ratio = np.sqrt(np.sqrt(np.maximum(var_x, 1e-12) / np.maximum(var_y, 1e-12)))
    return cd * ratio, cd / ratio


def outline_rays(t, theta, delta, n_rays=128):
    """Each hole's outline as its distance from the outline's centroid at angles 2 pi j / n
    (outline.py:220-251): the page outline solved per ray by Newton on the polar angle, the
    offset delta added along the normal (stretch 1 / cos of the ray-normal angle)."""
    cx, cy = t.centroid_px(theta)
    phi = 2.0 * np.pi * np.arange(int(n_rays)) / int(n_rays)
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


# ------------------------------------------------------------------------------------------ page
def greedy_colours(c: np.ndarray, reach: float) -> np.ndarray:
    """Holes within `reach` of each other get different colours; holes in index order take the
    smallest colour no neighbour holds (page.py:185-199)."""
    from scipy.spatial import cKDTree

    nbrs = cKDTree(c).query_ball_point(c, reach)
    colour = np.full(len(c), -1, dtype=np.int64)
    for i, nb in enumerate(nbrs):
        used = set(colour[nb].tolist())
        k = 0
        while k in used:
            k += 1
        colour[i] = k
    return colour


class OutlinePage:
    """One page through the outline fit (PageFit, page.py:78-937): a free level per hole, the
    far light on, hole state FP64 on the device."""

    def __init__(self, image, seeds, template: OutlineTemplate, *, chunk=2000, far_ramp=4.0, far_box=128,
                 far_upsample=2, start_sigma=1.0, device=None):
        from scipy.spatial import cKDTree

        dev = template.device if device is None else device
        self.device = dev
        img = torch.as_tensor(np.asarray(image, dtype=np.float64) if not isinstance(image, torch.Tensor) else image)
        self.img = img.to(device=dev, dtype=torch.float32).contiguous()
        self.h, self.w = (int(v) for v in self.img.shape)
        c = np.array(seeds, dtype=np.float64, copy=True).reshape(-1, 2)
        self.n = len(c)
        self.template = template
        self.chunk = int(chunk)
        d, _ = cKDTree(c).query(c, k=2)
        self.spacing = float(np.median(d[:, 1]))
        self.R = max(3, int(math.ceil(0.75 * self.spacing)))
        colour = greedy_colours(c, 1.5 * self.spacing)
        chunks = []
        for k in range(int(colour.max()) + 1):
            idx = np.flatnonzero(colour == k)
            chunks += [idx[s:s + self.chunk] for s in range(0, len(idx), self.chunk)]
        self.n_colours = int(colour.max()) + 1
        self.order = torch.from_numpy(np.concatenate(chunks).astype(np.int32)).to(dev)
        self.starts = np.cumsum([0] + [len(x) for x in chunks]).astype(np.int32)
        self.all_idx = torch.arange(self.n, dtype=torch.int32, device=dev)
        self.far_ramp, self.far_box, self.far_upsample = float(far_ramp), int(far_box), int(far_upsample)
        B = self.far_box
        out_y = max(0.0, -c[:, 1].min(), c[:, 1].max() - (self.h - 1))
        out_x = max(0.0, -c[:, 0].min(), c[:, 0].max() - (self.w - 1))
        self.my, self.mx = B // 2 + int(math.ceil(out_y)) + 4, B // 2 + int(math.ceil(out_x)) + 4
        self.hp, self.wp = _even_fast(self.h + 2 * self.my), _even_fast(self.w + 2 * self.mx)
        self.c = torch.from_numpy(c).to(dev)
        z = lambda: torch.zeros(self.n, dtype=torch.float64, device=dev)  # noqa: E731
        self.a, self.level, self.delta = z(), z(), z()
        self.delta_sd = torch.full((self.n,), float("nan"), dtype=torch.float64, device=dev)
        self.size_on, self.clip_px, self.var_line, self.rms_typ = True, None, (1.0, 0.0), None
        self.theta, self.start_sigma, self.history, self.converged = None, float(start_sigma), [], False
        self.log, self.timings = [], {}
        self._far_template = None

    # ---------------------------------------------------------------------------------- gathers
    def windows(self, idx=None, R=None, *, page=None, level=True, outputs=("valid", "data", "model"),
                radius_fine=0.0):
        """Per-window values of the holes idx (device int32; all by default), (n, (2R+1)^2)."""
        idx = self.all_idx if idx is None else idx
        R = self.R if R is None else int(R)
        n, P = int(idx.numel()), (2 * R + 1) ** 2
        dev = self.device
        out = {}
        for name in outputs:
            dt = {"valid": torch.uint8, "bin": torch.int32}.get(name, torch.float64)
            out[name] = torch.empty((n, P), dtype=dt, device=dev)
        t = self.template
        tm = self.tmpl if ("own" in out or "v0" in out) else None
        _lib.call("spx_outline_windows", _ptr(self.img), self.h, self.w, self.w, _ptr(page), _ptr(tm), t.N, t.U, R,
                  _ptr(idx), n, _ptr(self.c), _ptr(self.a), _ptr(self.level) if level else None, _ptr(self.delta),
                  float(radius_fine), _ptr(out.get("bin")), _ptr(out.get("valid")), _ptr(out.get("data")),
                  _ptr(out.get("model")), _ptr(out.get("own")), _ptr(out.get("v0")), stream_handle())
        if "valid" in out:
            out["valid"] = out["valid"].bool()
        return out

    def _ramp(self, r, r0):
        s = torch.clamp((r - r0) / self.far_ramp, 0.0, 1.0)
        return 0.5 * (1.0 + torch.cos(math.pi * s))

    # --------------------------------------------------------------------------------- template
    def set_theta(self, theta):
        """Templates of theta: near parts (inside r0, cosine ramp over far_ramp) for the per-hole
        render, far kernels for the FFT light, box doubling (page.py:301-369)."""
        self.theta = np.asarray(theta, float)
        t = self.template
        t0, td, t2 = t.render_quad(self.theta)
        self.rms_typ = None
        prof = t0[t.N // 2].abs().cpu().numpy()
        above = np.flatnonzero(prof > 1e-3 * prof.max())
        reach = (above[-1] - above[0] + 1) / (2 * t.U)
        cap = t.N // (2 * t.U) - 2
        if reach > cap and t.N // t.U < 96:
            self.log.append(f"  template box {t.N // t.U} -> {2 * (t.N // t.U)} px: the hole's light reaches "
                            f"{reach:.1f} px, past the {cap} px the box can render")
            self.template = OutlineTemplate(upsample=t.U, size_px=2 * (t.N // t.U), harmonics=t.ks, blur=t.blur,
                                            aa=t.aa, device=self.device)
            return self.set_theta(theta)
        r0 = float(min(int(math.ceil(reach)) + 1, cap, self.R + 3))
        self.t0_full = t0
        self.far_r0 = r0
        wn = self._ramp(torch.hypot(t.xx_d, t.yy_d), r0)
        t0n, tdn, t2n = t0 * wn, td * wn, t2 * wn
        tx, ty = t.derivatives_xy(t0n)
        tm = torch.zeros((t.N, t.N, 8), dtype=torch.float32, device=self.device)
        for k, g in enumerate((t0n, tdn, t2n, tx, ty)):
            tm[..., k] = g.to(torch.float32)
        self.tmpl = tm.contiguous()
        self.Rr = min(int(math.ceil(r0 + self.far_ramp)) + 1, cap)
        if self._far_template is None:
            self._far_template = OutlineTemplate(upsample=self.far_upsample, size_px=self.far_box, harmonics=t.ks,
                                                 blur=t.blur, aa=t.aa, device=self.device)
        ft, B = self._far_template, self.far_box
        k = torch.arange(B, device=self.device) * ft.U
        off = torch.arange(B, device=self.device, dtype=torch.float64) - B // 2
        wf = 1.0 - self._ramp(torch.hypot(off[None, :], off[:, None]), r0)
        self.far_fft = []
        for g in ft.render_quad(self.theta):
            pad = torch.zeros((self.hp, self.wp), dtype=torch.float64, device=self.device)
            pad[:B, :B] = g[k][:, k] * wf
            self.far_fft.append(torch.fft.rfft2(pad))

    def far_field(self) -> torch.Tensor:
        """The holes' far light (H, W): a, a d, a d^2 splatted and convolved with the far kernels."""
        grid = torch.zeros((3, self.hp, self.wp), dtype=torch.float64, device=self.device)
        _lib.call("spx_outline_far_splat", _ptr(self.c), _ptr(self.a), _ptr(self.delta), self.n, self.mx, self.my,
                  self.hp, self.wp, _ptr(grid), stream_handle())
        spec = sum(torch.fft.rfft2(grid[k]) * self.far_fft[k] for k in range(3))
        far = torch.fft.irfft2(spec, s=(self.hp, self.wp))
        oy, ox = self.my + self.far_box // 2, self.mx + self.far_box // 2
        return far[oy:oy + self.h, ox:ox + self.w]

    def _render_near(self, page, order, cache, c_old):
        t = self.template
        rl2 = float(4 * (self.Rr + 1) ** 2)  # the whole square window
        _lib.call("spx_hole_render", _ptr(page), self.h, self.w, _ptr(cache), _ptr(self.tmpl), t.N, t.U, self.Rr,
                  rl2, _ptr(order), int(order.numel()), _ptr(c_old), _ptr(self.c), _ptr(self.a), _ptr(self.delta),
                  stream_handle())

    def render_page(self) -> torch.Tensor:
        """Every hole's near light over its (2Rr+1)^2 window plus the far light (H, W) FP64."""
        page = torch.zeros(self.h * self.w, dtype=torch.float64, device=self.device)
        cache = torch.zeros(self.n * (2 * self.Rr + 1) ** 2, dtype=torch.float32, device=self.device)
        self._render_near(page, self.all_idx, cache, self.c.clone())
        return page.reshape(self.h, self.w) + self.far_field()

    # ------------------------------------------------------------------------------------ start
    def _stack(self, *, cleaned, page=None, good=None, radius_px=None):
        """(sums, wsum) (N*N) host arrays of the plain or neighbour-cleaned stack (spx_outline_stack)."""
        t = self.template
        nn = t.N * t.N
        acc = torch.zeros((2, nn), dtype=torch.float64, device=self.device)
        g = None if good is None else torch.from_numpy(np.asarray(good, dtype=np.uint8)).to(self.device)
        v0, v1 = self.var_line
        _lib.call("spx_outline_stack", _ptr(self.img), self.h, self.w, self.w, _ptr(page),
                  _ptr(self.tmpl) if cleaned else None, t.N, t.U, self.R, _ptr(self.all_idx), self.n, _ptr(g),
                  _ptr(self.c), _ptr(self.a), _ptr(self.level), _ptr(self.delta), float(v0), float(v1),
                  0.0 if radius_px is None else float(radius_px) * t.U, int(bool(cleaned)), _ptr(acc[0]), _ptr(acc[1]),
                  stream_handle())
        acc = acc.cpu().numpy()
        return acc[0], acc[1]

    def plain_stack(self) -> np.ndarray:
        """Mean grey level per fine bin of every window pixel's offset from its hole (page.py:488-505)."""
        t = self.template
        sums, cnt = self._stack(cleaned=False)
        with np.errstate(invalid="ignore", divide="ignore"):
            return (sums / cnt).reshape(t.N, t.N)

    def initialise(self):
        """Round start from the plain stack's half height; brightness and level by a 2 x 2 least
        squares per hole, twice, the second against the others' light (page.py:420-486)."""
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
        page = torch.zeros((self.h, self.w), dtype=torch.float64, device=self.device)
        for _ in range(2):
            win = self.windows(page=page, level=False, outputs=("valid", "data", "model", "v0"))
            wv = win["valid"].to(torch.float64)
            v0 = win["v0"]
            y = torch.where(win["valid"], win["data"] - (win["model"] - self.a[:, None] * v0), 0.0)
            s11, s12, s22 = (wv * v0 * v0).sum(1), (wv * v0).sum(1), wv.sum(1)
            r1, r2 = (wv * v0 * y).sum(1), (wv * y).sum(1)
            det = s11 * s22 - s12 * s12
            self.a = torch.clamp((r1 * s22 - r2 * s12) / det, min=1e-3)
            self.level = (s11 * r2 - s12 * r1) / det
            page = self.render_page()

    # ------------------------------------------------------------------------------- noise line
    def fit_noise(self, page):
        """Variance of the leftovers against the model level as a straight line, from the first
        three chunks of holes (page.py:507-535)."""
        n3 = min(self.n, 3 * self.chunk)
        win = self.windows(self.all_idx[:n3], page=page, outputs=("valid", "data", "model"))
        ok = win["valid"]
        res = (win["data"] - win["model"])[ok]
        lev = win["model"][ok]
        xs, ys = _level_bins(res, lev)
        xs, ys = np.asarray(xs), np.asarray(ys)
        if len(xs) >= 3:
            v1, v0 = np.polyfit(xs, ys, 1)
            v1, v0 = max(v1, 0.0), max(v0, 0.05 * ys.mean())
        else:
            v0, v1 = max(float(ys.mean()) if len(ys) else 1e-3, 1e-3), 0.0
        self.var_line = (float(v0), float(v1))

    def _opt(self, **kw):
        v0, v1 = self.var_line
        base = dict(nv0=v0, nv1=v1, rscale=0.0, max_move=0.25, max_delta=0.0, a_med=1.0, relax=1.0, size=0,
                    reserved=0)
        base.update(kw)
        return _lib.SpxHoleStepOptions(**base)

    def delta_information(self):
        """Per-hole standard error of the edge offset from the normal matrix at the current state
        (no step, no robust weights); (median, outline radius) (page.py:553-585)."""
        t = self.template
        page = self.render_page().reshape(-1).contiguous()
        stat = torch.zeros((self.n, 4), dtype=torch.float64, device=self.device)
        c_old = torch.empty((self.n, 2), dtype=torch.float64, device=self.device)
        opt = self._opt(max_move=0.0, max_delta=1e300, a_med=1e300, relax=0.0, size=1)
        _lib.call("spx_hole_step", _ptr(self.img), self.h, self.w, self.w, _ptr(page), _ptr(self.tmpl), t.N, t.U,
                  self.R, _ptr(self.all_idx), self.n, _ptr(self.c), _ptr(self.a), _ptr(self.level),
                  _ptr(self.delta), _ptr(c_old), ctypes.byref(opt), _ptr(stat), stream_handle())
        self.delta_sd = stat[:, 3].clone()
        return float(np.nanmedian(self.delta_sd.cpu().numpy())), 0.5 * t.cd_px(self.theta)

    # ------------------------------------------------------------------------------- hole stage
    def _sweep_stats(self, stat, size):
        """(median leftover rms, 99th-percentile centre move, 99th-percentile offset move) by
        NumPy's rules, from exact order statistics selected on the device."""
        n = self.n
        pos = 0.99 * (n - 1)
        lo, hi = int(math.floor(pos)), int(math.ceil(pos))
        frac = pos - lo
        out = torch.empty(6, dtype=torch.float64, device=self.device)
        cols = (ctypes.c_int64 * 6)(2, 2, 0, 0, 1, 1)
        ks = (ctypes.c_int64 * 6)((n - 1) // 2, n // 2, lo, hi, lo, hi)
        _lib.call("spx_select_kth", _ptr(stat), n, 4, cols, ks, 6, _ptr(out), stream_handle())
        v = out.cpu().tolist()

        def lerp(a, b):
            d = b - a
            return b - d * (1.0 - frac) if frac >= 0.5 else a + d * frac

        return 0.5 * (v[0] + v[1]), lerp(v[2], v[3]), lerp(v[4], v[5]) if size else 0.0

    def fit_holes(self, iterations=5, size=True, max_move=0.25, stop_px=0.002, robust_c=3.0):
        """Block Gauss-Seidel over the colour chunks (page.py:587-729): every chunk's step against
        the complete page model, its light re-rendered before the next chunk; the far light
        refreshed after every sweep; stops once 99 % of the holes move less than stop_px."""
        t = self.template
        max_delta = 0.3 * 0.5 * t.cd_px(self.theta)
        self.clip_px = float(max_delta)
        dev, n = self.device, self.n
        P = (2 * self.Rr + 1) ** 2
        page = torch.zeros(self.h * self.w, dtype=torch.float64, device=dev)
        cache = torch.zeros(n * P, dtype=torch.float32, device=dev)
        c_old = self.c[self.order.long()].contiguous()
        self._render_near(page, self.order, cache, c_old)
        far = self.far_field().reshape(-1)
        page += far
        if not self.history:
            self.fit_noise(page.reshape(self.h, self.w))
        a_med = _np_median(self.a)
        stat = torch.zeros((n, 4), dtype=torch.float64, device=dev)
        opt = self._opt(max_move=float(max_move), max_delta=float(max_delta), a_med=a_med, size=int(bool(size)))
        rl2 = float(4 * (self.Rr + 1) ** 2)
        starts = self.starts
        nchunk = len(starts) - 1
        sweeps = 0
        for _ in range(int(iterations)):
            sweeps += 1
            if self.rms_typ is None:
                # the first sweep with a new template: each chunk's robust scale is its own median
                # leftover on the page as it stands when the chunk's turn comes
                for j in range(nchunk):
                    b, m = int(starts[j]), int(starts[j + 1] - starts[j])
                    if m == 0:
                        continue
                    idx = self.order[b:b + m]
                    if robust_c:
                        win = self.windows(idx, page=page, outputs=("valid", "data", "model"))
                        res = torch.where(win["valid"], win["data"] - win["model"], 0.0)
                        cnt = win["valid"].sum(1).clamp(min=1).to(torch.float64)
                        s = _np_median(torch.sqrt((res * res).sum(1) / cnt))
                        opt.rscale = float(robust_c) * max(s, 1e-6)
                    else:
                        opt.rscale = 0.0
                    _lib.call("spx_hole_step", _ptr(self.img), self.h, self.w, self.w, _ptr(page), _ptr(self.tmpl),
                              t.N, t.U, self.R, _ptr(idx), m, _ptr(self.c), _ptr(self.a), _ptr(self.level),
                              _ptr(self.delta), _off(c_old, 2 * b), ctypes.byref(opt), _off(stat, 4 * b),
                              stream_handle())
                    _lib.call("spx_hole_render", _ptr(page), self.h, self.w, _off(cache, P * b), _ptr(self.tmpl), t.N,
                              t.U, self.Rr, rl2, _ptr(idx), m, _off(c_old, 2 * b), _ptr(self.c), _ptr(self.a),
                              _ptr(self.delta), stream_handle())
            else:
                opt.rscale = float(robust_c) * max(self.rms_typ, 1e-6) if robust_c else 0.0
                start = (ctypes.c_int32 * len(starts))(*starts.tolist())
                _lib.call("spx_hole_sweep", _ptr(self.img), self.h, self.w, self.w, _ptr(page), _ptr(self.tmpl), t.N,
                          t.U, self.R, self.Rr, rl2, _ptr(self.order), start, nchunk, _ptr(self.c), _ptr(self.a),
                          _ptr(self.level), _ptr(self.delta), _ptr(c_old), _ptr(cache), ctypes.byref(opt),
                          _ptr(stat), stream_handle())
            new_far = self.far_field().reshape(-1)
            page += new_far - far
            far = new_far
            if size:
                self.delta_sd[self.order.long()] = stat[:, 3]
            self.rms_typ, m99, d99 = self._sweep_stats(stat, size)
            self.log.append(f"    holes it {sweeps}: 99% move {m99:.4f} px, delta 99% {d99:.4f}, typical leftover "
                            f"rms {self.rms_typ:.2f}")
            if m99 < stop_px and d99 < stop_px:
                break
        self.sweeps_run = getattr(self, "sweeps_run", 0) + sweeps
        return page.reshape(self.h, self.w)

    # ------------------------------------------------------------------------------ shape stage
    def inside(self) -> np.ndarray:
        c = self.c.cpu().numpy()
        return ((c[:, 0] >= self.R + 1) & (c[:, 0] <= self.w - 2 - self.R) & (c[:, 1] >= self.R + 1)
                & (c[:, 1] <= self.h - 2 - self.R))

    def window_rms(self, page) -> torch.Tensor:
        """Leftover rms of every hole's window against the page model (device, n)."""
        win = self.windows(page=page, outputs=("valid", "data", "model"))
        left = torch.where(win["valid"], win["data"] - win["model"], 0.0)
        return torch.sqrt((left * left).sum(1) / win["valid"].sum(1).clamp(min=1).to(torch.float64))

    def good_holes(self, page):
        """(good, rms) host arrays: the holes the shape stage may stack. A good hole's window lies
        inside the page, its brightness is within 0.4 - 2.5 times and its edge offset within
        0.6 px of the typical hole's (the medians over the holes inside), and its leftover rms
        below twice the typical one (page.py:784-803)."""
        rms = self.window_rms(page)
        ins = torch.from_numpy(self.inside()).to(self.device)
        pool = ins if bool(ins.any()) else torch.ones_like(ins)
        a_typ = _np_median(self.a[pool])
        d_typ = _np_median(self.delta[pool])
        r_typ = _np_median(rms[pool])
        ok = ins & (self.a > 0.4 * a_typ) & (self.a < 2.5 * a_typ) & ((self.delta - d_typ).abs() < 0.6)
        ok &= rms < 2.0 * r_typ
        return ok.cpu().numpy(), rms.cpu().numpy()

    def cleaned_stack(self, page, good, radius_px=None):
        """The neighbour-cleaned average hole on the fine grid and its weight, (N, N) host arrays
        (NaN where no pixel fell): every good hole's window pixels without the other holes' light
        and the level, over its brightness, weighted by a^2 / variance (page.py:732-766)."""
        t = self.template
        total, weight = self._stack(cleaned=True, page=page.reshape(-1).contiguous(), good=good,
                                    radius_px=radius_px)
        mean = np.full(total.shape, np.nan)
        np.divide(total, weight, out=mean, where=weight > 0)
        return mean.reshape(t.N, t.N), weight.reshape(t.N, t.N)

    def fit_shape(self, stack, weight, stages):
        t = self.template
        ok = np.isfinite(stack) & (weight > 0)
        okd = torch.from_numpy(ok).to(self.device)
        y = torch.from_numpy(stack[ok]).to(self.device)
        sw = torch.from_numpy(np.sqrt(weight[ok])).to(self.device)
        return fit_stages(t, lambda ths: t.render_batch(ths, box_px=1.0 / t.U)[:, okd], y, sw, self.theta,
                          stages=stages, log=self.log.append)

    # -------------------------------------------------------------------------------------- run
    def _take_proposal(self, proposal, scale, weight):
        """Move theta `weight` of the way to the shape stage's proposal and the brightnesses by
        the matching power of its scale; returns the proposal's outline-CD change (px)."""
        t = self.template
        change = t.cd_px(proposal) - t.cd_px(self.theta)
        self.a = self.a * scale ** weight
        self.set_theta(self.theta + weight * (proposal - self.theta))
        return change

    def _choose_size(self, note):
        sd, radius = self.delta_information()
        self.size_on = sd < 0.3 * radius
        note(f"  edge offsets per hole {'fitted' if self.size_on else 'held at zero'}: their typical "
             f"standard error is {sd:.3f} px for an outline radius of {radius:.2f} px")

    def run(self, rounds=14, iterations=5, log=None, relax=0.5, tol_px=0.08 / 21.0, release_px=0.3 / 21.0,
            per_hole_size="auto"):
        """The page fit: alternate hole stage and shape stage under _RoundPlan for at most
        `rounds` rounds, then a last hole stage; returns the readings (_page_readings)."""
        t_fit = time.perf_counter()
        note = log if log is not None else (lambda s: None)
        self.initialise()
        plan = _RoundPlan(iterations, relax, tol_px, release_px)
        self.size_on = per_hole_size is True
        for rnd in range(int(rounds)):
            t_rnd = time.perf_counter()
            if rnd == 1 and per_hole_size == "auto":
                self._choose_size(note)
            page = self.fit_holes(iterations=plan.sweeps(rnd), size=self.size_on and rnd > 0)
            self.fit_noise(page)
            good, _ = self.good_holes(page)
            stack, weight = self.cleaned_stack(page, good, radius_px=plan.stack_radius(rnd, self.spacing))
            stages = plan.stages()
            proposal, scale, const, chi2 = self.fit_shape(stack, weight, stages)
            moved = self._take_proposal(proposal, scale, plan.step_weight(rnd))
            cd = self.template.cd_px(self.theta)
            self.history.append(dict(round=rnd + 1, stages=list(stages), outline_cd_px=cd, chi2=chi2,
                                     proposal_change_px=moved, const=const, good=int(good.sum()),
                                     var_line=self.var_line, mean_delta_px=float(self.delta.mean()),
                                     seconds=time.perf_counter() - t_rnd))
            note(f"  [{rnd + 1}] {stages[-1]} model: CD {cd:.4f} px after a {moved:+.4f} px proposal, "
                 f"stack chi2 {chi2:.3f} over {int(good.sum())} holes")
            for line in self.log:
                note(line)
            self.log = []
            verdict = plan.decide(rnd, chi2, moved)
            if verdict == "release":
                note(f"  the {plan.stages()[-1]} model is free from the next round on")
            elif verdict == "done":
                self.converged = True
                note(f"  outline settled in round {rnd + 1}")
                break
        page = self.fit_holes(iterations=2 * iterations, size=self.size_on)
        if not self.size_on:
            self.delta_information()
        good, rms = self.good_holes(page)
        stack, weight = self.cleaned_stack(page, good)
        self.timings["total"] = time.perf_counter() - t_fit
        out = _page_readings(self, good, rms)
        t = self.template
        out.update(stack=stack, stack_weight=weight, template=self.t0_full.cpu().numpy(),
                   template_box=t.render(self.theta, box_px=1.0 / t.U).cpu().numpy())
        return out


class _RoundPlan:
    """The round schedule of the page fit (the reference's PageFit.run, page.py:806-896, keeps
    the same numbers): `iterations` + 3 hole sweeps in the first three rounds and
    max(3, iterations // 2 + 1) after them; the first round stacks only the core (0.55 of the
    spacing) and takes the whole proposal, later rounds `relax` of it. The model starts round and
    is released one stage further ("super", then "full") once the stack's chi2 holds within 5 %
    over two rounds and the proposal moves the outline CD by less than `release_px`; with every
    stage free, the same steadiness and a move below `tol_px` end the fit."""

    ORDER = ("round", "super", "full")

    def __init__(self, iterations, relax, tol_px, release_px):
        self.iterations, self.relax = int(iterations), float(relax)
        self.tol_px, self.release_px = float(tol_px), float(release_px)
        self.level = 0
        self.last_chi2 = None

    def sweeps(self, rnd):
        return self.iterations + 3 if rnd < 3 else max(3, self.iterations // 2 + 1)

    def stack_radius(self, rnd, spacing):
        return 0.55 * spacing if rnd == 0 else None

    def step_weight(self, rnd):
        return 1.0 if rnd == 0 else self.relax

    def stages(self):
        return self.ORDER[:self.level + 1]

    def decide(self, rnd, chi2, moved):
        """'release', 'done' or 'go' after round `rnd` (0-based)."""
        steady = self.last_chi2 is not None and abs(chi2 - self.last_chi2) < 0.05 * self.last_chi2
        self.last_chi2 = chi2
        last = len(self.ORDER) - 1
        if rnd > 0 and steady and abs(moved) < self.release_px and self.level < last:
            self.level += 1
            self.last_chi2 = None
            return "release"
        if self.level == last and steady and abs(moved) < self.tol_px:
            return "done"
        return "go"


def _offset_area_coefficients(t, theta):
    """(c0, c1, c2) with area(d) = c0 + c1 d + c2 d^2 the exact area of the outline moved d px
    along its normal: in polar form r + d g with g = |(r, r')| / r, so the area
    (1/2) integral (r + d g)^2 over the angle is quadratic in d."""
    r, dr = t.radius_at(theta, t.angles)
    g = np.hypot(r, dr) / r
    return math.pi * np.mean(r * r), 2.0 * math.pi * np.mean(r * g), math.pi * np.mean(g * g)


def _page_readings(pg, good, rms):
    """The per-hole and page readings of a finished fit (the fields of PageFit.results,
    page.py:898-937): centres at the outline's centroid, edge offsets about the good holes'
    mean, each hole's CD (and its standard error) from the exact area of its offset outline."""
    t, theta = pg.template, pg.theta
    good = np.asarray(good, dtype=bool)
    d_raw = pg.delta.cpu().numpy()
    d_mean = float(np.mean(d_raw[good])) if good.any() else 0.0
    d = d_raw - d_mean
    c0, c1, c2 = _offset_area_coefficients(t, theta)
    area = np.maximum(c0 + d * (c1 + c2 * d), 1e-6)
    d_sd = pg.delta_sd.cpu().numpy()
    return dict(
        centers_xy=pg.c.cpu().numpy() + np.asarray(t.centroid_px(theta)), brightness=pg.a.cpu().numpy(),
        level=pg.level.cpu().numpy(), edge_offset_px=d, cd_px=2.0 * np.sqrt(area / math.pi),
        cd_outline_px=t.cd_px(theta), theta=np.array(theta, copy=True), delta_sd_px=d_sd,
        cd_sd_px=d_sd * (c1 + 2.0 * c2 * d) / np.sqrt(math.pi * area), per_hole_size=pg.size_on,
        edge_offset_mean_px=d_mean, clip_px=pg.clip_px, at_clip=np.abs(d_raw) >= 0.99 * (pg.clip_px or np.inf),
        describe=t.describe(theta, 1.0), good=good, rms=np.asarray(rms), var_line=pg.var_line, history=pg.history,
        seconds=pg.timings.get("total"), spacing_px=pg.spacing, window_px=pg.R, converged=pg.converged)


# ------------------------------------------------------------------------------------ public API
def fit_outline(image, seeds_xy, *, pixel_size: float = 1.0, unit: str = "px", blur: str = "gauss2",
                rounds: int = 14, invert: bool = False, n_rays: int = 128, backend: str = "auto", device: int = 0,
                log=None) -> OutlineFitResult:
    """Fit one outline and one blur to the whole page and every hole's position, brightness,
    level and size (reference subpx/outline.py:34-96, same arguments and checks). Dark holes on a
    bright page (``invert``) are reflected as max - image. Returns NumPy arrays in pixels."""
    img = np.asarray(image) if not isinstance(image, torch.Tensor) else image.detach().cpu().numpy()
    if img.ndim != 2 or img.size == 0 or not np.issubdtype(img.dtype, np.number) or not np.isfinite(img).all():
        raise ValueError("image must be a nonempty 2D array of finite intensities")
    seeds = np.asarray(seeds_xy, dtype=np.float64).reshape(-1, 2)
    if len(seeds) < MIN_HOLES or not np.isfinite(seeds).all():
        raise ValueError(f"the outline fit needs at least {MIN_HOLES} finite seeds")
    if blur not in BLUR_FAMILIES:
        raise ValueError("blur must be one of " + ", ".join(BLUR_FAMILIES))
    if int(rounds) != rounds or rounds < 1:
        raise ValueError("rounds must be a positive integer")
    if int(n_rays) != n_rays or n_rays < 16:
        raise ValueError("n_rays must be an integer of at least 16")
    if not (np.isfinite(pixel_size) and pixel_size > 0):
        raise ValueError("pixel_size must be positive")
    resolve_backend(backend)
    signal = img.astype(np.float64)
    top = float(signal.max())
    if invert:
        signal = top - signal
    with gpu_device(device) as dev:
        template = OutlineTemplate(upsample=UPSAMPLE, size_px=BLUR_FAMILIES[blur], blur=blur, device=dev)
        page = OutlinePage(signal, seeds, template, device=dev)
        res = page.run(rounds=int(rounds), iterations=5, log=log)
        template = page.template
        theta = np.asarray(res["theta"], dtype=float)
        delta = np.asarray(res["edge_offset_px"], dtype=float)
        cd_x, cd_y = ellipse_widths(template, theta, delta)
        rays = outline_rays(template, theta, delta, int(n_rays))
        d = res["describe"]
    level = np.asarray(res["level"], dtype=float)
    if invert:
        level = top - level
    fwhm = tuple(float(v) for v in d["blur_fwhm_nm"])
    summary = {
        "blur_family": blur, "rounds_max": int(rounds), "rounds_run": len(page.history),
        "converged": bool(res["converged"]), "per_hole_size": bool(res["per_hole_size"]),
        "outline_cd_px": float(res["cd_outline_px"]), "outline_extent_x_px": float(d["extent_x_nm"]),
        "outline_extent_y_px": float(d["extent_y_nm"]), "superellipse_p": float(d["superellipse"]["p"]),
        "outline_turn_deg": float(d["superellipse"]["phi_deg"]),
        "harmonics": {k: [float(v) for v in pair] for k, pair in d["harmonics"].items()},
        "blur_fwhm_px": list(fwhm), "blur_rms_px": [float(v) for v in d["blur_rms_nm"]],
        "holes_fitted": int(len(seeds)), "good_holes": int(np.sum(res["good"])),
        "spacing_px": float(res["spacing_px"]), "window_px": int(res["window_px"]),
        "template_box_px": int(template.N // template.U), "noise_variance": [float(v) for v in res["var_line"]],
        "edge_offset_clip_px": None if res["clip_px"] is None else float(res["clip_px"]),
        "edge_offset_mean_px": float(res["edge_offset_mean_px"]),
        "edge_offset_sd_median_px": float(np.nanmedian(res["delta_sd_px"])) if np.isfinite(
            res["delta_sd_px"]).any() else float("nan"),
        "at_clip": int(np.sum(res["at_clip"])), "seconds": float(res["seconds"]), "backend": "gpu",
        "theta": [float(v) for v in theta], "lengths": "px",
    }
    return OutlineFitResult(
        centers_xy=np.asarray(res["centers_xy"], dtype=float), cd_px=np.asarray(res["cd_px"], dtype=float),
        cd_x_px=cd_x, cd_y_px=cd_y, cd_se_px=np.asarray(res["cd_sd_px"], dtype=float), edge_offset_px=delta,
        brightness=np.asarray(res["brightness"], dtype=float), level=level, rays_px=rays,
        at_clip=np.asarray(res["at_clip"], dtype=bool), good=np.asarray(res["good"], dtype=bool),
        rms=np.asarray(res["rms"], dtype=float), outline_cd_px=float(res["cd_outline_px"]), blur_fwhm_px=fwhm,
        per_hole_size=bool(res["per_hole_size"]), converged=bool(res["converged"]), rounds=len(page.history),
        blur=blur, theta=theta, template=np.asarray(res["template"], dtype=float),
        stack=np.asarray(res["stack"], dtype=float), upsample=int(template.U), summary=summary,
        pixel_size=float(pixel_size), unit=str(unit),
    )


def render_outline_page(theta, centers, brightness, shape, *, level: float = 500.0, noise: float = 30.0,
                        edge_offset=None, blur: str = "gauss2", seed: int = 0, device: int = 0) -> torch.Tensor:
    """A synthetic page of the outline model on the device: every hole's light
    a (T0 + d T1 + d^2 T2) of the outline and blur `theta` (near windows + far light, as the fit's
    page model), a constant level and N(0, noise^2) pixel noise. (H, W) FP64 device tensor.
    Synthetic input for tests and benchmarks."""
    h, w = int(shape[0]), int(shape[1])
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 2)
    with gpu_device(device) as dev:
        t = OutlineTemplate(upsample=UPSAMPLE, size_px=BLUR_FAMILIES[blur], blur=blur, device=dev)
        pg = OutlinePage(torch.zeros((h, w), dtype=torch.float32, device=dev), c, t, device=dev)
        pg.set_theta(np.asarray(theta, float))
        pg.a = torch.as_tensor(np.asarray(brightness, dtype=np.float64)).to(dev).clone()
        if edge_offset is not None:
            pg.delta = torch.as_tensor(np.asarray(edge_offset, dtype=np.float64)).to(dev).clone()
        page = pg.render_page() + float(level)
        if noise:
            g = torch.Generator(device=dev)
            g.manual_seed(int(seed))
            page += float(noise) * torch.randn(page.shape, generator=g, dtype=torch.float64, device=dev)
        return page


_reading_templates: dict = {}


def _reading_template(blur, device):
    """The template the page readings (extents, blur widths) are taken on: fit_outline's first box
    of the blur form, no coverage sub-samples (outline.py:339-349)."""
    key = (blur, str(device))
    if key not in _reading_templates:
        _reading_templates[key] = OutlineTemplate(upsample=UPSAMPLE, size_px=BLUR_FAMILIES[blur], blur=blur, aa=1,
                                                  device=device)
    return _reading_templates[key]


def fit_outline_fast(image, seeds_xy=None, *, pixel_size: float = 1.0, unit: str = "px", blur: str = "gauss2",
                     invert: bool = False, n_rays: int = 128, backend: str = "auto", device: int = 0, log=None,
                     trace: dict | None = None, redrawn: bool = False):
    """One outline and one blur for the whole image in one pass (reference outline.py:361-493):
    the holes are averaged as they are (binned at their sub-pixel offsets), the neighbours' light is
    put into the model of that average, outline and blur are fitted to it; every hole is then
    centred against the complete page model and the fit repeated with those centres
    (``onepass.fit_images``). Page values, no size per hole; GPU only.

    image (H, W): bright holes (``invert=True``: dark holes, reflected as max - image). seeds_xy
    (N, 2) [x, y], or None for the method's own seed finder. blur: 'gauss2' or 'lens'. n_rays >= 16
    directions of the page's outline in ``rays_px``. ``trace`` receives the per-page record of the
    fit (``fit_images``' dict and ``image``: inverted, top, own_seeds); ``redrawn`` is accepted for
    the signature and adds nothing. Returns ``OutlineFastResult`` (lengths in px)."""
    from .types import OutlineFastResult

    started = time.perf_counter()
    img = np.asarray(image)
    if img.ndim != 2 or img.size == 0 or not np.issubdtype(img.dtype, np.number) or not np.isfinite(img).all():
        raise ValueError("image must be a nonempty 2D array of finite intensities")
    seeds = None
    if seeds_xy is not None:
        seeds = np.asarray(seeds_xy, dtype=np.float64).reshape(-1, 2)
        if len(seeds) < MIN_HOLES or not np.isfinite(seeds).all():
            raise ValueError(f"the outline fit needs at least {MIN_HOLES} finite seeds")
    if blur not in BLUR_FAMILIES:
        raise ValueError("blur must be one of " + ", ".join(BLUR_FAMILIES))
    if int(n_rays) != n_rays or n_rays < 16:
        raise ValueError("n_rays must be an integer of at least 16")
    if not (np.isfinite(pixel_size) and pixel_size > 0):
        raise ValueError("pixel_size must be positive")
    if backend not in ("auto", "gpu", "cpu"):
        raise ValueError("backend must be 'auto', 'gpu' or 'cpu'")
    if backend == "cpu" or not torch.cuda.is_available():
        raise RuntimeError("fit_outline_fast needs the GPU; it has no CPU path")
    resolve_backend("gpu")
    from . import onepass

    signal = img.astype(np.float64)
    top = float(signal.max())
    if invert:
        signal = top - signal
    with gpu_device(device) as dev:
        own = seeds is None
        if own:
            img_t = torch.from_numpy(signal).to(device=dev, dtype=torch.float32)
            seeds = onepass.find_seeds_device(img_t).cpu().numpy()
            if len(seeds) < MIN_HOLES:
                raise ValueError(f"the outline fit needs at least {MIN_HOLES} holes; {len(seeds)} were found")
        res = onepass.fit_images([signal], [seeds], blur=blur, device=device, log=log)[0]
        t = onepass.lib_template(blur, dev)
        theta = res["theta"]
        zero = np.zeros(1)
        cd = float(t.cd_px(theta))
        cd_x, cd_y = (float(v[0]) for v in ellipse_widths(t, theta, zero))
        rays = outline_rays(t, theta, zero, int(n_rays))[0]
        centers = res["centers_xy"] + np.asarray(t.centroid_px(theta))
        d = _reading_template(blur, dev).describe(theta, 1.0)
    if trace is not None:
        trace.update(res)
        trace["image"] = dict(inverted=bool(invert), top=top, own_seeds=own)
    level = top - res["level"] if invert else res["level"]
    fwhm = tuple(float(v) for v in d["blur_fwhm_nm"])
    summary = {
        "blur_family": blur, "outline_cd_px": cd,
        "outline_extent_x_px": float(d["extent_x_nm"]), "outline_extent_y_px": float(d["extent_y_nm"]),
        "superellipse_p": float(d["superellipse"]["p"]), "outline_turn_deg": float(d["superellipse"]["phi_deg"]),
        "harmonics": {k: [float(v) for v in pair] for k, pair in d["harmonics"].items()},
        "blur_fwhm_px": [float(v) for v in fwhm], "blur_rms_px": [float(v) for v in d["blur_rms_nm"]],
        "holes_fitted": int(len(seeds)), "holes_stacked": int(np.sum(res["good"])),
        "spacing_px": float(res["spacing_px"]), "window_px": int(res["window_px"]),
        "pair_reach_px": float(res["pair_reach_px"]), "noise_sd": float(res["noise_sd"]),
        "leftover_chi2": float(res["chi2"]) / max(int(res["bins"]), 1),
        "first_cd_px": float(t.cd_px(res["theta_first"])),
        "fit_iterations": int(res["iterations"]), "converged": bool(res["converged"]),
        "exchanged": bool(res["exchanged"]), "per_hole_size": False,
        "seconds": time.perf_counter() - started, "backend": "gpu", "theta": [float(v) for v in theta],
        "lengths": "px",
    }
    return OutlineFastResult(
        centers_xy=centers, brightness=res["brightness"], level=level, rms=res["rms"],
        good=np.asarray(res["good"], dtype=bool), outline_cd_px=cd, cd_x_px=cd_x, cd_y_px=cd_y, rays_px=rays,
        blur_fwhm_px=fwhm, theta=theta, stack=res["stack"], stack_weight=res["stack_weight"],
        upsample=int(res["upsample"]), blur=blur, summary=summary, pixel_size=float(pixel_size), unit=str(unit))
