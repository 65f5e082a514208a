"""One-pass outline fit on the GPU: the page's outline and blur in one pass over the image.

The reference's ``fit_outline_fast`` (subpx/outline.py:361-493) and its driver
``subpx/_gpu/outline_onepass.py`` with the page fitter ``subpx/_outline/onepass.py``. Per image:

    A  seeds -> plain stack -> pair maps -> first page fit (grid of starts, round + superellipse)
    B  hole pass with the size held (centre, brightness, level of every hole; two sweeps)
    C  plain stack again with B's centres -> pair maps -> final page fit from A's answer, all free;
       gauss2: a fit that ends with its two Gaussians exchanged is run again from the exchanged labels

- plain stack: every stacked hole's window binned at its offset on a quarter-pixel grid, the plain
  mean (``spx_onepass_stack``); each hole against it (``spx_onepass_versus``);
- pair maps Gamma_q: for each of the 16 sub-pixel classes, where the stacked holes' neighbours and
  the holes themselves stand (``spx_onepass_pair_hist`` + a real FFT + ``spx_onepass_pair_crop``):
  the neighbours' light is part of the stack's model;
- page fit (``Fitter``): S(b) = scale Re sum_f T^(f) Gamma_q(f) exp(2 pi i f . u_b) / L^2 + const,
  T^ = outline transform (trapezoid rule along the outline, ``spx_onepass_outline_hat``) x blur
  transfer function x pixel box (``spx_onepass_coeff``), the model by class
  (``spx_onepass_synth``), scale and constant solved exactly and the cost, gradient and
  Gauss-Newton matrix of the free parameters per page (``spx_onepass_project``), all FP64 on the
  device; the Levenberg-Marquardt iterations (bounds as an active set, gain-ratio damping, weak
  priors on the second Gaussian) run there too (``spx_onepass_lm_*``), pages in lockstep, the
  host reading one flag per iteration;
- hole pass: the hole-stage kernels of ``holes.HolePage`` with templates synthesised from the
  fitted outline and blur (the polygon's exact transform on a band, ``spx_onepass_band``, and an
  inverse FFT).

There is no CPU path: the library must load and a GPU must be present.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, stream_handle
from .holes import HolePage, _even_fast

STACK_BINS = 4  # stack bins per px
PAIR_BINS = 8  # pair-map grid points per px
F_MAX = 0.6  # band of the page fit, cycles per px
N_POLY = 256  # outline points of the page fit
PAIR_REACH_MARGIN_PX = 13.0  # neighbours within the window's half width + this are in the stack's model
FIRST_STAGES, FIRST_ITERATIONS, FIRST_FTOL = ("round", "super"), (10, 12), 1e-6
FINAL_ITERATIONS = 30
CENTRING_SWEEPS = 2
GRID_SIZES = np.linspace(0.4, 1.8, 15)  # grid of starts: outline half widths relative to the start
GRID_WIDTHS = np.linspace(0.45, 2.2, 12)  # and blur widths, px
BLUR_KIND = {"gauss2": 0, "lens": 1, "gauss1": 2}


def _p(t):
    return None if t is None else int(t.data_ptr())


def _dev(a, dev, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


def _cplx(a, dev):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex128))).to(dev)


def _lerp(a, b, t):
    """NumPy's linear interpolation between neighbouring order statistics."""
    d = b - a
    return b - d * (1.0 - t) if t >= 0.5 else a + d * t


def percentiles_device(x: torch.Tensor, qs) -> list[float]:
    """NumPy's linear percentiles of a device FP64 vector from exact order statistics
    (``spx_select_kth``)."""
    n = int(x.numel())
    pos = []
    for q in qs:
        qf = float(q) / 100.0
        vi = n * qf + (1.0 + qf * (1.0 - 1.0 - 1.0)) - 1.0  # numpy's virtual index, alpha = beta = 1
        lo = int(math.floor(vi))
        lo = min(max(lo, 0), n - 1)
        pos.append((lo, min(lo + 1, n - 1), vi - math.floor(vi)))
    ks = [k for lo, hi, _ in pos for k in (lo, hi)]
    out = torch.empty(len(ks), dtype=torch.float64, device=x.device)
    _lib.call("spx_select_kth", _p(x), n, 1, None, (ctypes.c_int64 * len(ks))(*ks), len(ks), _p(out),
              stream_handle())
    v = out.cpu().tolist()
    return [_lerp(v[2 * i], v[2 * i + 1], t) if hi != lo else v[2 * i] for i, (lo, hi, t) in enumerate(pos)]


# ---------------------------------------------------------------------------------------- seeds
def gaussian_taps(sigma: float, truncate: float = 2.5) -> np.ndarray:
    """Normalised Gaussian taps w[0..r], r = int(truncate sigma + 0.5) (scipy's gaussian_filter1d)."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (float(sigma) ** 2) * x * x)
    return (phi / phi.sum())[r:]


def find_seeds_device(img: torch.Tensor, smooth_sigma: float = 1.0, nms: int = 2,
                      rel_threshold: float = 0.3) -> torch.Tensor:
    """Bright blobs' centres (N, 2) [x, y] FP64 on the device (outline_onepass.py:463-501): local
    maxima (+-nms px) of the smoothed page above lo + rel_threshold (hi - lo) (lo, hi = 5th and
    99.8th percentiles of every 4th pixel), none on the frame; while the blobs are wider than the
    window (their size read from the area above the level per maximum) the window and the
    smoothing widen (3 tries at most); each maximum is refined by a parabola through the log of
    the height above lo - 0.02 (hi - lo) along x and along y. img: (H, W) FP32 on the device."""
    H, W = (int(v) for v in img.shape)
    dev = img.device
    sh = stream_handle
    sub = img[::4, ::4].reshape(-1).to(torch.float64).contiguous()
    lo, hi = percentiles_device(sub, (5.0, 99.8))
    thr = lo + rel_threshold * (hi - lo)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    _lib.call("spx_onepass_count_above", _p(img), H, W, 2, float(thr), _p(count), sh())
    n2 = ((H + 1) // 2) * ((W + 1) // 2)
    above = float(count.item()) / n2 * (H * W)
    tmp = torch.empty((H, W), dtype=torch.float64, device=dev)
    sm = torch.empty((H, W), dtype=torch.float64, device=dev)
    flags = torch.empty((H, W), dtype=torch.uint8, device=dev)
    rows = torch.empty(H, dtype=torch.int32, device=dev)
    n = 0
    counts = None
    for _ in range(3):
        taps = _dev(gaussian_taps(smooth_sigma), dev)
        _lib.call("spx_onepass_smooth", _p(img), H, W, _p(taps), int(taps.numel()) - 1, _p(tmp), _p(sm), sh())
        _lib.call("spx_onepass_peaks", _p(sm), H, W, int(nms), float(thr), _p(flags), _p(rows), sh())
        counts = rows.cpu().numpy().astype(np.int64)
        n = int(counts.sum())
        if n == 0:
            break
        width = math.sqrt(4.0 * above / (math.pi * n))
        if width <= 2 * nms + 1:
            break
        nms = int(math.ceil(0.6 * width))
        smooth_sigma = max(smooth_sigma, 0.2 * width)
    if n == 0:
        return torch.zeros((0, 2), dtype=torch.float64, device=dev)
    offsets = _dev(np.concatenate([[0], np.cumsum(counts)[:-1]]), dev, torch.int32)
    seeds = torch.empty((n, 2), dtype=torch.float64, device=dev)
    _lib.call("spx_onepass_peak_list", _p(sm), _p(flags), _p(offsets), H, W, float(lo - 0.02 * (hi - lo)), _p(seeds),
              sh())
    return seeds


# ------------------------------------------------------------------------------------ cell list
class CellList:
    """Holes sorted by square cells of side cs (outline_onepass.py:504-519); host bookkeeping,
    the order and the cells' first positions uploaded as int32."""

    def __init__(self, c: np.ndarray, h: int, w: int, cs: float, dev):
        self.cs = float(cs)
        self.ncx = int(math.ceil(w / cs)) + 1
        self.ncy = int(math.ceil(h / cs)) + 1
        cx = np.clip(np.floor(c[:, 0] / cs), 0, self.ncx - 1).astype(np.int64)
        cy = np.clip(np.floor(c[:, 1] / cs), 0, self.ncy - 1).astype(np.int64)
        cell = cy * self.ncx + cx
        order = np.argsort(cell, kind="stable")
        start = np.searchsorted(cell[order], np.arange(self.ncx * self.ncy + 1))
        self.order = _dev(order, dev, torch.int32)
        self.start = _dev(start, dev, torch.int32)

    def nearest(self, c: torch.Tensor) -> np.ndarray:
        out = torch.empty(int(c.shape[0]), dtype=torch.float64, device=c.device)
        _lib.call("spx_onepass_nearest", _p(c), int(c.shape[0]), _p(self.order), _p(self.start), self.ncx, self.ncy,
                  self.cs, _p(out), stream_handle())
        return out.cpu().numpy()


# ---------------------------------------------------------------------------------- page fitter
def stage_map(t, stage):
    """(free theta indices, tie matrix (P, nf)) of a stage: round = circle under an isotropic blur,
    super = superellipse and turn with the whole blur, full = everything (model.py:270-290)."""
    from .outline import _stage_parameters

    free, ties = _stage_parameters(t, stage)
    tm = np.zeros((t.n_params, len(free)))
    for i, k in enumerate(free):
        tm[k, i] = 1.0
    for k, src in ties.items():
        tm[k, free.index(src)] = 1.0
    return free, tm


class Fitter:
    """Outline and blur of a batch of pages from their stacks (onepass.py:165-623): device
    evaluation, host Levenberg-Marquardt. One per (template, window R, box L)."""

    def __init__(self, t, *, box_px: int, f_max: float = F_MAX, U: int = STACK_BINS, R: int, n_poly: int = N_POLY,
                 device=None):
        self.t, self.blur, self.ks = t, t.blur, t.ks
        self.kind = BLUR_KIND[t.blur]
        self.n_shape, self.n_blur, self.P = t.n_shape, t.n_blur, t.n_params
        self.L, self.U, self.R, self.n_poly = int(box_px), int(U), int(R), int(n_poly)
        dev = device
        self.device = dev
        self.phi = 2.0 * np.pi * np.arange(self.n_poly) / self.n_poly
        kmax = int(math.floor(f_max * self.L))
        KY, KX = np.meshgrid(np.arange(-kmax, kmax + 1), np.arange(0, kmax + 1), indexing="ij")
        keep = (KY * KY + KX * KX) <= (f_max * self.L) ** 2
        self.kmax, self.nky, self.nkx = kmax, 2 * kmax + 1, kmax + 1
        sel = np.flatnonzero(keep.ravel())
        self.F = len(sel)
        fmap = np.full(self.nky * self.nkx, -1, dtype=np.int32)
        fmap[sel] = np.arange(self.F, dtype=np.int32)
        fx = (KX / self.L).ravel()[sel]
        fy = (KY / self.L).ravel()[sel]
        self.fx, self.fy = _dev(fx, dev), _dev(fy, dev)
        self.pix = _dev(np.sinc(fx) * np.sinc(fy), dev)
        self.dbl = _dev(np.where(KX.ravel()[sel] == 0, 1.0, 2.0), dev)
        self.fmap = _dev(fmap, dev, torch.int32)
        # bins n in [-nmax, nmax] per axis (offset n / U px), by class n mod U, padded with -1
        self.nmax = int(math.floor((self.R - 0.5) * self.U))
        n = np.arange(-self.nmax, self.nmax + 1)
        cls = np.mod(n, self.U)
        self.npts = max(int(np.sum(cls == q)) for q in range(self.U))
        idx = np.full((self.U, self.npts), -1, dtype=np.int64)
        for q in range(self.U):
            where = np.flatnonzero(cls == q)
            idx[q, :len(where)] = where
        self.idx, self.pad = idx, idx >= 0
        u = np.where(self.pad, n[np.clip(idx, 0, None)], 0) / self.U
        kyv = np.arange(-kmax, kmax + 1) / self.L
        kxv = np.arange(0, kmax + 1) / self.L
        self.Ey = _cplx(np.exp(2j * np.pi * u[:, :, None] * kyv[None, None]) * self.pad[:, :, None], dev)
        self.ExT = _cplx((np.exp(2j * np.pi * u[:, :, None] * kxv[None, None]) * self.pad[:, :, None])
                         .transpose(0, 2, 1) / self.L ** 2, dev)
        self.nbins = self.U * self.U * self.npts * self.npts
        self.lo, self.hi = t.bounds()
        self.evals = 0

    # bins by class
    def to_classes(self, grid: torch.Tensor):
        """(Ms, Ms) device bins (hole centre in the middle) -> (U, U, npts, npts) by sub-pixel class,
        0 where padded; and the mask of real bins (device bool)."""
        pos = grid.shape[-1] // 2 - self.nmax + np.clip(self.idx, 0, None)
        py = torch.from_numpy(pos[:, None, :, None]).to(grid.device)
        px = torch.from_numpy(pos[None, :, None, :]).to(grid.device)
        mask = torch.from_numpy(self.pad[:, None, :, None] & self.pad[None, :, None, :]).to(grid.device)
        return torch.where(mask, grid[py, px], torch.zeros((), dtype=grid.dtype, device=grid.device)), mask

    # the model
    def geometry(self, theta: torch.Tensor):
        """The outline of every row of theta (rows, P) (device): geo (4, rows, n) = [x, y, dx/dphi,
        dy/dphi], gk (rows, n, n_shape) = r dr/d(shape parameter), area (rows)."""
        rows, n = int(theta.shape[0]), self.n_poly
        geo = torch.empty((4, rows, n), dtype=torch.float64, device=self.device)
        gk = torch.empty((rows, n, self.n_shape), dtype=torch.float64, device=self.device)
        area = torch.empty(rows, dtype=torch.float64, device=self.device)
        ks = (ctypes.c_int32 * len(self.ks))(*self.ks)
        _lib.call("spx_onepass_lm_geometry", _p(theta), rows, self.P, n, len(self.ks), ks, _p(geo), _p(gk), _p(area),
                  stream_handle())
        return geo, gk, area

    def outline_hat(self, theta: torch.Tensor, grads: bool) -> torch.Tensor:
        """O^ and (grads) its shape derivatives at the band: (rows, F, 1 + K, 2) FP64."""
        geo, gk, area = self.geometry(theta)
        rows = int(theta.shape[0])
        K = self.n_shape if grads else 0
        out = torch.empty((rows, self.F, 1 + K, 2), dtype=torch.float64, device=self.device)
        _lib.call("spx_onepass_outline_hat", _p(geo), _p(gk), _p(area), rows, self.n_poly, K,
                  2.0 * math.pi / self.n_poly, _p(self.fx), _p(self.fy), self.F, _p(out), stream_handle())
        return out

    def coeff(self, theta: torch.Tensor, blur: torch.Tensor, tm: torch.Tensor | None) -> torch.Tensor:
        """Band coefficients (B, 1 + nf, F, 2) of the template and its free-direction derivatives
        (tm (P, nf) device, or None); blur (B, n_blur) the rows' blur parts."""
        B = int(theta.shape[0])
        grads = tm is not None
        ohat = self.outline_hat(theta, grads)
        nf = int(tm.shape[1]) if grads else 0
        out = torch.empty((B, 1 + nf, self.F, 2), dtype=torch.float64, device=self.device)
        _lib.call("spx_onepass_coeff", _p(ohat), self.n_shape if grads else 0, _p(blur), self.kind, self.n_blur,
                  _p(tm), nf, _p(self.fx), _p(self.fy), _p(self.pix), B, self.F, _p(out), stream_handle())
        return out

    def synth(self, coeff: torch.Tensor, gamma: torch.Tensor, cls_sel: torch.Tensor | None = None) -> torch.Tensor:
        """The stack's model by class: (B, C, slots, npts, npts), slots = U^2 (or the page's class)."""
        B, C = int(coeff.shape[0]), int(coeff.shape[1])
        slots = 1 if cls_sel is not None else self.U * self.U
        M = torch.empty((B, C, slots, self.npts, self.npts), dtype=torch.float64, device=self.device)
        _lib.call("spx_onepass_synth", _p(coeff), B, C, self.F, _p(self.fmap), _p(self.dbl), _p(gamma), self.U,
                  _p(cls_sel), _p(self.Ey), _p(self.ExT), self.npts, self.nky, self.nkx, _p(M), stream_handle())
        return M

    def evaluate(self, th: np.ndarray, data: dict, tm: np.ndarray | None = None):
        """Host (cost, scale, const[, g (B, nf), A (B, nf, nf)]) of parameter vectors th (B, P):
        scale and constant solved exactly, their directions projected off the derivatives (no
        priors: the fit adds them)."""
        self.evals += 1
        B = data["B"]
        theta = _dev(np.asarray(th, dtype=np.float64), self.device)
        blur = theta[:, self.n_shape:].contiguous()
        tmd = None if tm is None else _dev(tm, self.device)
        ev = self._evaluate_rows(theta, blur, tmd, data)
        o = ev.cpu().numpy()
        if tm is None:
            return o[:, 0], o[:, 1], o[:, 2]
        nf = tm.shape[1]
        return o[:, 0], o[:, 1], o[:, 2], o[:, 3:3 + nf], o[:, 3 + nf:].reshape(B, nf, nf)

    def _evaluate_rows(self, theta, blur, tmd, data, out=None) -> torch.Tensor:
        """spx_onepass_project's (B, 3 + nf + nf^2) for device rows theta (B, P): the outline, the
        band coefficients, the model by class, the variable projection."""
        B = data["B"]
        M = self.synth(self.coeff(theta, blur, tmd), data["gamma"])
        C = int(M.shape[1])
        nf = C - 1
        if out is None:
            out = torch.empty((B, 3 + nf + nf * nf), dtype=torch.float64, device=self.device)
        _lib.call("spx_onepass_project", _p(M), _p(data["y"]), _p(data["w"]), B, C, self.nbins, _p(out),
                  stream_handle())
        return out

    def prepare(self, gamma: torch.Tensor, y: torch.Tensor, w: torch.Tensor) -> dict:
        B = int(y.shape[0])
        return dict(gamma=gamma.contiguous(), y=y.reshape(B, -1).contiguous(), w=w.reshape(B, -1).contiguous(), B=B)

    # fitting
    def grid_start(self, base: np.ndarray, data: dict, keep: list | None = None) -> np.ndarray:
        """The round stage's grid of starts (onepass.py:422-486): 15 outline sizes x 12 blur widths
        compared on each page's most populated sub-pixel class; the cheapest per page."""
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
        U2, nb = self.U * self.U, self.npts * self.npts
        wsum = data["w"].reshape(B, U2, nb).sum(dim=2)
        top = torch.argmax(wsum, dim=1).to(torch.int32).contiguous()
        tha_d = _dev(tha.reshape(B * 15, -1), self.device)
        ohat = self.outline_hat(tha_d, False)
        blur = _dev(ths[:, j:], self.device)
        coeff = torch.empty((B, 180, self.F, 2), dtype=torch.float64, device=self.device)
        _lib.call("spx_onepass_grid_coeff", _p(ohat), 15, _p(blur), 12, self.kind, self.n_blur, _p(self.fx),
                  _p(self.fy), _p(self.pix), B, self.F, _p(coeff), stream_handle())
        M = self.synth(coeff, data["gamma"], top)
        table = torch.empty((B, 180), dtype=torch.float64, device=self.device)
        _lib.call("spx_onepass_grid_costs", _p(M), _p(data["y"]), _p(data["w"]), _p(top), B, 180, U2, nb, _p(table),
                  stream_handle())
        self.evals += 1
        tab = table.cpu().numpy()
        out = base.copy()
        for b in range(B):
            ia, isg = divmod(int(np.argmin(tab[b])), 12)
            out[b, 0] = out[b, 1] = tha[b, ia, 0]
            out[b, j:] = ths[isg, j:]
            if keep is not None:
                keep.append(dict(sizes_px=np.exp(tha[b, :, 0]), widths_px=np.exp(ths[:, j]),
                                 cost=tab[b].reshape(15, 12).copy(), pick=(ia, isg)))
        return out

    def default_prior(self):
        """Weak Gaussian priors (chi-square units) as rows (L, mean, sd) of linear forms of theta: the
        second Gaussian's widths twice the first's, its weight 5 % (onepass.py:488-502)."""
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

    def fit(self, theta0, gamma, y, w, stages=("round", "super", "full"), grid=True, max_iter=(12, 20, 30), ftol=1e-7,
            xtol=1e-7, keep: dict | None = None):
        """Staged Levenberg-Marquardt for a batch of pages in lockstep (onepass.py:504-623): bounds as
        an active set, Nielsen's gain-ratio damping, weak priors on the second Gaussian. theta0 (B, P)
        host; gamma (B, U^2, nky, nkx) complex, y, w (B, U, U, npts, npts) device, w in 1 / variance
        units. The iterations run on the device (spx_onepass_lm_*); the host reads one flag per
        iteration. Returns theta (B, P), scale, const, cost (host), [(stage, iterations, stopped (B,))]."""
        data = self.prepare(gamma, y, w)
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
            free, tm = stage_map(self.t, stage)
            base = theta.copy()
            if stage == "round":
                base[:, 1] = base[:, 0]
                base[:, 2] = 0.0
                base[:, 3] = 0.0
                base[:, 4:self.n_shape] = 0.0
                if grid:
                    base = self.grid_start(base, data, keep=None if keep is None else keep.setdefault("grid", []))
            lo, hi = self.lo[free] + 1e-6, self.hi[free] - 1e-6
            fixed = base * (1.0 - tm.sum(axis=1))[None, :]
            pj = PL @ tm
            pa = (pj.T * pw[None, :]) @ pj
            q0 = np.clip(base[:, free], lo, hi)
            st, buf = self._stage(B, fixed, tm, lo, hi, PL, pm, pw, pj, pa, q0, ftol, xtol)
            _lib.call("spx_onepass_lm_trial", ctypes.byref(st), 1, stream_handle())
            self._evaluate_rows(buf["theta"], buf["blur"], buf["tm"], data, out=buf["ev"])
            _lib.call("spx_onepass_lm_init", ctypes.byref(st), stream_handle())
            self.evals += 1
            its = 0
            for its in range(1, cap + 1):
                _lib.call("spx_onepass_lm_trial", ctypes.byref(st), 0, stream_handle())
                self._evaluate_rows(buf["theta"], buf["blur"], buf["tm"], data, out=buf["ev"])
                _lib.call("spx_onepass_lm_update", ctypes.byref(st), stream_handle())
                self.evals += 1
                if int(buf["alldone"].item()):
                    break
            q = buf["q"].cpu().numpy()
            theta = fixed + q @ tm.T
            cost, scale, const = (buf[k].cpu().numpy() for k in ("cost", "scale", "cst"))
            info.append((stage, its, buf["done"].cpu().numpy().astype(bool)))
        return theta, scale, const, cost, info

    def _stage(self, B, fixed, tm, lo, hi, PL, pm, pw, pj, pa, q0, ftol, xtol):
        """The device state of one stage: constants uploaded, work arrays allocated; returns the
        SpxLmStage and the tensors behind its pointers (kept alive by the caller)."""
        dev = self.device
        nf = tm.shape[1]
        npr = PL.shape[0]
        f64 = lambda a: _dev(np.asarray(a, dtype=np.float64), dev)  # noqa: E731
        z = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=dev)  # noqa: E731
        buf = dict(fixed=f64(fixed), tm=f64(tm), lo=f64(lo), hi=f64(hi), PL=f64(PL if npr else np.zeros((1, self.P))),
                   pm=f64(pm if npr else np.zeros(1)), pw=f64(pw if npr else np.zeros(1)),
                   pj=f64(pj if npr else np.zeros((1, nf))), pa=f64(pa), ev=z(B, 3 + nf + nf * nf), q=f64(q0),
                   qn=z(B, nf), theta=z(B, self.P), blur=z(B, self.n_blur), cost=z(B), scale=z(B), cst=z(B),
                   g=z(B, nf), A=z(B, nf, nf), lam=z(B), nu=z(B),
                   done=torch.zeros(B, dtype=torch.int32, device=dev),
                   alldone=torch.zeros(1, dtype=torch.int32, device=dev))
        st = _lib.SpxLmStage(B=B, P=self.P, nf=nf, nshape=self.n_shape, nb=self.n_blur, npr=npr,
                             ftol=float(ftol), xtol=float(xtol),
                             **{k: _p(v) for k, v in buf.items()})
        return st, buf


def second_is_narrower(t, theta) -> bool:
    """gauss2: the fitted second Gaussian is the narrower of the two (onepass.py:627-630)."""
    j = t.n_shape
    return bool(theta[j + 2] + theta[j + 3] < theta[j] + theta[j + 1])


def exchanged(t, theta) -> np.ndarray:
    """gauss2 with its two Gaussians exchanged (s1 <-> s2, w2 -> 1 - w2), inside the bounds
    (onepass.py:633-644)."""
    j = t.n_shape
    lo, hi = t.bounds()
    theta = np.asarray(theta, dtype=float)
    out = theta.copy()
    out[j:j + 2], out[j + 2:j + 4] = theta[j + 2:j + 4], theta[j:j + 2]
    out[j + 4] = -theta[j + 4]
    return np.clip(out, lo + 1e-6, hi - 1e-6)


# ----------------------------------------------------------------------------------------- page
_templates: dict = {}


def lib_template(blur: str, device):
    """The parameter bookkeeping template (outline_onepass.py:440-445), one per (blur, device)."""
    from .outline import OutlineTemplate

    key = (blur, str(device))
    if key not in _templates:
        _templates[key] = OutlineTemplate(upsample=8, size_px=24, blur=blur, device=device)
    return _templates[key]


class OnePassPage(HolePage):
    """One image through steps A to C (outline_onepass.py:520-935). The hole pass is HolePage's
    (spx_hole_sweep / spx_hole_render / spx_far_splat) with the fitted templates: this class sets
    every attribute that pass reads (img, h, w, n, c, a, level, delta, noise, spacing, R, the
    template and far-light fields of set_template) by the one-pass rules, so HolePage.__init__
    (its own spacing and start rules) is not called."""

    def __init__(self, image, seeds=None, *, blur: str = "gauss2", device: int = 0):
        resolve_backend("gpu")
        self.device = int(device)
        self.blur = blur
        self.U, self.pair_U, self.f_max = STACK_BINS, PAIR_BINS, F_MAX
        with gpu_device(self.device) as dev:
            self.dev = dev
            img = image if isinstance(image, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(image))
            self.img = img.to(device=dev, dtype=torch.float32).contiguous()
            self.h, self.w = (int(v) for v in self.img.shape)
            self.t = lib_template(blur, dev)
            if seeds is None:
                self.c = find_seeds_device(self.img)
            else:
                self.c = _dev(np.asarray(seeds, dtype=np.float64).reshape(-1, 2), dev)
            self.n = int(self.c.shape[0])
            if self.n < 2:
                raise ValueError(f"{self.n} holes found: no page fit")
            c = self.c.cpu().numpy()
            self.cells = CellList(c, self.h, self.w, 16.0, dev)
            nn = self.cells.nearest(self.c)
            found = nn < 1e20
            self.spacing = float(np.median(nn[found])) if found.any() else 16.0
            if self.spacing > 15.0:  # sparse holes: wider cells
                self.cells = CellList(c, self.h, self.w, 2.2 * self.spacing, dev)
                nn = self.cells.nearest(self.c)
                self.spacing = float(np.median(nn[nn < 1e20]))
            self.nearest_px = nn
            self.R = max(3, int(math.ceil(0.75 * self.spacing)))
            self.Ms = 2 * int(math.ceil((self.R + 0.5) * self.U)) + 1
            self.reach = float(self.R) + PAIR_REACH_MARGIN_PX
            self.L = int(8 * math.ceil((2 * self.reach + 8) / 8))
            self.a = torch.ones(self.n, dtype=torch.float64, device=dev)
            self.level = torch.zeros(self.n, dtype=torch.float64, device=dev)
            self.delta = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self._chunks = None
        self.tmpl = None

    # stack
    def interior(self) -> np.ndarray:
        m = self.R + 1
        c = self.c.cpu().numpy()
        return (c[:, 0] >= m) & (c[:, 0] <= self.w - 1 - m) & (c[:, 1] >= m) & (c[:, 1] <= self.h - 1 - m)

    def stack_of(self, use: np.ndarray):
        """Plain mean of the windows of the holes `use` on the quarter-pixel bins: (stack, pixels)
        device (Ms, Ms), NaN where no pixel fell (outline_onepass.py:582-592)."""
        Ms = self.Ms
        stack = torch.empty((Ms, Ms), dtype=torch.float64, device=self.dev)
        wsum = torch.empty((Ms, Ms), dtype=torch.float64, device=self.dev)
        u = _dev(np.asarray(use, dtype=np.uint8), self.dev, torch.uint8)
        _lib.call("spx_onepass_stack", _p(self.img), self.h, self.w, _p(self.c), self.n, _p(u), self.R, self.U, Ms,
                  _p(stack), _p(wsum), stream_handle())
        return stack, wsum

    def versus(self, stack: torch.Tensor, core_px: float) -> np.ndarray:
        """Per hole (n, 6) host: brightness, level, leftover rms, centre step dx, dy, pixels."""
        out = torch.empty((self.n, 6), dtype=torch.float64, device=self.dev)
        _lib.call("spx_onepass_versus", _p(self.img), self.h, self.w, _p(self.c), self.n, _p(stack), self.R, self.U,
                  self.Ms, float(core_px * core_px), _p(out), stream_handle())
        return out.cpu().numpy()

    def build_stack(self, core_frac: float = 0.5):
        """Stack of the interior holes; every hole against it; the stack again without the holes of
        unusual brightness or leftover; every hole against that (outline_onepass.py:594-625)."""
        with gpu_device(self.device):
            inside = self.interior()
            if not inside.any():
                raise ValueError(f"no hole lies {self.R + 1} px or more inside the image (its window): "
                                 "nothing to stack")
            self.core = max(2.0, core_frac * self.spacing)
            stack, _ = self.stack_of(inside)
            o = self.versus(stack, self.core)
            med = float(np.median(o[inside, 0]))
            a = np.maximum(o[:, 0] / med, 1e-3)
            typical = float(np.median(o[inside, 2]))
            use = inside & (a > 0.4) & (a < 2.5) & (o[:, 2] < 2.0 * typical) & (o[:, 5] > 0)
            if not use.any():
                raise ValueError("no hole of usual brightness and leftover to stack")
            self.a_host = a
            self.stack, self.wsum = self.stack_of(use)
            o = self.versus(self.stack, self.core)
            self.a_rel = np.maximum(o[:, 0], 1e-3)
            self.lev_rel = o[:, 1]
            self.rms = o[:, 2]
            self.noise = float(np.median(self.rms[use]))
            self.use = use

    # pair map
    def pair_map(self, sf: Fitter) -> torch.Tensor:
        """Gamma_q (U^2, nky, nkx) complex on the fitter's band (outline_onepass.py:663-698): the
        stacked holes' neighbours (brightness over the stacked mean, each hole's class weighted
        by 1 / its stacked holes) splatted per class, a real FFT, the crop to the band."""
        U, Ug = self.U, self.pair_U
        Ng = self.L * Ug
        with gpu_device(self.device) as dev:
            # bound to names: a temporary freed before the launch could hand its block to the next one
            a_d = _dev(self.a_host, dev)
            use_d = _dev(np.asarray(self.use, dtype=np.uint8), dev, torch.uint8)
            hist = torch.empty((U * U, Ng, Ng), dtype=torch.float64, device=dev)
            counts = torch.empty(U * U, dtype=torch.float64, device=dev)
            amean = float(self.a_host[self.use].mean())
            _lib.call("spx_onepass_pair_hist", _p(self.c), _p(a_d), amean, _p(use_d), self.n, U, _p(self.cells.order),
                      _p(self.cells.start), self.cells.ncx, self.cells.ncy, self.cells.cs, self.reach, Ug, Ng,
                      _p(hist), _p(counts), stream_handle())
            spec = torch.fft.rfft2(hist).contiguous()
            gamma = torch.empty((U * U, sf.nky, sf.nkx), dtype=torch.complex128, device=dev)
            _lib.call("spx_onepass_pair_crop", _p(spec), U * U, Ng, Ug, self.L, sf.kmax, _p(gamma), stream_handle())
            self.holes_per_class = counts.reshape(U, U)
            return gamma

    def stack_data(self, sf: Fitter):
        """The stack and its weights (1 / variance of a bin's mean) by class (outline_onepass.py:700-706)."""
        y, mask = sf.to_classes(torch.nan_to_num(self.stack, nan=0.0))
        w, _ = sf.to_classes(self.wsum)
        return y, torch.where(mask, w, torch.zeros((), dtype=w.dtype, device=w.device)) / self.noise ** 2

    def start_theta(self) -> np.ndarray:
        """A round first guess from the stack's second moment about the centre (filled bins within
        0.6 spacings, light above the stack's 10th percentile): half width from
        blob width^2 = blur width^2 + (outline / 1.1)^2 with a 1 px blur (outline_onepass.py:708-728)."""
        st = self.stack.cpu().numpy()
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

    # templates
    def _polygons(self, theta, n_poly, deltas):
        phi = 2.0 * np.pi * np.arange(n_poly) / n_poly
        r, dr = self.t.radius_at(theta, phi)
        out = []
        for d in deltas:
            rd = r + d * np.hypot(r, dr) / r if d else r
            out.append(np.stack([rd * np.cos(phi), rd * np.sin(phi)]))
        return np.stack(out)

    def band_grids(self, theta, box_px: int, f_max: float, upsample: int, n_poly: int, deltas) -> torch.Tensor:
        """Real (len(deltas), N, N) templates of theta's outline moved by each delta, N = box U:
        sum over the band of T^(f) exp(2 pi i f . u) / L^2 at u = (n - N / 2) / U (onepass.py:101-126)."""
        L, Uu = int(box_px), int(upsample)
        N = L * Uu
        kmax = int(math.floor(f_max * L))
        rows = len(deltas)
        xy = _dev(self._polygons(theta, n_poly, deltas), self.dev)
        blur = _dev(np.repeat(np.asarray(theta, float)[None, self.t.n_shape:], rows, axis=0), self.dev)
        full = torch.zeros((rows, N, N // 2 + 1), dtype=torch.complex128, device=self.dev)
        _lib.call("spx_onepass_band", _p(xy), rows, n_poly, _p(blur), BLUR_KIND[self.blur], self.t.n_blur, L, Uu, kmax,
                  0.0, _p(full), stream_handle())
        return torch.fft.irfft2(full, s=(N, N)) * (Uu * Uu)

    def set_template(self, theta, scale: float, const: float, far_box: int = 128, far_ramp: float = 2.0,
                     tU: int = 8):
        """Near templates [T0, dT/d delta, (1/2) d2T/d delta2, Tx, Ty] of theta on the fine grid with
        their reach, render half width and far kernel; every hole's brightness and level in the
        model's units (outline_onepass.py:733-793)."""
        self.theta = np.asarray(theta, float)
        tbox = max(32, int(8 * math.ceil(2 * (self.R + 6) / 8)))
        with gpu_device(self.device) as dev:
            h = 0.1
            t0g, tp, tm = self.band_grids(self.theta, tbox, 0.75, tU, 256, (0.0, h, -h))
            td = (tp - tm) / (2 * h)
            t2 = (tp + tm - 2 * t0g) / (2 * h * h)
            N = tbox * tU
            prof = t0g[N // 2].abs().cpu().numpy()
            above = np.nonzero(prof > 1e-3 * prof.max())[0]
            reach_px = float(above[-1] - above[0] + 1) / (2 * tU)
            Rr = min(int(math.ceil(reach_px)) + 1, N // (2 * tU) - 2, self.R + 3)
            xg = (torch.arange(N, device=dev, dtype=torch.float64) - N // 2) / tU
            rr = torch.hypot(xg[None, :], xg[:, None])
            r0 = float(Rr)
            wn = 0.5 * (1 + torch.cos(math.pi * torch.clamp((rr - r0) / far_ramp, 0.0, 1.0)))
            t0n, tdn, t2n = t0g * wn, td * wn, t2 * wn
            Rr = min(int(math.ceil(r0 + far_ramp)) + 1, N // (2 * tU) - 2)
            f1 = torch.fft.fftfreq(N, d=1.0 / tU, device=dev, dtype=torch.float64)
            sp = torch.fft.fft2(torch.fft.ifftshift(t0n))
            tx = torch.fft.fftshift(torch.fft.ifft2(sp * (2j * math.pi * f1[None, :])).real)
            ty = torch.fft.fftshift(torch.fft.ifft2(sp * (2j * math.pi * f1[:, None])).real)
            tmpl = torch.zeros((N, N, 8), dtype=torch.float32, device=dev)
            for k, gr in enumerate((t0n, tdn, t2n, tx, ty)):
                tmpl[..., k] = gr.to(torch.float32)
            self.tmpl, self.tN, self.tU, self.Rr = tmpl.contiguous(), N, tU, Rr
            self.r0, self.r_light = r0, r0 + far_ramp
            b0 = self.band_grids(self.theta, far_box, 0.49, 1, 128, (0.0,))[0]
            B = far_box
            xb = torch.arange(B, device=dev, dtype=torch.float64) - B // 2
            rb = torch.hypot(xb[None, :], xb[:, None])
            wf = 1.0 - 0.5 * (1 + torch.cos(math.pi * torch.clamp((rb - r0) / far_ramp, 0.0, 1.0)))
            c = self.c.cpu().numpy()
            out_x = max(0.0, -float(c[:, 0].min()), float(c[:, 0].max()) - (self.w - 1))
            out_y = max(0.0, -float(c[:, 1].min()), float(c[:, 1].max()) - (self.h - 1))
            my, mx = B // 2 + int(math.ceil(out_y)) + 4, B // 2 + int(math.ceil(out_x)) + 4
            hp, wp = _even_fast(self.h + 2 * my), _even_fast(self.w + 2 * mx)
            pad = torch.zeros((hp, wp), dtype=torch.float32, device=dev)
            pad[:B, :B] = (b0 * wf).to(torch.float32)
            self.far = dict(B=B, my=my, mx=mx, hp=hp, wp=wp, kf=torch.fft.rfft2(pad), kernel=pad[:B, :B])
            self.a = _dev(np.maximum(self.a_rel * scale, 1e-3), dev)
            self.level = _dev(self.a_rel * const + self.lev_rel, dev)
            self.delta = torch.zeros(self.n, dtype=torch.float64, device=dev)
            self.cd0 = self.t.cd_px(self.theta)

    def good_mask(self) -> np.ndarray:
        """Holes fit for the stack after the hole pass: window inside the page, usual brightness and
        leftover (outline_onepass.py:889-903)."""
        inside = self.interior()
        a = self.a.cpu().numpy()
        rms = self.hole_rms.cpu().numpy()
        med = np.median(a[inside])
        typical = np.median(rms[inside])
        return inside & (a > 0.4 * med) & (a < 2.5 * med) & (rms < 2.0 * typical)

    def restack(self, good: np.ndarray, scale: float):
        """The plain stack again from the hole pass's centres (outline_onepass.py:627-640)."""
        if not good.any():
            raise ValueError("no hole of usual brightness and leftover to stack after the hole pass")
        with gpu_device(self.device):
            self.model_a = self.a.cpu().numpy()
            self.a_host = np.maximum(self.model_a / scale, 1e-3)
            self.use = good
            self.stack, self.wsum = self.stack_of(good)
            self.a_rel = self.a_host
            self.noise = float(np.median(self.hole_rms.cpu().numpy()[good]))


# --------------------------------------------------------------------------------------- driver
def _fitter_key(pg: OnePassPage):
    return (pg.blur, pg.R, pg.U, pg.f_max, pg.L, pg.device)


def page_fits(pages, theta0=None, stages=("round", "super", "full"), grid=True, max_iter=(12, 20, 30), ftol=1e-7,
              cache=None):
    """The page fits of several pages, grouped by window size, each group in lockstep
    (outline_onepass.py:940-990). Per page (theta, scale, const, record)."""
    cache = {} if cache is None else cache
    out = [None] * len(pages)
    groups: dict = {}
    for i, pg in enumerate(pages):
        groups.setdefault(_fitter_key(pg), []).append(i)
    for key, members in groups.items():
        pg0 = pages[members[0]]
        with gpu_device(pg0.device) as dev:
            if key not in cache:
                cache[key] = Fitter(pg0.t, box_px=pg0.L, f_max=pg0.f_max, U=pg0.U, R=pg0.R, device=dev)
            sf = cache[key]
            gam, ys, ws, th = [], [], [], []
            for i in members:
                pg = pages[i]
                gam.append(pg.pair_map(sf))
                y, w = pg.stack_data(sf)
                ys.append(y)
                ws.append(w)
                th.append(pg.start_theta() if theta0 is None else theta0[i])
            ys, ws, gam = torch.stack(ys), torch.stack(ws), torch.stack(gam)
            theta, scale, const, cost, info = sf.fit(np.stack(th), gam, ys, ws, stages=stages, grid=grid,
                                                     max_iter=max_iter, ftol=ftol)
            bins = (ws > 0).reshape(len(members), -1).sum(dim=1).cpu().numpy()
        for k, i in enumerate(members):
            out[i] = (theta[k], float(scale[k]), float(const[k]),
                      dict(chi2=float(cost[k]), bins=int(bins[k]), iterations=int(info[-1][1]),
                           converged=bool(info[-1][2][k])))
    return out


def in_order(pages, blur, fits, cache=None):
    """gauss2: pages whose final fit ended with the second Gaussian the narrower are fitted again from
    the exchanged labels, together; per page the end with the lower cost is kept
    (outline_onepass.py:993-1022). Sets ``exchanged`` in every record."""
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


def fit_images(images, seeds=None, *, blur: str = "gauss2", exchange: bool = True, device: int = 0, log=None):
    """Steps A to C for a list of images (bright holes; the page fits run in lockstep). ``seeds``: one
    (N, 2) array per image (None: the method's own seed finder). One dict per image, host arrays
    (outline_onepass.py:1038-1119): per hole ``centers_xy`` (after B), ``brightness``, ``level``,
    ``rms``, ``good``; per page ``theta``, ``scale``, ``const``, ``chi2``, ``bins``, ``iterations``,
    ``converged``, ``exchanged``, ``theta_first``, ``stack``, ``stack_weight``, ``upsample``,
    ``spacing_px``, ``window_px``, ``pair_reach_px``, ``noise_sd``, ``seconds``."""
    if blur not in BLUR_KIND or blur == "gauss1":
        raise ValueError("blur must be 'gauss2' or 'lens'")
    started = time.perf_counter()

    def say(text):
        if log:
            torch.cuda.synchronize()
            log(f"{text} ({time.perf_counter() - started:.2f} s)")

    pages = []
    for k, image in enumerate(images):
        pg = OnePassPage(image, None if seeds is None else seeds[k], blur=blur, device=device)
        pg.build_stack()
        pages.append(pg)
    say("A: seeds and stack")
    cache: dict = {}
    first = page_fits(pages, stages=FIRST_STAGES, max_iter=FIRST_ITERATIONS, ftol=FIRST_FTOL, cache=cache)
    say("A: first page fit")
    for pg, (theta, scale, const, _) in zip(pages, first):
        pg.set_template(theta, scale, const)
        pg.passes = pg.fit_holes(sweeps=CENTRING_SWEEPS, size=False)
        pg.restack(pg.good_mask(), scale)
    say("B: hole pass and stack")
    fits = page_fits(pages, theta0=[f[0] for f in first], stages=("full",), grid=False,
                     max_iter=(FINAL_ITERATIONS,), cache=cache)
    fits = in_order(pages, blur if exchange else None, fits, cache=cache)
    say("C: final page fit")
    out = []
    for pg, f0, (theta, scale, const, record) in zip(pages, first, fits):
        out.append(dict(
            centers_xy=pg.c.cpu().numpy(), brightness=pg.model_a.copy(), level=pg.level.cpu().numpy(),
            rms=pg.hole_rms.cpu().numpy(), good=pg.use.copy(), theta=np.asarray(theta, float), scale=scale,
            const=const, theta_first=np.asarray(f0[0], float), stack=pg.stack.cpu().numpy(),
            stack_weight=pg.wsum.cpu().numpy(), upsample=pg.U, spacing_px=pg.spacing, window_px=pg.R,
            pair_reach_px=pg.reach, noise_sd=pg.noise, **record))
    seconds = (time.perf_counter() - started) / max(len(images), 1)
    for o in out:
        o["seconds"] = seconds
    return out
