"""Outline hole stage on the GPU: every hole's centre, brightness, level and edge offset against
the complete page model, one colour class at a time (block Gauss-Seidel).

    page model  P(p) = sum_i a_i [T0 + delta_i T1 + delta_i^2 T2](p - c_i) + far light
    hole i      data(p) = P(p) + level_i   on its (2R+1)^2 window

Mirrors the reference's hole pass of the one-pass outline fit, ``GpuPage.fit_holes``
(subpx/_gpu/outline_onepass.py:828-887), the GPU form of the library's
``PageFit.fit_holes`` (subpx/_outline/page.py:587-729):

- spacing = median nearest-hole distance, window R = max(3, ceil(0.75 spacing))
  (outline_onepass.py:535-548);
- colour classes: cells of 1.5 spacings, colour = cell parity + 4 x rank in the cell
  (outline_onepass.py:504-519, 808-826);
- per colour: ``spx_hole_step`` (normal equations + bounded Gauss-Newton step, the
  reference's hole_normal + hole_update) then ``spx_hole_render`` (old light out, new in);
- far light: the holes' brightness splatted with Keys weights (``spx_far_splat``) and
  convolved with the far kernel by an FFT, refreshed between sweeps.

The template is an erf-edged disk made on the device (``spx_hole_templates``), with the
near window, reach and far kernel of the reference's ``set_template``
(outline_onepass.py:730-793). The reference builds its templates from the fitted outline
and blur (subpx/_outline/onepass.py); that shape fit is not part of this module.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from . import _lib
from .backends import gpu_device, resolve_backend, stream_handle


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else int(t.data_ptr())


def template_box(R: int) -> int:
    """Template box in px for the window half width R (outline_onepass.py:734)."""
    return max(32, int(8 * math.ceil(2 * (R + 6) / 8)))


def near_reach(radius: float, sigma: float, R: int, U: int = 8, far_ramp: float = 2.0):
    """(N, r0, Rr, r_light) of set_template (outline_onepass.py:745-757) for the disk template:
    the reach is where the centre row of T0 falls below 1e-3 of its peak."""
    N = template_box(R) * U
    s2 = math.sqrt(2.0) * sigma
    prof = [0.5 * math.erfc((abs((j - N // 2) / U) - radius) / s2) for j in range(N)]
    top = max(prof)
    above = [j for j, v in enumerate(prof) if v > 1e-3 * top]
    reach_px = float(above[-1] - above[0] + 1) / (2 * U)
    Rr = min(int(math.ceil(reach_px)) + 1, N // (2 * U) - 2, R + 3)
    r0 = float(Rr)
    Rr = min(int(math.ceil(r0 + far_ramp)) + 1, N // (2 * U) - 2)
    return N, r0, Rr, r0 + far_ramp


def _even_fast(n: int) -> int:
    """Smallest even m >= n with prime factors 2, 3, 5, 7 only (FFT-friendly padded size)."""
    m = n + (n & 1)
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 2


def colour_classes(c: np.ndarray, h: int, w: int, spacing: float) -> list[np.ndarray]:
    """Index arrays of the colour classes in colour order: cells of 1.5 spacings, colour =
    (cell x parity) + 2 (cell y parity) + 4 (rank of the hole in its cell, by index)
    (outline_onepass.py:504-519, 808-826). Holes of one colour are >= 1.5 spacings apart."""
    cs = 1.5 * float(spacing)
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
    order = np.argsort(col, kind="stable")
    counts = np.bincount(col, minlength=int(col.max()) + 1)
    out, s = [], 0
    for k in counts:
        if k:
            out.append(order[s:s + int(k)])
        s += int(k)
    return out


@dataclass
class HolePassResult:
    """Hole state after the hole pass, NumPy, lengths in px: centres (n, 2) [x, y], brightness,
    level, edge offset, per-hole leftover rms of the last sweep, per-sweep log."""

    centers_xy: np.ndarray
    brightness: np.ndarray
    level: np.ndarray
    edge_offset_px: np.ndarray
    hole_rms: np.ndarray
    log: list
    meta: dict[str, Any] = field(default_factory=dict)


class HolePage:
    """One page through the hole stage (the reference's GpuPage hole pass,
    subpx/_gpu/outline_onepass.py:521-887, without its stack and page-fit steps)."""

    def __init__(self, image, seeds, *, noise: float, brightness=None, level=None, window_px: int | None = None,
                 device: int = 0):
        resolve_backend("gpu")
        self.device = int(device)
        with gpu_device(self.device) as dev:
            img = torch.as_tensor(np.asarray(image) if not isinstance(image, torch.Tensor) else image)
            if img.dim() != 2 or img.numel() == 0:
                raise ValueError("image must be a nonempty 2D array")
            self.img = img.to(device=dev, dtype=torch.float32).contiguous()
            self.h, self.w = (int(v) for v in self.img.shape)
            c = np.asarray(seeds, dtype=np.float64).reshape(-1, 2)
            if len(c) < 2 or not np.isfinite(c).all():
                raise ValueError("the hole pass needs at least 2 finite seeds")
            self.n = len(c)
            self.c = torch.from_numpy(c.copy()).to(dev)
            self.a = (torch.ones(self.n, dtype=torch.float64, device=dev) if brightness is None else
                      torch.as_tensor(np.asarray(brightness, dtype=np.float64)).to(dev).clone())
            self.level = (torch.zeros(self.n, dtype=torch.float64, device=dev) if level is None else
                          torch.as_tensor(np.asarray(level, dtype=np.float64)).to(dev).clone())
            self.delta = torch.zeros(self.n, dtype=torch.float64, device=dev)
            if not (np.isfinite(noise) and noise > 0):
                raise ValueError("noise must be a positive pixel noise sigma")
            self.noise = float(noise)
            self.spacing = self._spacing(c)
            self.R = int(window_px) if window_px else max(3, int(math.ceil(0.75 * self.spacing)))
        self._chunks = None
        self.tmpl = None

    # ------------------------------------------------------------------------------- geometry
    def _spacing(self, c: np.ndarray) -> float:
        """Median distance to the nearest other hole (outline_onepass.py:535-547): a first search
        within 48 px, repeated within 3.3 spacings when the holes are sparse (spacing > 15 px);
        holes with no neighbour found do not count, 16 px when none has one."""
        from .fit import seed_neighbours_device

        seeds = torch.from_numpy(c.astype(np.float32)).to(self.c.device)

        def median_nearest(radius):
            nb = seed_neighbours_device(seeds, (self.h, self.w), radius, 1).cpu().numpy()[:, 0]
            has = nb >= 0
            if not has.any():
                return None
            return float(np.median(np.hypot(c[has, 0] - c[nb[has], 0], c[has, 1] - c[nb[has], 1])))

        spacing = median_nearest(48.0)
        if spacing is None:
            return 16.0
        if spacing > 15.0:
            spacing = median_nearest(max(48.0, 3.3 * spacing)) or spacing
        return spacing

    def colours(self) -> list[torch.Tensor]:
        if self._chunks is None:
            c = self.c.cpu().numpy()
            self._chunks = [torch.from_numpy(ix.astype(np.int32)).to(self.c.device)
                            for ix in colour_classes(c, self.h, self.w, self.spacing)]
        return self._chunks

    # ------------------------------------------------------------------------------- template
    def set_template(self, radius: float, sigma: float, *, U: int = 8, far_ramp: float = 2.0,
                     far_box: int = 128):
        """Disk template of ``radius`` px and edge blur ``sigma`` px on the fine grid, its near window,
        render half width and far kernel (outline_onepass.py:730-793); edge offsets start at 0."""
        if not (radius > 0 and sigma > 0):
            raise ValueError("radius and sigma must be positive")
        N, r0, Rr, r_light = near_reach(float(radius), float(sigma), self.R, int(U), float(far_ramp))
        with gpu_device(self.device) as dev:
            tm = torch.empty(N * N * 8, dtype=torch.float32, device=dev)
            far = torch.empty((far_box, far_box), dtype=torch.float32, device=dev)
            _lib.call("spx_hole_templates", float(radius), float(sigma), N, int(U), r0, float(far_ramp),
                      int(far_box), _ptr(tm), _ptr(far), stream_handle())
            c = self.c.cpu().numpy()
            B = int(far_box)
            out_x = max(0.0, -float(c[:, 0].min()), float(c[:, 0].max()) - (self.w - 1))
            out_y = max(0.0, -float(c[:, 1].min()), float(c[:, 1].max()) - (self.h - 1))
            my, mx = B // 2 + int(math.ceil(out_y)) + 4, B // 2 + int(math.ceil(out_x)) + 4
            hp, wp = _even_fast(self.h + 2 * my), _even_fast(self.w + 2 * mx)
            pad = torch.zeros((hp, wp), dtype=torch.float32, device=dev)
            pad[:B, :B] = far
            self.far = dict(B=B, my=my, mx=mx, hp=hp, wp=wp, kf=torch.fft.rfft2(pad), kernel=far)
        self.tmpl, self.tN, self.tU, self.Rr = tm, N, int(U), Rr
        self.r0, self.r_light = r0, r_light
        self.cd0 = 2.0 * float(radius)
        self.delta.zero_()

    def far_field(self) -> torch.Tensor:
        """The holes' far light on the page (H, W) FP64: brightness x (1 + 4 delta / cd0) splatted
        with Keys weights and convolved with the far kernel (outline_onepass.py:795-806)."""
        f = self.far
        grid = torch.zeros((f["hp"], f["wp"]), dtype=torch.float32, device=self.c.device)
        _lib.call("spx_far_splat", _ptr(self.c), _ptr(self.a), _ptr(self.delta), float(self.cd0), self.n,
                  int(f["mx"]), int(f["my"]), int(f["hp"]), int(f["wp"]), _ptr(grid), stream_handle())
        out = torch.fft.irfft2(torch.fft.rfft2(grid) * f["kf"], s=(f["hp"], f["wp"]))
        oy, ox = f["my"] + f["B"] // 2, f["mx"] + f["B"] // 2
        return out[oy:oy + self.h, ox:ox + self.w].to(torch.float64)

    # ------------------------------------------------------------------------------ hole pass
    def fit_holes(self, sweeps: int = 2, size: bool = False, max_move: float = 0.25, robust_c: float = 3.0,
                  relax: float = 1.0, keep: dict | None = None, far: bool = True,
                  timing: dict | None = None) -> list:
        """Hole stage for a fixed number of sweeps (outline_onepass.py:828-887): centre, brightness
        and level of every hole (``size``: its edge offset too). Returns per-sweep statistics;
        ``keep`` receives ``light`` (the page model the last sweep ended with, device (H, W));
        ``timing`` receives ``sweep_ms``, the device time of each sweep's step + render launches."""
        if self.tmpl is None:
            raise RuntimeError("set_template must run before fit_holes")
        if int(sweeps) != sweeps or sweeps < 1:
            raise ValueError("sweeps must be a positive integer")
        chunks = self.colours()
        H, W, N, U, R, Rr = self.h, self.w, self.tN, self.tU, self.R, self.Rr
        P = (2 * Rr + 1) ** 2
        max_delta = 0.3 * 0.5 * self.cd0
        rl2 = float((self.r_light + 0.75) ** 2)
        sh = stream_handle
        with gpu_device(self.device) as dev:
            # holes in colour order: order[start[j]:start[j+1]] is colour j; c_old, the light cache
            # and the statistics follow the same order
            order = torch.cat(chunks).contiguous()
            start_np = np.cumsum([0] + [int(ix.shape[0]) for ix in chunks]).astype(np.int32)
            start = (ctypes.c_int32 * len(start_np))(*start_np.tolist())
            ncol = len(chunks)
            page = torch.zeros(H * W, dtype=torch.float64, device=dev)
            cache = torch.zeros(self.n * P, dtype=torch.float32, device=dev)
            c_old = self.c[order.long()].contiguous()
            # first light of every hole (the cache is empty, so nothing is taken out)
            _lib.call("spx_hole_render", _ptr(page), H, W, _ptr(cache), _ptr(self.tmpl), N, U, Rr, rl2, _ptr(order),
                      self.n, _ptr(c_old), _ptr(self.c), _ptr(self.a), _ptr(self.delta), sh())
            far_l = self.far_field().reshape(-1) if far else None
            if far:
                page += far_l
            rms_typ = self.noise
            a_med = float(np.median(self.a.cpu().numpy()))
            stat = torch.zeros((self.n, 4), dtype=torch.float64, device=dev)
            opt = _lib.SpxHoleStepOptions(nv0=self.noise ** 2, nv1=0.0, rscale=0.0, max_move=float(max_move),
                                          max_delta=float(max_delta), a_med=a_med, relax=float(relax),
                                          size=int(bool(size)), reserved=0)
            log = []
            for it in range(int(sweeps)):
                opt.rscale = float(robust_c) * max(rms_typ, 1e-6) if robust_c else 0.0
                if timing is not None:
                    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                    ev[0].record()
                _lib.call("spx_hole_sweep", _ptr(self.img), H, W, W, _ptr(page), _ptr(self.tmpl), N, U, R, Rr, rl2,
                          _ptr(order), start, ncol, _ptr(self.c), _ptr(self.a), _ptr(self.level), _ptr(self.delta),
                          _ptr(c_old), _ptr(cache), ctypes.byref(opt), _ptr(stat), sh())
                if timing is not None:
                    ev[1].record()
                    timing.setdefault("events", []).append(ev)
                if far and it + 1 < sweeps:
                    new_far = self.far_field().reshape(-1)
                    page += new_far - far_l
                    far_l = new_far
                rms_typ, m99, d99 = self._sweep_stats(stat)
                log.append(dict(sweep=it + 1, move99=m99, delta99=d99, rms=rms_typ))
            if timing is not None:
                torch.cuda.current_stream().synchronize()
                timing["sweep_ms"] = [e0.elapsed_time(e1) for e0, e1 in timing.pop("events", [])]
            self.hole_rms = torch.empty(self.n, dtype=torch.float64, device=dev)
            self.hole_rms[order.long()] = stat[:, 2]
            self.stat = stat
            if keep is not None:
                keep["light"] = page.reshape(H, W)
        return log

    def _sweep_stats(self, stat: torch.Tensor):
        """(median leftover rms, 99th percentile centre move, 99th percentile edge-offset move)
        of a sweep, by NumPy's rules (median: mean of the middle pair; percentile: linear
        interpolation), from six exact order statistics selected on the device."""
        n = self.n
        pos = 0.99 * (n - 1)
        lo, hi = int(math.floor(pos)), int(math.ceil(pos))
        t = pos - lo
        out = torch.empty(6, dtype=torch.float64, device=stat.device)
        cols = (ctypes.c_int64 * 6)(2, 2, 0, 0, 1, 1)
        ks = (ctypes.c_int64 * 6)((n - 1) // 2, n // 2, lo, hi, lo, hi)
        _lib.call("spx_select_kth", _ptr(stat), n, 4, cols, ks, 6, _ptr(out), stream_handle())
        v = out.cpu().tolist()

        def lerp(a, b):  # numpy's _lerp
            d = b - a
            return b - d * (1.0 - t) if t >= 0.5 else a + d * t

        return 0.5 * (v[0] + v[1]), lerp(v[2], v[3]), lerp(v[4], v[5])

    def result(self, log=None) -> HolePassResult:
        return HolePassResult(
            centers_xy=self.c.cpu().numpy(), brightness=self.a.cpu().numpy(), level=self.level.cpu().numpy(),
            edge_offset_px=self.delta.cpu().numpy(),
            hole_rms=self.hole_rms.cpu().numpy() if hasattr(self, "hole_rms") else np.full(self.n, np.nan),
            log=list(log or []),
            meta={"spacing_px": self.spacing, "window_px": self.R, "render_px": getattr(self, "Rr", None),
                  "colours": len(self.colours()), "template_N": getattr(self, "tN", None)})


def fit_holes(image, seeds, *, radius: float, sigma: float, noise: float, brightness=None, level=None,
              sweeps: int = 2, size: bool = False, max_move: float = 0.25, robust_c: float = 3.0, relax: float = 1.0,
              window_px: int | None = None, far: bool = True, device: int = 0) -> HolePassResult:
    """The hole stage on one page with an erf-disk template of ``radius`` and ``sigma`` px:
    centre, brightness and level of every hole (``size``: its edge offset too) after ``sweeps``
    Gauss-Seidel sweeps against the complete page model (outline_onepass.py:828-887)."""
    pg = HolePage(image, seeds, noise=noise, brightness=brightness, level=level, window_px=window_px, device=device)
    pg.set_template(radius, sigma)
    log = pg.fit_holes(sweeps=sweeps, size=size, max_move=max_move, robust_c=robust_c, relax=relax, far=far)
    return pg.result(log)


def render_model_page(centers, brightness, *, radius: float, sigma: float, shape, background: float = 500.0,
                      noise: float = 30.0, edge_offset=None, seed: int = 0, device: int = 0) -> torch.Tensor:
    """A synthetic page of the hole-stage model on the device: every hole's near light
    (spx_hole_render from an empty cache) and far light, plus a constant background and
    N(0, noise^2) pixel noise. Returns a float32 (H, W) device tensor. Synthetic input for
    tests and benchmarks."""
    h, w = int(shape[0]), int(shape[1])
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 2)
    pg = HolePage(torch.zeros((h, w), dtype=torch.float32, device=torch.device("cuda", int(device))), c,
                  noise=max(float(noise), 1e-3), brightness=brightness, device=device)
    pg.set_template(radius, sigma)
    if edge_offset is not None:
        pg.delta.copy_(torch.as_tensor(np.asarray(edge_offset, dtype=np.float64)))
    H, W, N, U, Rr = pg.h, pg.w, pg.tN, pg.tU, pg.Rr
    rl2 = float((pg.r_light + 0.75) ** 2)
    with gpu_device(pg.device) as dev:
        order = torch.arange(pg.n, dtype=torch.int32, device=dev)
        page = torch.zeros(H * W, dtype=torch.float64, device=dev)
        cache = torch.zeros(pg.n * (2 * Rr + 1) ** 2, dtype=torch.float32, device=dev)
        _lib.call("spx_hole_render", _ptr(page), H, W, _ptr(cache), _ptr(pg.tmpl), N, U, Rr, rl2, _ptr(order), pg.n,
                  _ptr(pg.c), _ptr(pg.c), _ptr(pg.a), _ptr(pg.delta), stream_handle())
        page = page.reshape(H, W) + pg.far_field() + float(background)
        if noise > 0:
            g = torch.Generator(device=dev)
            g.manual_seed(int(seed))
            page = page + float(noise) * torch.randn((H, W), dtype=torch.float64, device=dev, generator=g)
        return page.to(torch.float32)

