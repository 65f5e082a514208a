"""Deblending: ``detect_centers(..., background="model")``.

Keeps the reference's behaviour (subpx/centers.py:180-194, 225-316; subpx/_cpu/deblend.py:
421-522, with the GPU twin subpx/_gpu/deblend.py): the image (holes made bright: ``invert``
negates it) is modelled as a smooth background map plus one 2-D Gaussian per detected hole;
background map and neighbour-cleaned Gaussian fits alternate (relaxation 0.6) until 99 % of the
centres and log-widths move less than 1e-3 (at most ``deblend_iterations``); every hole is then
refined again on its cell with the background and all other holes removed, by the same method
(logquad / weighted / none: the batch refinements of the GPU backend on the cell's bright core;
the Voronoi estimators on the cell near the fitted Gaussian). With ``tile_h`` the page is
deblended in bands of tile_h rows widened by ``overlap`` on each side, each hole taken from the
band that owns its row.

The pixel work is the library's (spx_deblend_init, spx_gauss_fit, spx_gauss_render,
spx_deblend_clean, spx_tile_medians, spx_tile_map, spx_cell_refine, spx_deblend_window,
spx_voronoi_labels / spx_voronoi_rois for the cells, spx_radial_symmetry /
spx_isophote_curvature); the host runs the iteration on per-hole parameter arrays: the status
rules, the free-pixel thresholds and the interpolation of tiles without free pixels.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .backends import stream_handle
from .types import CenterResult

DEBLEND_METHODS = ("weighted", "logquad", "logquadratic", "none", "radial_symmetry", "isophote_curvature")
FWHM_K = math.sqrt(2.0 * math.log(2.0))
RENDER_SIGMAS = 4.0
STATUS_NAMES = ("ok", "fit_diverged", "non_gaussian_fit", "center_outside_cell")


def _ptr(t):
    return None if t is None else int(t.data_ptr())


def sigma_major(p: np.ndarray) -> np.ndarray:
    """Long-axis sigma of Gaussians [A, x0, y0, u1, u2, u3]: 1 / sqrt of the smaller eigenvalue of
    their precision L L^T, L = [[e^u1, 0], [u2, e^u3]]."""
    p = np.asarray(p, dtype=np.float64).reshape(-1, 6)
    e1, u2, e3 = np.exp(p[:, 3]), p[:, 4], np.exp(p[:, 5])
    P = np.empty((len(p), 2, 2))
    P[:, 0, 0] = e1 * e1
    P[:, 0, 1] = P[:, 1, 0] = e1 * u2
    P[:, 1, 1] = u2 * u2 + e3 * e3
    lam = np.linalg.eigvalsh(P)[:, 0] if len(p) else np.zeros(0)
    return 1.0 / np.sqrt(np.maximum(lam, 1e-300))


def auto_tile(seeds: np.ndarray, shape) -> tuple[int, float]:
    """Tile side of about four pitches (the median nearest-seed distance), within [16, max side]."""
    from scipy.spatial import cKDTree

    if len(seeds) < 2:
        return int(max(shape)), float("nan")
    pitch = float(np.median(cKDTree(seeds).query(seeds, k=2)[0][:, 1]))
    return int(np.clip(round(4 * pitch), 16, max(shape))), pitch


def _tile_centres(tile: int, shape):
    H, W = shape
    yc = np.array([(j * tile + min((j + 1) * tile, H) - 1) / 2 for j in range(-(-H // tile))])
    xc = np.array([(i * tile + min((i + 1) * tile, W) - 1) / 2 for i in range(-(-W // tile))])
    return yc, xc


class _Cells:
    """A band's cells on the device: masks (N, Hm, Wm) u8, origins (N, 2) int32, mask sizes."""

    def __init__(self, img_u16: torch.Tensor, seeds: np.ndarray, pad: int, half: int):
        from .voronoi import _stack_shape, voronoi_labels_device

        dev = img_u16.device
        H, W = img_u16.shape
        self.n = len(seeds)
        rows = np.array([[x - half, y - half, 2 * half + 1, 2 * half + 1, x, y] for x, y in seeds],
                        dtype=np.float64).reshape(-1, 6)
        rows_t = torch.from_numpy(rows).to(dev)
        labels, _ = voronoi_labels_device(torch.from_numpy(np.ascontiguousarray(seeds)).to(dev), (H, W), areas=False)
        self.Hm, self.Wm = _stack_shape(rows, H, W, int(pad))
        fill = torch.empty((self.n, self.Hm, self.Wm), dtype=torch.float32, device=dev)
        self.masks = torch.empty((self.n, self.Hm, self.Wm), dtype=torch.uint8, device=dev)
        self.origins = torch.empty((self.n, 2), dtype=torch.int32, device=dev)
        _lib.call("spx_voronoi_rois", _ptr(img_u16), H, W, img_u16.stride(0), _ptr(labels), _ptr(rows_t), self.n,
                  int(pad), self.Hm, self.Wm, _ptr(fill), _ptr(self.masks), _ptr(self.origins), stream_handle())
        del fill, labels
        self.masks_host = self.masks.cpu().numpy().astype(bool)
        self.origins_host = self.origins.cpu().numpy().astype(np.int64)
        self.pixels = self.masks_host.reshape(self.n, -1).sum(axis=1)
        self.caps = np.maximum(1.0, np.sqrt(self.pixels / np.pi))


def _status(p: np.ndarray, cells: _Cells, sig: np.ndarray) -> np.ndarray:
    """Status codes (reference subpx/_cpu/deblend.py:262-277): 1 non-finite, 2 not a hole
    (A <= 0 or as wide as the cell), 3 centre off the cell, else 0."""
    n = len(p)
    h, w = cells.Hm, cells.Wm
    st = np.zeros(n, dtype=np.int32)
    bad = ~np.isfinite(p).all(axis=1)
    st[bad] = 1
    st[~bad & ((p[:, 0] <= 0) | (sig >= 0.999 * cells.caps))] = 2
    c = np.where(np.isfinite(p[:, 1:3]), p[:, 1:3], 0.0)
    iy = np.round(c[:, 1]).astype(int).clip(0, h - 1)
    ix = np.round(c[:, 0]).astype(int).clip(0, w - 1)
    inside = (p[:, 1] >= -0.5) & (p[:, 1] <= w - 0.5) & (p[:, 2] >= -0.5) & (p[:, 2] <= h - 0.5)
    inside &= cells.masks_host[np.arange(n), iy, ix]
    st[(st == 0) & ~inside] = 3
    return st


class _Deblender:
    """The iteration of one band (reference subpx/_cpu/deblend.py:421-522) on device arrays."""

    def __init__(self, signal: torch.Tensor, seeds: np.ndarray, cells: _Cells, *, max_iter: int = 30,
                 free_fraction: float = 0.05, free_distance_fwhm: float = 1.25):
        self.signal = signal
        self.dev = signal.device
        self.H, self.W = signal.shape
        self.seeds = seeds
        self.cells = cells
        self.max_iter = int(max_iter)
        self.free_fraction = float(free_fraction)
        self.free_distance_fwhm = float(free_distance_fwhm)
        self.tile = auto_tile(seeds, (self.H, self.W))[0]
        self.cap = int(max(cells.Hm, cells.Wm))
        n = len(seeds)
        px = np.stack([np.clip(np.rint(seeds[:, 0]), 0, self.W - 1), np.clip(np.rint(seeds[:, 1]), 0, self.H - 1)],
                      axis=1)
        self.dist = torch.empty((self.H, self.W), dtype=torch.float64, device=self.dev)
        px_t = torch.from_numpy(np.ascontiguousarray(px)).to(self.dev)  # held until the call's kernel has run
        _lib.call("spx_voronoi_labels", _ptr(px_t), n, self.H, self.W, None, None, _ptr(self.dist), stream_handle())
        torch.cuda.current_stream(self.dev).synchronize()
        del px_t
        yc, xc = _tile_centres(self.tile, (self.H, self.W))
        self.yc, self.xc = yc, xc
        self.yc_t = torch.from_numpy(yc).to(self.dev)
        self.xc_t = torch.from_numpy(xc).to(self.dev)
        self.model = torch.empty((self.H, self.W), dtype=torch.float64, device=self.dev)
        self.background = torch.empty_like(self.model)
        self.work = torch.empty((n, cells.Hm, cells.Wm), dtype=torch.float64, device=self.dev)
        self.fit_out = torch.empty((n, 6), dtype=torch.float64, device=self.dev)
        self.info = {}

    def fit(self, data: torch.Tensor, init: np.ndarray) -> np.ndarray:
        c = self.cells
        init_t = torch.from_numpy(np.ascontiguousarray(init, dtype=np.float64)).to(self.dev)
        _lib.call("spx_gauss_fit", _ptr(data), _ptr(c.masks), c.n, c.Hm, c.Wm, _ptr(init_t), self.max_iter,
                  _ptr(self.fit_out), stream_handle())
        return self.fit_out.cpu().numpy()

    def _global(self, p):
        g = p.copy()
        g[:, 1:3] += self.cells.origins_host
        return g

    def render(self, p, sig):
        self.pg = torch.from_numpy(self._global(p)).to(self.dev)
        self.radius = torch.from_numpy(np.clip(np.ceil(RENDER_SIGMAS * sig), 2, self.cap)).to(self.dev)
        _lib.call("spx_gauss_render", _ptr(self.pg), _ptr(self.radius), self.cells.n, self.H, self.W,
                  _ptr(self.model), stream_handle())

    def _tile_medians(self, model_below, dist_atleast):
        ny, nx = len(self.yc), len(self.xc)
        med = torch.empty((ny, nx), dtype=torch.float64, device=self.dev)
        cnt = torch.empty((ny, nx), dtype=torch.int64, device=self.dev)
        _lib.call("spx_tile_medians", _ptr(self.signal), _ptr(self.model), _ptr(self.dist), self.H, self.W,
                  self.tile, float(model_below), float(dist_atleast), _ptr(med), _ptr(cnt), stream_handle())
        return med.cpu().numpy(), cnt.cpu().numpy()

    def background_of(self, p, st, sig):
        """The background map from the residual's tile medians over free pixels (reference
        subpx/_cpu/deblend.py:323-393, 459-466)."""
        good = (st == 0) & (p[:, 0] > 0)
        amp = float(np.median(p[good, 0])) if good.any() else 1.0
        fwhm = float(np.median(2 * FWHM_K * sig[good])) if good.any() else 0.0
        free_med, free_cnt = self._tile_medians(self.free_fraction * amp, self.free_distance_fwhm * fwhm)
        measured = free_cnt >= max(16, int(0.1 * self.tile * self.tile))
        info = {"tiles": int(measured.size), "measured_tiles": int(measured.sum())}
        if measured.all():
            med, info["source"] = free_med, "measured"
        elif measured.any():
            from scipy.interpolate import LinearNDInterpolator
            from scipy.spatial import Delaunay, cKDTree

            grid = np.stack(np.meshgrid(self.yc, self.xc, indexing="ij"), axis=-1)
            known, vals = grid[measured], free_med[measured]
            try:
                tri = Delaunay(known) if len(vals) >= 3 else None
            except Exception:  # collinear measured tiles
                tri = None
            med = free_med.copy()
            missing = ~measured
            med[missing] = np.nan
            if tri is not None:
                med[missing] = LinearNDInterpolator(tri, vals)(grid[missing])
            still = missing & np.isnan(med)
            med[still] = vals[cKDTree(known).query(grid[still])[1]]
            info["source"] = "interpolated"
        else:
            med, _ = self._tile_medians(math.inf, -math.inf)
            info["source"] = "inferred"
            if np.isnan(med).all():
                resid = (self.signal - self.model).cpu().numpy()
                self.background.fill_(float(np.median(resid)))
                self.info = info
                return
            med = np.where(np.isnan(med), np.nanmedian(med), med)
        info["tile_medians"] = med
        self.info = info
        med_t = torch.from_numpy(np.ascontiguousarray(med)).to(self.dev)
        _lib.call("spx_tile_map", _ptr(med_t), _ptr(self.yc_t), _ptr(self.xc_t), len(self.yc), len(self.xc), self.H,
                  self.W, _ptr(self.background), stream_handle())

    def clean(self):
        c = self.cells
        _lib.call("spx_deblend_clean", _ptr(self.signal), _ptr(self.background), _ptr(self.model), _ptr(self.pg),
                  _ptr(self.radius), _ptr(c.masks), _ptr(c.origins), c.n, c.Hm, c.Wm, self.H, self.W,
                  _ptr(self.work), stream_handle())
        return self.work

    def run(self, iterations: int, relaxation: float = 0.6, tolerance: float = 1e-3) -> dict:
        c = self.cells
        n = c.n
        seeds_t = torch.from_numpy(np.ascontiguousarray(self.seeds)).to(self.dev)
        init_t = torch.empty((n, 6), dtype=torch.float64, device=self.dev)
        _lib.call("spx_deblend_init", _ptr(self.signal), self.H, self.W, _ptr(c.masks), _ptr(c.origins),
                  _ptr(seeds_t), n, c.Hm, c.Wm, _ptr(self.work), _ptr(init_t), stream_handle())
        params = self.fit(self.work, init_t.cpu().numpy())
        sig = sigma_major(params)
        status = _status(params, c, sig)
        self.render(params, sig)
        used, converged = 0, False
        for used in range(1, int(iterations) + 1):
            self.background_of(params, status, sig)
            fitted = self.fit(self.clean(), params)
            step = np.where(np.isfinite(fitted), fitted - params, 0.0)
            params = params + relaxation * step
            sig = sigma_major(params)
            status = _status(params, c, sig)
            self.render(params, sig)
            ok = (status == 0) & np.isfinite(fitted).all(axis=1)
            if not ok.any():
                break
            if np.percentile(np.abs(step[ok][:, 1:]).max(axis=1), 99) < tolerance:
                converged = True
                break
        self.background_of(params, status, sig)
        cleaned = self.clean()
        return dict(params=params, status=status, sigma=sig, cleaned=cleaned, iterations=used, converged=converged,
                    background_info=dict(self.info))


def deblend_band(img_u16: torch.Tensor, signal: torch.Tensor, seeds: np.ndarray, *, pad: int = 3,
                 iterations: int = 30):
    """Cells and deblended state of one band (reference subpx/centers.py:225-240)."""
    seeds = np.asarray(seeds, dtype=np.float64).reshape(-1, 2)
    pitch = auto_tile(seeds, tuple(signal.shape))[1]
    half = int(np.ceil(0.6 * pitch)) if np.isfinite(pitch) else 16
    cells = _Cells(img_u16, seeds, pad, half)
    st = _Deblender(signal, seeds, cells).run(iterations)
    return cells, st


def _band_centers(img_u16, signal, seeds, refine, *, pad, iterations, upsample_factor):
    """(centres (N, 2) band px, quality or None, cell areas, statuses, state) of one band."""
    from .voronoi import _estimate

    cells, st = deblend_band(img_u16, signal, seeds, pad=pad, iterations=iterations)
    dev = signal.device
    n = cells.n
    origins = cells.origins_host.astype(np.float64)
    quality = None
    if refine in ("radial_symmetry", "isophote_curvature"):
        p = st["params"]
        reach = np.where(np.isfinite(p).all(axis=1), 2 * FWHM_K * st["sigma"], np.inf)
        loc = torch.from_numpy(np.ascontiguousarray(seeds - origins)).to(dev)
        reach_t = torch.from_numpy(reach).to(dev)
        rois = torch.empty((n, cells.Hm, cells.Wm), dtype=torch.float32, device=dev)
        win = torch.empty((n, cells.Hm, cells.Wm), dtype=torch.uint8, device=dev)
        _lib.call("spx_deblend_window", _ptr(st["cleaned"]), _ptr(cells.masks), _ptr(loc), _ptr(reach_t), n, cells.Hm,
                  cells.Wm, int(pad), _ptr(rois), _ptr(win), stream_handle())
        local_t, q_t = _estimate(rois, win, refine, upsample_factor, None, 0.5)
        local, quality = local_t.cpu().numpy(), q_t.cpu().numpy()
    else:
        amp = torch.from_numpy(np.ascontiguousarray(st["params"][:, 0])).to(dev)
        out = torch.empty((n, 2), dtype=torch.float64, device=dev)
        method = 1 if refine in ("logquad", "logquadratic") else 0
        _lib.call("spx_cell_refine", _ptr(st["cleaned"]), _ptr(cells.masks), n, cells.Hm, cells.Wm, method, _ptr(amp),
                  0.2, _ptr(out), stream_handle())
        local = out.cpu().numpy()
    return local + origins, quality, cells.pixels.astype(np.int64), st["status"], st


def refine_deblended(img_t: torch.Tensor, result: CenterResult, *, refine: str, invert: bool, pad: int,
                     iterations: int, upsample_factor: int, tile_h, overlap: int) -> CenterResult:
    """Re-refine a detection's centres on deblended cells (reference subpx/centers.py:270-316)."""
    signal_full = img_t.to(torch.float64) * (-1.0 if invert else 1.0)
    seeds = np.asarray(result.centers_xy, dtype=np.float64).reshape(-1, 2)
    H = int(img_t.shape[0])
    bands = [(0, H)] if not tile_h else [(y0, min(H, y0 + int(tile_h))) for y0 in range(0, H, int(tile_h))]
    centers, qualities, areas, statuses, used, sources = [], [], [], [], 0, set()
    for y0, y1 in bands:
        lo, hi = (0, H) if not tile_h else (max(0, y0 - int(overlap)), min(H, y1 + int(overlap)))
        own = (seeds[:, 1] >= y0) & (seeds[:, 1] < y1)
        if not own.any():
            continue
        inside = (seeds[:, 1] >= lo) & (seeds[:, 1] < hi)
        local_seeds = seeds[inside] - [0, lo]
        found, quality, area, status, st = _band_centers(
            img_t[lo:hi], signal_full[lo:hi].contiguous(), local_seeds, refine, pad=int(pad),
            iterations=int(iterations), upsample_factor=int(upsample_factor))
        keep = own[inside]
        centers.append(found[keep] + [0, lo])
        areas.append(area[keep])
        statuses.append(status[keep])
        if quality is not None:
            qualities.append(np.asarray(quality, dtype=np.float64)[keep])
        used = max(used, st["iterations"])
        sources.add(st["background_info"].get("source"))
    centers = np.concatenate(centers) if centers else np.zeros((0, 2))
    finite = np.isfinite(centers).all(axis=1)
    status = np.concatenate(statuses) if statuses else np.zeros(0, np.int32)
    codes, counts = np.unique(status, return_counts=True)
    meta = dict(result.meta)
    meta.update(background="model", deblend_iterations=int(iterations), deblend_iterations_used=int(used),
                deblend_status_counts={STATUS_NAMES[int(c)]: int(k) for c, k in zip(codes, counts)},
                background_source=sorted(s for s in sources if s),
                deblend_dropped_nonfinite=int((~finite).sum()))
    if refine in ("radial_symmetry", "isophote_curvature"):
        key = "radial_symmetry_residual" if refine == "radial_symmetry" else "isophote_curvature_spread"
        meta[key] = (np.concatenate(qualities) if qualities else np.zeros(0))[finite]
        meta["voronoi_cell_area"] = (np.concatenate(areas) if areas else np.zeros(0, np.int64))[finite].astype(np.int64)
    return CenterResult(centers_xy=centers[finite], method=result.method, backend=result.backend, meta=meta)
