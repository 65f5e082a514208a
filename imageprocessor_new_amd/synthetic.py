"""Synthetic contact-hole arrays with exactly known parameters.

Conventions follow the reference generator (reference subpx/synthetic_arrays.py:22-24,
195-374): points are [x, y] pixels with pixel centres at integers; holes are bright on a
dark background; noise is photon (Poisson) noise plus Gaussian read noise, rounded to
integers. Unlike the reference (Fourier-domain rendering of a lens-blurred outline) each
hole here is the erf-edged disk or axis-aligned ellipse the fits model, point-sampled at
pixel centres in float64:

    I(x, y) = B + sum_h A_h * erfc((rho_h - r_h) / (sqrt(2) s_h)) / 2,
    rho_h = sqrt((x - x_h)^2 + ((y - y_h) / q_h)^2)   (q_h = 1 for disks)

so a fit's error against the truth is the method's own. Draw order from
``numpy.random.default_rng(seed)``: centre jitter (N, 2), radii, edge sigmas, ellipse
flags, axis ratios, then the noise rows top to bottom (Poisson, then Gaussian, per chunk).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np
from scipy.special import erfc


@dataclass
class HoleArray:
    image: np.ndarray          # (H, W) uint16
    centers_xy: np.ndarray     # (N, 2) float64 exact centres
    radius: np.ndarray         # (N,) x semi-axis
    sigma: np.ndarray          # (N,) edge blur sigma
    q: np.ndarray              # (N,) y/x axis ratio, 1 for disks
    ellipse: np.ndarray        # (N,) bool
    amplitude: float
    background: float
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def inside(self) -> np.ndarray:
        """Holes whose whole disk plus 3 sigma lies inside the image."""
        H, W = self.image.shape
        reach = self.radius + 3.0 * self.sigma
        x, y = self.centers_xy[:, 0], self.centers_xy[:, 1]
        return (x - reach >= 0) & (x + reach <= W - 1) & (y - reach >= 0) & (y + reach <= H - 1)


def square_lattice(H: int, W: int, pitch: float, origin: float | None = None) -> np.ndarray:
    o = pitch / 2.0 if origin is None else origin
    xs = np.arange(o, W, pitch)
    ys = np.arange(o, H, pitch)
    gx, gy = np.meshgrid(xs, ys)
    return np.stack([gx.ravel(), gy.ravel()], axis=1).astype(np.float64)


def hex_lattice(H: int, W: int, pitch: float, origin: float | None = None) -> np.ndarray:
    """Staggered array: rows pitch*sqrt(3)/2 apart, odd rows shifted by pitch/2."""
    o = pitch / 2.0 if origin is None else origin
    dy = pitch * np.sqrt(3.0) / 2.0
    pts = []
    for i, y in enumerate(np.arange(o, H - o / 2, dy)):
        xs = np.arange(o + (pitch / 2.0 if i % 2 else 0.0), W - o / 2, pitch)
        pts.append(np.stack([xs, np.full_like(xs, y)], axis=1))
    return np.concatenate(pts, axis=0).astype(np.float64)


def render_clean(shape, centers, radius, sigma, q, amplitude, background) -> np.ndarray:
    """Noise-free float64 image of the holes (exact model, point-sampled)."""
    H, W = shape
    img = np.full((H, W), float(background), dtype=np.float64)
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 2)
    n = centers.shape[0]
    if n == 0:
        return img
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (n,))
    sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n,))
    q = np.broadcast_to(np.asarray(q, dtype=np.float64), (n,))
    amp = np.broadcast_to(np.asarray(amplitude, dtype=np.float64), (n,))
    reach = np.ceil(radius * np.maximum(q, 1.0) + 9.0 * sigma).astype(np.int64) + 1
    for w in np.unique(reach):
        ids = np.nonzero(reach == w)[0]
        off = np.arange(-w, w + 1)
        for c0 in range(0, ids.size, 2048):
            blk = ids[c0:c0 + 2048]
            ix = np.rint(centers[blk, 0]).astype(np.int64)
            iy = np.rint(centers[blk, 1]).astype(np.int64)
            dx = (ix[:, None] + off[None, :]) - centers[blk, 0:1]          # (b, S)
            dy = (iy[:, None] + off[None, :]) - centers[blk, 1:2]          # (b, S)
            rho = np.sqrt(dx[:, None, :] ** 2 + (dy[:, :, None] / q[blk, None, None]) ** 2)
            u = (rho - radius[blk, None, None]) / (np.sqrt(2.0) * sigma[blk, None, None])
            val = amp[blk, None, None] * 0.5 * erfc(u)                     # (b, S(y), S(x))
            for j, h in enumerate(blk):
                x0, y0 = ix[j] - w, iy[j] - w
                xa, ya = max(x0, 0), max(y0, 0)
                xb, yb = min(x0 + 2 * w + 1, W), min(y0 + 2 * w + 1, H)
                if xa >= xb or ya >= yb:
                    continue
                img[ya:yb, xa:xb] += val[j, ya - y0:yb - y0, xa - x0:xb - x0]
    return img


def add_noise(clean: np.ndarray, rng: np.random.Generator, read_sigma: float = 30.0,
              poisson: bool = True, chunk_rows: int = 1024) -> np.ndarray:
    """Poisson(I) + N(0, read_sigma^2), rounded half-to-even, clipped to uint16."""
    H, W = clean.shape
    out = np.empty((H, W), dtype=np.uint16)
    for y0 in range(0, H, chunk_rows):
        blk = clean[y0:y0 + chunk_rows]
        v = rng.poisson(np.maximum(blk, 0.0)).astype(np.float64) if poisson else blk.copy()
        if read_sigma > 0:
            v += rng.normal(0.0, read_sigma, size=blk.shape)
        out[y0:y0 + chunk_rows] = np.clip(np.rint(v), 0, 65535).astype(np.uint16)
    return out


def render_hole_array(shape, centers, *, radius, sigma, q=1.0, amplitude=3000.0,
                      background=500.0, seed=0, noise=True, read_sigma=30.0,
                      rng: np.random.Generator | None = None) -> HoleArray:
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 2)
    n = centers.shape[0]
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (n,)).copy()
    sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n,)).copy()
    q = np.broadcast_to(np.asarray(q, dtype=np.float64), (n,)).copy()
    clean = render_clean(shape, centers, radius, sigma, q, amplitude, background)
    rng = np.random.default_rng(seed) if rng is None else rng
    if noise:
        image = add_noise(clean, rng, read_sigma=read_sigma)
    else:
        image = np.clip(np.rint(clean), 0, 65535).astype(np.uint16)
    return HoleArray(image, centers, radius, sigma, q, q != 1.0, float(amplitude),
                     float(background), meta={"seed": seed, "noise": bool(noise)})


# Workloads of the contact-hole benchmarks (sizes in pixels).
WORKLOADS = {
    # 2048^2 square array, pitch 32, r = 8, s = 1.2 -> 4,096 holes (logquad detection).
    "square2048": dict(H=2048, W=2048, layout="square", pitch=32.0, r=(8.0, 8.0), s=(1.2, 1.2),
               ellipse_frac=0.0),
    # 8192^2 square array, pitch 32, r ~ U[6, 10], s ~ U[0.8, 1.6] -> 65,536 holes.
    "square8192": dict(H=8192, W=8192, layout="square", pitch=32.0, r=(6.0, 10.0), s=(0.8, 1.6),
               ellipse_frac=0.0),
    # 16384^2 staggered array, ~200k holes, r ~ U[4, 20], 10 % ellipses q ~ U[0.8, 1].
    "hex16384": dict(H=16384, W=16384, layout="hex", pitch=39.37, r=(4.0, 20.0), s=(0.8, 1.6),
               ellipse_frac=0.1),
}


def workload_holes(name: str, seed: int = 0, *, H: int | None = None, W: int | None = None,
                   jitter: float = 0.3) -> dict:
    """Hole parameters of one benchmark workload, drawn from ``default_rng(seed)`` in the order
    centre jitter, radii, edge sigmas, ellipse flags, axis ratios (the noise is drawn after
    them). Returns dict(H, W, centers (N, 2), radius, sigma, q, ellipse, rng)."""
    cfg = dict(WORKLOADS[name])
    H = int(H or cfg["H"])
    W = int(W or cfg["W"])
    lat = square_lattice(H, W, cfg["pitch"]) if cfg["layout"] == "square" else hex_lattice(H, W, cfg["pitch"])
    rng = np.random.default_rng(seed)
    n = lat.shape[0]
    centers = lat + rng.normal(0.0, jitter, size=(n, 2))
    radius = rng.uniform(cfg["r"][0], cfg["r"][1], size=n)
    sigma = rng.uniform(cfg["s"][0], cfg["s"][1], size=n)
    ell = rng.random(n) < cfg["ellipse_frac"]
    qv = rng.uniform(0.8, 1.0, size=n)
    q = np.where(ell, qv, 1.0)
    return dict(H=H, W=W, centers=centers, radius=radius, sigma=sigma, q=q, ellipse=ell, rng=rng, cfg=cfg)


def make_workload(name: str, seed: int = 0, *, H: int | None = None, W: int | None = None,
                  noise: bool = True, jitter: float = 0.3) -> HoleArray:
    """Render one benchmark workload on the host; H/W override the size (same pitch and draws)."""
    d = workload_holes(name, seed, H=H, W=W, jitter=jitter)
    arr = render_hole_array((d["H"], d["W"]), d["centers"], radius=d["radius"], sigma=d["sigma"], q=d["q"],
                            seed=None, noise=noise, rng=d["rng"])
    arr.ellipse = d["ellipse"]
    arr.meta.update(config=name, seed=seed, H=d["H"], W=d["W"], pitch=d["cfg"]["pitch"], layout=d["cfg"]["layout"])
    return arr


@dataclass
class DevicePage:
    """A workload rendered on the GPU: ``image`` (H, W) uint16 device tensor plus the exact hole
    parameters (host arrays)."""
    image: Any
    centers_xy: np.ndarray
    radius: np.ndarray
    sigma: np.ndarray
    q: np.ndarray
    ellipse: np.ndarray
    amplitude: float
    background: float
    meta: dict[str, Any] = field(default_factory=dict)

    def patch_half(self, lo: int = 7, hi: int = 31) -> np.ndarray:
        """Half side of each hole's fit patch: side = 2 ceil(r_max + 3 s) + 1 clamped to
        [2 lo + 1, 2 hi + 1] (15..63 px by default), r_max the larger semi-axis."""
        reach = self.radius * np.maximum(self.q, 1.0) + 3.0 * self.sigma
        return np.clip(np.ceil(reach), lo, hi).astype(np.int16)


def render_workload_device(name: str, seed: int = 0, *, H: int | None = None, W: int | None = None,
                           jitter: float = 0.3, amplitude: float = 3000.0, background: float = 500.0,
                           read_sigma: float = 30.0, device: int = 0) -> DevicePage:
    """The workload's holes (``workload_holes``, same draws as ``make_workload``) rendered on the
    GPU by spx_render_page: point-sampled erf disks / ellipses, Poisson(I) + N(0, read_sigma^2)
    noise from a counter-based stream keyed by ``seed`` (not numpy's noise stream)."""
    d = workload_holes(name, seed, H=H, W=W, jitter=jitter)
    n = d["centers"].shape[0]
    holes = np.column_stack([d["centers"], d["radius"], d["sigma"], d["q"], np.full(n, amplitude)])
    img = render_page(holes, d["H"], d["W"], background=background, read_sigma=read_sigma, seed=seed,
                      device=device)
    return DevicePage(img, d["centers"], d["radius"], d["sigma"], d["q"], d["ellipse"], float(amplitude),
                      float(background), meta=dict(config=name, seed=seed, H=d["H"], W=d["W"],
                                                   pitch=d["cfg"]["pitch"], layout=d["cfg"]["layout"]))


# ---------------------------------------------------------------------------------------
# Large pages (32768^2, ~1.1M holes): hole lists drawn here, pixels rendered on the GPU.

def mixed_page_holes(seed: int = 0, H: int = 32768, W: int = 32768, region: int = 2048,
                     pitch=(24.0, 48.0), r_frac=(0.2, 0.3), sigma=(0.8, 1.6), jitter: float = 0.3,
                     amplitude: float = 3000.0):
    """Hole list of a page tiled into ``region`` x ``region`` blocks, each a square or a
    staggered array with its own pitch in [pitch[0], pitch[1]] (the mixed large-page workload:
    32768^2, pitch 24-48 px, ~1.2M holes). The pitch is drawn so that the hole density
    1 / pitch^2 is uniform over its range (pitch = U[1/48^2, 1/24^2]^(-1/2)): a uniform pitch
    would put ~1.0M holes on the page. Holes keep half a pitch from their block's edges;
    r ~ U[r_frac] * pitch, sigma ~ U[sigma], centre jitter N(0, jitter^2). Returns (holes (N, 6)
    float32 [x, y, r, sigma, q = 1, A], layouts (nby, nbx) 0 square / 1 staggered, pitches
    (nby, nbx))."""
    rng = np.random.default_rng(seed)
    nby, nbx = -(-H // region), -(-W // region)
    layouts = rng.integers(0, 2, size=(nby, nbx))
    dens = rng.uniform(1.0 / pitch[1] ** 2, 1.0 / pitch[0] ** 2, size=(nby, nbx))
    pitches = 1.0 / np.sqrt(dens)
    out = []
    for by in range(nby):
        for bx in range(nbx):
            p = float(pitches[by, bx])
            h = min(region, H - by * region)
            w = min(region, W - bx * region)
            lat = square_lattice(h, w, p) if layouts[by, bx] == 0 else hex_lattice(h, w, p)
            keep = ((lat[:, 0] >= 0.5 * p) & (lat[:, 0] <= w - 0.5 * p) & (lat[:, 1] >= 0.5 * p)
                    & (lat[:, 1] <= h - 0.5 * p))
            lat = lat[keep] + np.array([bx * region, by * region], dtype=np.float64)
            n = lat.shape[0]
            c = lat + rng.normal(0.0, jitter, size=(n, 2))
            r = rng.uniform(r_frac[0], r_frac[1], size=n) * p
            s = rng.uniform(sigma[0], sigma[1], size=n)
            out.append(np.column_stack([c, r, s, np.ones(n), np.full(n, amplitude)]))
    holes = np.concatenate(out, axis=0).astype(np.float32)
    return holes, layouts, pitches


def render_page(holes, H: int, W: int, *, background: float = 500.0, read_sigma: float = 30.0,
                seed: int = 0, device: int = 0):
    """Render a hole list (N, 6) [x, y, r, sigma, q, A] into an (H, W) uint16 device tensor
    (spx_render_page: point-sampled erf disks, Poisson photon noise + read noise)."""
    import torch

    from . import _lib
    from .backends import gpu_device, stream_handle

    with gpu_device(device) as dev:
        h = torch.from_numpy(np.ascontiguousarray(holes, dtype=np.float32)).to(dev)
        img = torch.empty((int(H), int(W)), dtype=torch.uint16, device=dev)
        _lib.call("spx_render_page", int(seed) & 0xFFFFFFFFFFFFFFFF, int(h.data_ptr()), int(h.shape[0]), int(H),
                  int(W), float(background), float(read_sigma), int(img.data_ptr()), int(W), stream_handle())
        torch.cuda.current_stream().synchronize()
        return img
