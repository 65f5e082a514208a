"""Learned-start batched fitting: training data, training, and the fit with fallback.

* ``page_training_set``: whole pages of a benchmark workload rendered on the GPU (by default
  the 16384^2 staggered array: r U[4, 20], 10 % ellipses, ~200k holes per page) with seeds
  disjoint from the evaluation seeds; every hole's fit patch is labelled by the plain LM
  (moment start, 30 iterations, step below 1e-6 px, Voronoi-cell mask) and the converged
  fits become the CNN targets.
* ``render_tiles`` / ``training_set``: single-hole tiles (spx_render_tiles) with the hole
  distributions of the benchmark arrays (``SPEC_STAGGERED``, ``SPEC_SQUARE``), the cheaper
  data of the tests.
* ``fit_learned``: CNN proposal -> at most ``lm_iters`` LM steps from it -> features whose
  refinement did not converge are refitted by the plain LM from the moment start. Because
  a converged LM from either start sits at the same minimum of the same cost, every
  result meets the plain fit's criteria; the fallback count is reported.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .backends import stream_handle
from .cnn import LearnedStart
from .fit import cell_radius, fit_erf_disks_device, seed_neighbours_device

TILE = 64
COLS = 256

SPEC_STAGGERED = dict(layout=1, pitch=39.37, jitter=0.3, r=(4.0, 20.0), s=(0.8, 1.6), ellipse_frac=0.1)
SPEC_SQUARE = dict(layout=0, pitch=32.0, jitter=0.3, r=(6.0, 10.0), s=(0.8, 1.6), ellipse_frac=0.0)


def tile_spec(spec: dict, amplitude=3000.0, background=500.0, read_sigma=30.0) -> _lib.SpxTileSpec:
    return _lib.SpxTileSpec(int(spec["layout"]), COLS, float(spec["pitch"]), float(spec["jitter"]),
                            float(spec["r"][0]), float(spec["r"][1]), float(spec["s"][0]), float(spec["s"][1]),
                            float(spec["ellipse_frac"]), float(amplitude), float(background), float(read_sigma))


@dataclass
class Tiles:
    img: torch.Tensor        # (rows * TILE, COLS * TILE) uint16
    truth: torch.Tensor      # (n, 7) [x0, y0, r, s, A, B, q] of each centre hole
    seeds: torch.Tensor      # (n, 2) float32: rounded true centres
    half: torch.Tensor       # (n,) int16
    model: torch.Tensor      # (n,) uint8


def half_for(r: torch.Tensor, s: torch.Tensor, fixed: int | None = None) -> torch.Tensor:
    if fixed is not None:
        return torch.full_like(r, int(fixed), dtype=torch.int16)
    return torch.clamp(torch.ceil(r + 3.0 * s) + 2, 7, 31).to(torch.int16)


def render_tiles(n: int, seed: int, spec: dict, device, fixed_half: int | None = None) -> Tiles:
    rows = (n + COLS - 1) // COLS
    img = torch.empty((rows * TILE, COLS * TILE), dtype=torch.uint16, device=device)
    truth = torch.empty((n, 7), dtype=torch.float32, device=device)
    sp = tile_spec(spec)
    import ctypes

    _lib.call("spx_render_tiles", int(seed), int(n), TILE, ctypes.byref(sp), int(img.data_ptr()),
              img.stride(0), int(truth.data_ptr()), stream_handle())
    seeds = torch.round(truth[:, :2]).contiguous()
    half = half_for(truth[:, 2], truth[:, 3], fixed_half).contiguous()
    model = (truth[:, 6] != 1.0).to(torch.uint8).contiguous()
    return Tiles(img, truth, seeds, half, model)


def plain_fit(img, seeds, half, model, init=None, max_iter=30, step_tol=1e-6, neighbours=None, tile_waves=0):
    opts = _lib.fit_options(max_iter=max_iter, step_tol=step_tol, tile_waves=tile_waves)
    return fit_erf_disks_device(img, seeds, half, model, init, options=opts, neighbours=neighbours)


def training_set(net: LearnedStart, tiles: Tiles):
    """(x, targets, weights) on the device: inputs of every tile whose plain fit converged."""
    res = plain_fit(tiles.img, tiles.seeds, tiles.half, tiles.model)
    ok = res["status"] == 0
    x, norm = net.prepare(tiles.img, tiles.seeds, tiles.half)
    t = net.encode(res["params"], norm)
    w = torch.ones_like(t)
    w[:, 6] = torch.where(tiles.model.bool(), 1.0, 0.0)  # q is only learned on ellipses
    keep = torch.nonzero(ok).squeeze(1)
    return x[keep].contiguous(), t[keep].contiguous(), w[keep].contiguous()


@dataclass
class PageFeatures:
    """Fit inputs of every hole of a device page: seeds at the rounded true centres, patch half
    sides from the radius and edge width (side 2 ceil(r + 3 s) + 1 in 15..63), the model (ellipse
    holes take the 7-parameter fit) and the Voronoi neighbour lists."""
    seeds: torch.Tensor
    half: torch.Tensor
    model: torch.Tensor
    nbr: torch.Tensor


def page_features(page, device) -> PageFeatures:
    seeds = torch.from_numpy(np.rint(page.centers_xy).astype(np.float32)).to(device)
    half = torch.from_numpy(page.patch_half()).to(device)
    model = torch.from_numpy(page.ellipse.astype(np.uint8)).to(device)
    nbr = seed_neighbours_device(seeds, page.image.shape, cell_radius(int(half.max().item())))
    return PageFeatures(seeds, half, model, nbr)


def page_training_set(net: LearnedStart, n_patches: int = 1_000_000, seed: int = 10_000,
                      workload: str = "hex16384", device="cuda"):
    """(x, targets, weights, info) on the device from whole rendered pages of ``workload`` with
    page seeds seed, seed + 1, ...: the first ``n_patches`` holes whose plain fit converged."""
    from . import synthetic

    xs, ts, ws = [], [], []
    have, pages, total = 0, 0, 0
    k = seed
    while have < n_patches:
        page = synthetic.render_workload_device(workload, seed=k, device=torch.device(device).index or 0)
        f = page_features(page, device)
        res = plain_fit(page.image, f.seeds, f.half, f.model, neighbours=f.nbr)
        ok = torch.nonzero(res["status"] == 0).squeeze(1)
        x, norm = net.prepare(page.image, f.seeds, f.half)
        t = net.encode(res["params"], norm)
        w = torch.ones_like(t)
        w[:, 6] = torch.where(f.model.bool(), 1.0, 0.0)  # q is only learned on ellipses
        take = ok[: n_patches - have]
        xs.append(x[take].contiguous())
        ts.append(t[take].contiguous())
        ws.append(w[take].contiguous())
        have += int(take.numel())
        total += int(f.seeds.shape[0])
        pages += 1
        k += 1
        del page, f, res, x, norm, t, w
    info = {"pages": pages, "page_seeds": [seed, k - 1], "holes_seen": total, "samples": have,
            "workload": workload}
    return torch.cat(xs), torch.cat(ts), torch.cat(ws), info


@dataclass
class FitTiming:
    seconds: float
    propose_s: float = 0.0
    refine_s: float = 0.0
    fallback_s: float = 0.0
    n_fallback: int = 0
    extra: dict = field(default_factory=dict)


def fit_learned(net: LearnedStart, img, seeds, half, model, *, lm_iters: int = 4, step_tol: float = 1e-6,
                neighbours=None, tile_waves: int = 0, bucket_mask: int = 0):
    """Learned start + <= lm_iters LM steps, plain-LM fallback for the rest: the CNN proposal
    (spx_cnn_propose), then one fit launch (spx_fit_batch_cells with learned_iters) in which a
    feature that the residual judge does not accept after lm_iters attempts from the proposal is
    fitted again from the moment start by the same wave.

    Returns (result dict like fit_erf_disks_device plus ``fallback`` (K,) bool = iters >
    lm_iters, FitTiming). Synchronises the device. ``iters`` of a fallback feature counts the
    learned attempts plus the refit's.
    """
    sync = torch.cuda.synchronize
    sync()
    t0 = time.perf_counter()
    init = net.propose(img, seeds, half)
    sync()
    t1 = time.perf_counter()
    opts = _lib.fit_options(max_iter=30, step_tol=step_tol, learned_iters=int(lm_iters), tile_waves=tile_waves,
                            bucket_mask=bucket_mask)
    res = fit_erf_disks_device(img, seeds, half, model, init, options=opts, neighbours=neighbours)
    sync()
    t2 = time.perf_counter()
    counters = np.zeros(5, dtype=np.uint64)
    _lib.call("spx_fit_counters", int(counters.ctypes.data), 5, stream_handle())
    res["fallback"] = res["iters"] > int(lm_iters)
    return res, FitTiming(t2 - t0, t1 - t0, t2 - t1, 0.0, int(counters[4]))


def train_learned_start(n_tiles: int = 1 << 20, epochs: int = 20, batch: int = 4096, lr: float = 1e-3,
                        seed: int = 10_000, spec: dict | None = None, device="cuda", log_every: int = 0,
                        chunk: int = 1 << 18, mix: list | None = None):
    """Render training tiles (seeds from `seed` on), label them with the plain LM and train.

    ``mix``: optional list of (spec, fraction) pairs; default is the staggered-array
    distribution only. Returns (net, info dict).
    """
    net = LearnedStart(device=device, seed=seed)
    specs = mix if mix is not None else [(spec or SPEC_STAGGERED, 1.0)]
    xs, ts, ws = [], [], []
    t0 = time.perf_counter()
    k = 0
    for sp, frac in specs:
        todo = int(n_tiles * frac)
        while todo > 0:
            n = min(chunk, todo)
            tiles = render_tiles(n, seed + k, sp, device)
            x, t, w = training_set(net, tiles)
            xs.append(x)
            ts.append(t)
            ws.append(w)
            del tiles
            todo -= n
            k += 1
    x = torch.cat(xs)
    t = torch.cat(ts)
    w = torch.cat(ws)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log = net.train(x, t, w, epochs=epochs, batch=batch, lr=lr, seed=seed, log_every=log_every)
    info = {"tiles": n_tiles, "samples": int(x.shape[0]), "data_seconds": t_data, "train_seconds": log.seconds,
            "losses": log.losses, "epochs": epochs, "batch": batch, "lr": lr}
    return net, info


def train_on_pages(n_patches: int = 1_000_000, epochs: int = 20, batch: int = 4096, lr: float = 1e-3,
                   seed: int = 10_000, workload: str = "hex16384", device="cuda", log_every: int = 0):
    """The learned start of the uneven-batch workload: ``n_patches`` plain-LM-labelled patches
    from pages of ``workload`` rendered with page seeds from ``seed`` on (disjoint from the
    evaluation pages' seeds below 10,000), ``epochs`` of Adam at ``lr``. Returns (net, info)."""
    net = LearnedStart(device=device, seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, t, w, dinfo = page_training_set(net, n_patches=n_patches, seed=seed, workload=workload, device=device)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log = net.train(x, t, w, epochs=epochs, batch=batch, lr=lr, seed=seed, log_every=log_every)
    info = dict(dinfo, data_seconds=t_data, train_seconds=log.seconds, losses=log.losses, epochs=epochs,
                batch=batch, lr=lr)
    del x, t, w
    return net, info
