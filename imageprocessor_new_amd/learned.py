"""Learned-start batched fitting: training data, training, and the fit with fallback.

* ``render_tiles``: synthetic tiles on the GPU (spx_render_tiles) with the hole
  distributions of the benchmark arrays (``SPEC_STAGGERED``: r U[4, 20], 10 % ellipses on a
  39.37 px staggered array; ``SPEC_SQUARE``: r U[6, 10] on a 32 px square array).
* ``training_set``: the plain LM labels every tile (moment start, 30 iterations, step
  below 1e-6 px); converged fits become the CNN targets.
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
from .fit import fit_erf_disks_device

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


def plain_fit(img, seeds, half, model, init=None, max_iter=30, step_tol=1e-6):
    opts = _lib.fit_options(max_iter=max_iter, step_tol=step_tol)
    return fit_erf_disks_device(img, seeds, half, model, init, options=opts)


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
class FitTiming:
    seconds: float
    propose_s: float = 0.0
    refine_s: float = 0.0
    fallback_s: float = 0.0
    n_fallback: int = 0
    extra: dict = field(default_factory=dict)


def fit_learned(net: LearnedStart, img, seeds, half, model, *, lm_iters: int = 4, step_tol: float = 1e-6):
    """Learned start + <= lm_iters LM steps, plain-LM fallback for the rest.

    Returns (result dict like fit_erf_disks_device, FitTiming). Synchronises the device.
    """
    sync = torch.cuda.synchronize
    sync()
    t0 = time.perf_counter()
    init = net.propose(img, seeds, half)
    sync()
    t1 = time.perf_counter()
    res = plain_fit(img, seeds, half, model, init=init, max_iter=lm_iters, step_tol=step_tol)
    sync()
    t2 = time.perf_counter()
    bad = torch.nonzero(res["status"] != 0).squeeze(1)
    nb = int(bad.numel())
    if nb:
        sub = plain_fit(img, seeds[bad].contiguous(), half[bad].contiguous(), model[bad].contiguous(),
                        step_tol=step_tol)
        for k in ("params", "centers", "rms", "iters", "status"):
            res[k][bad] = sub[k]
    sync()
    t3 = time.perf_counter()
    return res, FitTiming(t3 - t0, t1 - t0, t2 - t1, t3 - t2, nb)


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
