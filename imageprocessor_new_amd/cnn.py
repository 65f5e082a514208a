"""Learned initial guess for the erf-disk fits: a small CNN run and trained on the GPU.

Network (task-defined; no reference counterpart): the fit patch of each feature resampled
to 32 x 32 and normalised (spx_cnn_prepare), four 3x3 stride-2 convolutions
1 -> 16 -> 32 -> 64 -> 64 with ReLU, then fc 256 -> 128 (ReLU) -> 7. Its outputs decode to
a start [x0, y0, r, sigma, A, B, q] for the LM (spx_cnn_forward). Every layer runs as a
BF16 GEMM on the matrix cores with FP32 accumulation; training (spx_cnn_train_step) is
mean squared error on normalised targets with Adam, all in hand-written HIP kernels.

Learned-start fitting (``fit_learned``): the proposal is refined by at most ``lm_iters``
LM steps; a feature whose refinement has not converged is refitted by the plain LM from
the moment start (the fallback), so every result meets the plain fit's criteria.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .backends import stream_handle

LAYERS = [(9 * 1, 16), (9 * 16, 32), (9 * 32, 64), (9 * 64, 64), (256, 128), (128, 7)]


def _p(t):
    return None if t is None else int(t.data_ptr())


def param_count() -> int:
    return int(_lib.load().spx_cnn_param_count())


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, in the library's flat layout (weights K x N, bias N)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (k, n) in enumerate(LAYERS):
        std = np.sqrt(2.0 / k) if i < len(LAYERS) - 1 else np.sqrt(1.0 / k)
        parts.append(rng.normal(0.0, std, size=(k, n)).astype(np.float32).ravel())
        parts.append(np.zeros(n, dtype=np.float32))
    theta = np.concatenate(parts)
    return theta


@dataclass
class TrainLog:
    losses: list = field(default_factory=list)
    seconds: float = 0.0
    samples: int = 0


class LearnedStart:
    """CNN parameters on a device plus the Adam state; propose() and train()."""

    def __init__(self, theta: np.ndarray | None = None, device: torch.device | str = "cuda", seed: int = 0):
        n = param_count()
        theta = init_params(seed) if theta is None else np.asarray(theta, dtype=np.float32)
        if theta.shape != (n,):
            raise ValueError(f"theta must have {n} values")
        self.device = torch.device(device)
        self.theta = torch.from_numpy(theta.copy()).to(self.device)
        self.theta_bf = torch.empty(n, dtype=torch.int16, device=self.device)
        self.adam_m = torch.zeros(n, dtype=torch.float32, device=self.device)
        self.adam_v = torch.zeros(n, dtype=torch.float32, device=self.device)
        self.step = 0
        self._sync_bf16()

    def _sync_bf16(self):
        _lib.call("spx_cnn_to_bf16", _p(self.theta), self.theta.numel(), _p(self.theta_bf), stream_handle())

    # ---- data ------------------------------------------------------------------------
    def prepare(self, img: torch.Tensor, seeds: torch.Tensor, half: torch.Tensor):
        """(x (K, 32, 32) BF16 bits as int16, norm (K, 5) f32) on the device."""
        K = int(seeds.shape[0])
        x = torch.empty((K, 32, 32), dtype=torch.int16, device=self.device)
        norm = torch.empty((K, 5), dtype=torch.float32, device=self.device)
        H, W = img.shape
        _lib.call("spx_cnn_prepare", _p(img), H, W, img.stride(0), _p(seeds), _p(half), K, _p(x), _p(norm),
                  stream_handle())
        return x, norm

    def encode(self, params: torch.Tensor, norm: torch.Tensor) -> torch.Tensor:
        K = int(params.shape[0])
        t = torch.empty((K, 7), dtype=torch.float32, device=self.device)
        _lib.call("spx_cnn_encode", _p(params), _p(norm), K, _p(t), stream_handle())
        return t

    # ---- inference ---------------------------------------------------------------------
    def forward_raw(self, x: torch.Tensor) -> torch.Tensor:
        K = int(x.shape[0])
        out = torch.empty((K, 7), dtype=torch.float32, device=self.device)
        _lib.call("spx_cnn_forward", _p(x), K, _p(self.theta), _p(self.theta_bf), None, _p(out), stream_handle())
        return out

    def propose(self, img: torch.Tensor, seeds: torch.Tensor, half: torch.Tensor) -> torch.Tensor:
        """Start parameters (K, 7) [x0, y0, r, sigma, A, B, q] for each feature (device): patch
        preparation, network and decoding in one launch (spx_cnn_propose)."""
        K = int(seeds.shape[0])
        out = torch.empty((K, 7), dtype=torch.float32, device=self.device)
        H, W = img.shape
        _lib.call("spx_cnn_propose", _p(img), H, W, img.stride(0), _p(seeds), _p(half), K, _p(self.theta),
                  _p(self.theta_bf), _p(out), stream_handle())
        return out

    def propose_two_step(self, img: torch.Tensor, seeds: torch.Tensor, half: torch.Tensor) -> torch.Tensor:
        """The same proposal by spx_cnn_prepare + spx_cnn_forward (the training path's steps)."""
        x, norm = self.prepare(img, seeds, half)
        K = int(seeds.shape[0])
        out = torch.empty((K, 7), dtype=torch.float32, device=self.device)
        _lib.call("spx_cnn_forward", _p(x), K, _p(self.theta), _p(self.theta_bf), _p(norm), _p(out),
                  stream_handle())
        return out

    # ---- training ----------------------------------------------------------------------
    def train_step(self, x: torch.Tensor, t: torch.Tensor, w: torch.Tensor | None = None, lr: float = 1e-3,
                   want_loss: bool = False) -> float | None:
        self.step += 1
        loss = ctypes.c_float(0.0)
        _lib.call("spx_cnn_train_step", _p(x), _p(t), _p(w), int(x.shape[0]), _p(self.theta), _p(self.theta_bf),
                  _p(self.adam_m), _p(self.adam_v), self.step, float(lr),
                  ctypes.byref(loss) if want_loss else None, stream_handle())
        return float(loss.value) if want_loss else None

    def grad(self, x: torch.Tensor, t: torch.Tensor, w: torch.Tensor | None = None):
        """(gradient (P,) f32 device tensor, loss) of the mean squared error on a batch."""
        g = torch.empty_like(self.theta)
        loss = ctypes.c_float(0.0)
        _lib.call("spx_cnn_grad", _p(x), _p(t), _p(w), int(x.shape[0]), _p(self.theta), _p(self.theta_bf), _p(g),
                  ctypes.byref(loss), stream_handle())
        return g, float(loss.value)

    def train(self, x: torch.Tensor, t: torch.Tensor, w: torch.Tensor | None = None, *, epochs: int = 20,
              batch: int = 4096, lr: float = 1e-3, seed: int = 0, log_every: int = 0) -> TrainLog:
        import time

        n = int(x.shape[0])
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        log = TrainLog()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            xs, ts = x[perm], t[perm]
            ws = w[perm] if w is not None else None
            last = None
            for b0 in range(0, n - batch + 1, batch):
                sl = slice(b0, b0 + batch)
                want = (b0 + 2 * batch > n)
                last = self.train_step(xs[sl].contiguous(), ts[sl].contiguous(),
                                       ws[sl].contiguous() if ws is not None else None, lr=lr, want_loss=want)
                log.samples += batch
            log.losses.append(last)
            if log_every and (ep + 1) % log_every == 0:
                print(f"[cnn] epoch {ep + 1}/{epochs} loss {last:.5f}", flush=True)
        torch.cuda.synchronize()
        log.seconds = time.perf_counter() - t0
        return log

    def state(self) -> np.ndarray:
        return self.theta.detach().cpu().numpy()

    def save(self, path) -> None:
        np.savez_compressed(path, theta=self.state(), step=np.int64(self.step))

    @classmethod
    def load(cls, path, device="cuda") -> "LearnedStart":
        with np.load(path, allow_pickle=False) as z:
            obj = cls(z["theta"], device=device)
            obj.step = int(z["step"])
        return obj
