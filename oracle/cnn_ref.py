"""CPU reference of the learned-start CNN (forward, loss gradient, patch preparation).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). The network has no reference
counterpart; this restates the library's definition in float64 with torch on the CPU
(autograd for the gradient) and numpy, on the library's flat parameter layout:
layer l has a (K_l x N_l) weight matrix (row index (ky * 3 + kx) * Cin + c for a 3x3
convolution) followed by N_l biases; activations are NHWC.

``bf16_round`` reproduces the GPU's round-to-nearest-even BF16 so a test can feed both
sides the same operands. Pinned by central finite differences of the loss, torch's BF16
rounding, a hand-summed convolution output, and for ``prepare`` by torch's bilinear
resampling (align_corners=False), the exact resample of an affine ramp with its closed-form
ring mean and range, and a flat patch mapping to zero (tests/test_oracle_cnn.py).
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [(9 * 1, 16), (9 * 16, 32), (9 * 32, 64), (9 * 64, 64), (256, 128), (128, 7)]
CONV_CIN = [1, 16, 32, 64]


def bf16_round(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(a) -> np.ndarray:
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def unpack(theta):
    """List of (W (K, N), b (N,)) torch float64 tensors (views of a leaf when requires_grad)."""
    out = []
    o = 0
    for k, n in LAYERS:
        W = theta[o:o + k * n].reshape(k, n)
        o += k * n
        b = theta[o:o + n]
        o += n
        out.append((W, b))
    return out


def forward(theta: torch.Tensor, x: torch.Tensor, round_acts: bool = False) -> torch.Tensor:
    """x (B, 32, 32) float64 -> (B, 7). round_acts: BF16-round every hidden activation."""
    layers = unpack(theta)
    a = x.reshape(-1, 1, 32, 32)
    for l in range(4):
        W, b = layers[l]
        cin = CONV_CIN[l]
        cout = W.shape[1]
        w = W.reshape(3, 3, cin, cout).permute(3, 2, 0, 1)  # (cout, cin, ky, kx)
        a = F.relu(F.conv2d(a, w, b, stride=2, padding=1))
        if round_acts:
            a = torch.from_numpy(bf16_round(a.detach().numpy()).astype(np.float64))
    a = a.permute(0, 2, 3, 1).reshape(a.shape[0], -1)  # NHWC flatten
    W, b = layers[4]
    a = F.relu(a @ W + b)
    if round_acts:
        a = torch.from_numpy(bf16_round(a.detach().numpy()).astype(np.float64))
    W, b = layers[5]
    return a @ W + b


def loss_and_grad(theta_np, x_np, t_np, w_np=None):
    """Mean squared error (mean over B x 7 outputs) and its gradient, float64 autograd."""
    theta = torch.tensor(np.asarray(theta_np, dtype=np.float64), requires_grad=True)
    x = torch.tensor(np.asarray(x_np, dtype=np.float64))
    t = torch.tensor(np.asarray(t_np, dtype=np.float64))
    y = forward(theta, x)
    w = torch.ones_like(y) if w_np is None else torch.tensor(np.asarray(w_np, dtype=np.float64))
    loss = (w * (y - t) ** 2).sum() / y.numel()
    loss.backward()
    return float(loss), theta.grad.numpy()


def prepare(img, seed_xy, half):
    """Reference of spx_cnn_prepare for one feature: (x (32, 32) float64, norm (5,))."""
    g = np.asarray(img, dtype=np.float64)
    H, W = g.shape
    hs = max(1, min(int(half), 31))
    side = 2 * hs + 1
    cx = float(np.rint(np.float32(seed_xy[0])))
    cy = float(np.rint(np.float32(seed_xy[1])))
    rows = np.clip(np.arange(side) + int(cy) - hs, 0, H - 1)
    cols = np.clip(np.arange(side) + int(cx) - hs, 0, W - 1)
    P = g[rows[:, None], cols[None, :]]
    ring = np.concatenate([P[0, :], P[-1, :], P[1:-1, 0], P[1:-1, -1]])
    B0 = ring.mean()
    A0 = max(P.max() - B0, 1.0)
    scale = side / 32.0
    out = np.zeros((32, 32))
    for i in range(32):
        for j in range(32):
            px = min(max((j + 0.5) * scale - 0.5, 0.0), side - 1)
            py = min(max((i + 0.5) * scale - 0.5, 0.0), side - 1)
            x0 = min(int(px), max(side - 2, 0))
            y0 = min(int(py), max(side - 2, 0))
            fx, fy = px - x0, py - y0
            x1, y1 = min(x0 + 1, side - 1), min(y0 + 1, side - 1)
            v = (1 - fy) * ((1 - fx) * P[y0, x0] + fx * P[y0, x1]) + fy * ((1 - fx) * P[y1, x0] + fx * P[y1, x1])
            out[i, j] = (v - B0) / A0
    return out, np.array([cx, cy, side, B0, A0])
