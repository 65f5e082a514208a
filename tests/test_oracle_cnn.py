"""Pins of the CNN reference (oracle/cnn_ref.py): finite differences and BF16 rounding."""

import numpy as np
import torch

from oracle import cnn_ref


def _theta(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for k, n in cnn_ref.LAYERS:
        parts.append(rng.normal(0, np.sqrt(2.0 / k), size=k * n))
        parts.append(rng.normal(0, 0.05, size=n))
    return np.concatenate(parts)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = _theta(1)
    x = rng.normal(0, 1, size=(3, 32, 32))
    t = rng.normal(0, 1, size=(3, 7))
    loss, g = cnn_ref.loss_and_grad(theta, x, t)
    idx = np.concatenate([rng.choice(144, 4, replace=False),                       # conv1 weights
                          160 + rng.choice(4608, 4, replace=False),                # conv2 weights
                          160 + 4640 + 18496 + 36928 + rng.choice(32768, 4, replace=False),  # fc1
                          [len(theta) - 1]])                                        # last bias
    h = 1e-6
    for i in idx:
        tp = theta.copy(); tp[i] += h
        tm = theta.copy(); tm[i] -= h
        fd = (cnn_ref.loss_and_grad(tp, x, t)[0] - cnn_ref.loss_and_grad(tm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_round_matches_torch():
    v = np.random.default_rng(2).normal(0, 100, size=10000).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(cnn_ref.bf16_round(v), ref)


def test_conv_layout_against_direct_sum():
    # conv1 output pixel (0, 1) channel 3 computed by hand from the flat layout
    rng = np.random.default_rng(3)
    theta = _theta(3)
    x = rng.normal(0, 1, size=(1, 32, 32))
    W = theta[:144].reshape(9, 16)
    b = theta[144:160]
    acc = b[3]
    for ky in range(3):
        for kx in range(3):
            yy, xx = 2 * 0 + ky - 1, 2 * 1 + kx - 1
            if 0 <= yy < 32 and 0 <= xx < 32:
                acc += W[ky * 3 + kx, 3] * x[0, yy, xx]
    th = torch.tensor(theta)
    layers = cnn_ref.unpack(th)
    Wt = layers[0][0].reshape(3, 3, 1, 16).permute(3, 2, 0, 1)
    out = torch.nn.functional.conv2d(torch.tensor(x).reshape(1, 1, 32, 32), Wt, layers[0][1], stride=2, padding=1)
    assert abs(out[0, 3, 0, 1].item() - acc) < 1e-12
