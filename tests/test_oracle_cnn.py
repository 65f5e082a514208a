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


def _interp_reference(P):
    """The 32 x 32 bilinear resample of a square patch by torch's own routine (align_corners
    False: source = (dst + 0.5) * side / 32 - 0.5, clamped at the borders), an implementation
    independent of the oracle's loops."""
    t = torch.tensor(np.asarray(P, dtype=np.float64))[None, None]
    return torch.nn.functional.interpolate(t, size=(32, 32), mode="bilinear", align_corners=False)[0, 0].numpy()


def test_prepare_matches_library_bilinear_resample():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 4000, size=(160, 170)).astype(np.uint16)
    for seed, half in (((80.3, 70.6), 15), ((40.0, 50.0), 7), ((100.5, 90.4), 31), ((60.0, 64.0), 4)):
        x, norm = cnn_ref.prepare(img, seed, half)
        cx, cy, side, B0, A0 = norm
        P = img[int(cy) - half:int(cy) + half + 1, int(cx) - half:int(cx) + half + 1].astype(np.float64)
        assert P.shape == (side, side)
        ring = np.concatenate([P[0], P[-1], P[1:-1, 0], P[1:-1, -1]])
        assert B0 == ring.mean() and A0 == max(P.max() - B0, 1.0)
        np.testing.assert_allclose(x, (_interp_reference(P) - B0) / A0, rtol=0, atol=1e-12)


def test_prepare_reproduces_an_affine_ramp():
    # bilinear resampling is exact on affine data: at output (i, j) the patch coordinate is
    # (j + 0.5) side / 32 - 0.5 (clamped to [0, side - 1]); the ring mean of a ramp is its
    # centre value and its maximum lies half a side off in both directions
    H, W = 200, 220
    a, bx, by = 1000.0, 3.0, -2.0
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.rint(a + bx * xx + by * yy + 1000).astype(np.uint16)  # integer-valued, exact
    for (sx, sy), half in (((100.2, 90.7), 15), ((60.0, 60.0), 31), ((30.4, 150.0), 9)):
        x, norm = cnn_ref.prepare(img, (sx, sy), half)
        cx, cy = np.rint(sx), np.rint(sy)
        side = 2 * half + 1
        centre = a + 1000 + bx * cx + by * cy
        A0 = (abs(bx) + abs(by)) * half
        p = np.clip((np.arange(32) + 0.5) * side / 32.0 - 0.5, 0, side - 1)
        expect = (bx * (p[None, :] - half) + by * (p[:, None] - half)) / A0
        assert norm[3] == centre and norm[4] == A0
        np.testing.assert_allclose(x, expect, rtol=0, atol=1e-12)


def test_prepare_of_a_flat_patch_is_zero():
    img = np.full((64, 64), 777, np.uint16)
    x, norm = cnn_ref.prepare(img, (32.0, 30.0), 12)
    assert norm[3] == 777.0 and norm[4] == 1.0
    assert np.array_equal(x, np.zeros((32, 32)))
