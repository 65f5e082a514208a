"""Pins of the window log-quadratic / weighted-centroid oracle (oracle/logquad_window.py)."""

import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import logquad_window as lw


def _spot_page(centres, quad, *, shape=(96, 128), amp=900.0, base=(300.0, 0.35, -0.2), support=3.2):
    """Float page: plane background + exactly log-quadratic spots A exp(-q(dx, dy)), zero beyond
    `support` px, so every pixel of the background ring lies on the plane."""
    H, W = shape
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    g = base[0] + base[1] * xx + base[2] * yy
    labels = np.zeros(shape, np.int32)
    qa, qb, qc = quad
    for i, (x0, y0) in enumerate(centres):
        dx, dy = xx - x0, yy - y0
        spot = amp * np.exp(-(qa * dx * dx + qb * dy * dy + qc * dx * dy))
        spot[np.hypot(dx, dy) > support] = 0.0
        g += spot
        labels[spot > 0.3 * amp] = i + 1
    rows, ids = [], []
    for lab in range(1, len(centres) + 1):
        ys, xs = np.nonzero(labels == lab)
        rows.append([xs.min(), ys.min(), xs.max() - xs.min() + 1, ys.max() - ys.min() + 1, xs.mean(), ys.mean()])
        ids.append(lab)
    return g, labels, np.asarray(rows, np.float32), np.asarray(ids)


@pytest.mark.parametrize("quad", [(0.18, 0.18, 0.0), (0.25, 0.12, 0.0), (0.2, 0.15, 0.08)])
def test_exact_recovery_of_log_quadratic_spots(quad):
    rng = np.random.default_rng(4)
    centres = [(20 + 24 * i + rng.uniform(-0.5, 0.5), 20 + 28 * j + rng.uniform(-0.5, 0.5))
               for i in range(4) for j in range(3)]
    g, labels, rows, ids = _spot_page(centres, quad)
    c, ok = lw.refine_logquad_window(g, labels, rows, ids)
    assert ok.all()
    np.testing.assert_allclose(c, np.asarray(centres), atol=1e-7)


def test_flat_background_without_plane():
    centres = [(30.3, 40.8), (70.45, 41.1)]
    g, labels, rows, ids = _spot_page(centres, (0.2, 0.2, 0.0), base=(250.0, 0.0, 0.0))
    c, ok = lw.refine_logquad_window(g, labels, rows, ids, plane_bg=False)
    assert ok.all()
    np.testing.assert_allclose(c, np.asarray(centres), atol=1e-7)


def test_sloped_background_needs_the_plane():
    # without the plane a 0.35 count/px slope biases the centre; with it the fit is exact
    centres = [(30.3, 40.8)]
    g, labels, rows, ids = _spot_page(centres, (0.2, 0.2, 0.0), base=(250.0, 3.0, 0.0))
    c_plane, ok = lw.refine_logquad_window(g, labels, rows, ids)
    c_flat, ok2 = lw.refine_logquad_window(g, labels, rows, ids, plane_bg=False)
    assert ok.all() and ok2.all()
    assert np.abs(c_plane - centres).max() < 1e-7
    assert np.abs(c_flat - centres).max() > 1e-3


def test_convex_spot_is_refused_and_returns_centroid():
    # log-convex crater A exp(+0.1 r^2): H is positive definite, so the peak gate refuses it
    centres = [(30.3, 40.8)]
    g, labels, rows, ids = _spot_page(centres, (-0.1, -0.1, 0.0), amp=200.0)
    c, ok = lw.refine_logquad_window(g, labels, rows, ids)
    assert not ok[0]
    np.testing.assert_array_equal(c[0], rows[0, 4:6].astype(np.float64))


def test_lower_median_and_percentile_index_match_numpy():
    rng = np.random.default_rng(0)
    for n in range(1, 200):
        v = rng.integers(0, 50, n).astype(np.float64)
        assert lw.lower_median(v) == np.quantile(v, 0.5, method="lower")
        k = max(int(math.floor(0.99 * (n - 1.0))), 0)
        assert np.sort(v)[k] == np.quantile(v, 0.99, method="lower")
    assert lw.lower_median([]) == math.inf


def test_clipped_dilation_matches_maximum_filter_reflect():
    rng = np.random.default_rng(1)
    for shape, rad in [((13, 13), 2), ((9, 9), 1), ((7, 11), 3)]:
        m = rng.random(shape) < 0.15
        ref = ndimage.maximum_filter(m.astype(np.uint8), size=2 * rad + 1, mode="reflect") > 0
        np.testing.assert_array_equal(lw.dilate_clipped(m, rad), ref)


def test_plane_fit_matches_lstsq():
    rng = np.random.default_rng(2)
    x = rng.integers(4000, 4013, 60).astype(np.float64)
    y = rng.integers(8000, 8013, 60).astype(np.float64)
    g = rng.integers(400, 600, 60).astype(np.float64)
    a, b, c, ok = lw.plane_fit(x, y, g)
    ref = np.linalg.lstsq(np.c_[x, y, np.ones_like(x)], g, rcond=None)[0]
    assert ok
    pred = a * x + b * y + c
    np.testing.assert_allclose(pred, np.c_[x, y, np.ones_like(x)] @ ref, atol=1e-6)
    assert not lw.plane_fit(x[:2], y[:2], g[:2])[3]


def test_weighted_centroid_closed_forms():
    g = np.full((20, 20), 100, np.uint16)
    labels = np.zeros((20, 20), np.int32)
    # two-pixel blob: weights 300 and 100 above the border median 100
    g[8, 8], g[8, 9] = 400, 200
    labels[8, 8:10] = 1
    c, ok = lw.refine_weighted_rows(g, labels, [[8, 8, 2, 1, 1]], pad=3)
    assert ok[0]
    np.testing.assert_allclose(c[0], [8 + 100 / 400, 8.0])
    # blob at the background level: falls back to the mask mean
    g[8, 8], g[8, 9] = 100, 100
    c, ok = lw.refine_weighted_rows(g, labels, [[8, 8, 2, 1, 1]], pad=3)
    np.testing.assert_allclose(c[0], [8.5, 8.0])
    # symmetric 3x3 blob
    g[:] = 50
    g[4:7, 11:14] = [[60, 80, 60], [80, 120, 80], [60, 80, 60]]
    labels[:] = 0
    labels[4:7, 11:14] = 2
    c, ok = lw.refine_weighted_rows(g, labels, [[11, 4, 3, 3, 2]], pad=3)
    np.testing.assert_allclose(c[0], [12.0, 5.0])
