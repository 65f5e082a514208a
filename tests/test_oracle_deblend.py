"""Pins of the deblending oracle (oracle/deblend.py) by the reference's own known answers
(reference tests/test_deblend.py:46-57 and :177-192) and by calculus: the Gaussian Jacobian
against central differences, the precision's eigenvalue against the closed form, the tile-map
interpolation against the separable linear rule it must reproduce on planar tiles."""

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import deblend as db
from oracle import voronoi_refine as vr

FWHM, BACKGROUND, AMPLITUDE = 4.0, 20.0, 50.0


def lattice(ratio, shape=(120, 240), snr=None, seed=0):
    """Staggered Gaussian holes of FWHM 4 px at pitch ratio x FWHM, 15 px hole-free margin."""
    rng = np.random.default_rng(seed)
    pitch = ratio * FWHM
    sigma = FWHM / (2 * np.sqrt(2 * np.log(2)))
    xs = np.arange(15, shape[1] - 15, pitch)
    ys = np.arange(15, shape[0] - 15, pitch)
    xx, yy = np.meshgrid(xs, ys)
    xx[1::2] += pitch / 2
    c = np.column_stack((xx.ravel(), yy.ravel()))
    c = c[c[:, 0] < shape[1] - 15]
    y, x = np.indices(shape, dtype=float)
    img = np.full(shape, BACKGROUND)
    for cx, cy in c:
        img += AMPLITUDE * np.exp(-0.5 * ((x - cx) ** 2 + (y - cy) ** 2) / sigma ** 2)
    if snr is not None:
        img += rng.normal(0, AMPLITUDE / snr, shape)
    return img, c


def cells(img, centers, ratio):
    labels = vr.voronoi_labels(centers, img.shape)
    half = int(np.ceil(0.6 * ratio * FWHM))
    rows = np.array([[x - half, y - half, 2 * half + 1, 2 * half + 1, x, y] for x, y in centers])
    _, masks, origins = vr.extract_rois(img, labels, rows, pad=3)
    return masks, np.asarray(origins, dtype=np.float64)


def test_jacobian_and_sigma_major():
    rng = np.random.default_rng(0)
    p = np.column_stack([rng.uniform(10, 50, 5), rng.uniform(3, 8, 5), rng.uniform(3, 8, 5),
                         rng.uniform(-1, 0.3, 5), rng.uniform(-0.4, 0.4, 5), rng.uniform(-1, 0.3, 5)])
    yy, xx = np.indices((11, 12), dtype=float)
    dx, dy = xx.ravel()[None] - p[:, 1, None], yy.ravel()[None] - p[:, 2, None]
    J = db.jacobian(p, dx, dy)
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1e-6
        hi = db.gaussian(p + e, xx.ravel()[None] - (p + e)[:, 1, None], yy.ravel()[None] - (p + e)[:, 2, None])[0]
        lo = db.gaussian(p - e, xx.ravel()[None] - (p - e)[:, 1, None], yy.ravel()[None] - (p - e)[:, 2, None])[0]
        np.testing.assert_allclose(J[:, :, k], (hi - lo) / 2e-6, atol=1e-5 * AMPLITUDE)
    a, b, c = db.precision(p)
    lam_min = (a + c) / 2 - np.sqrt(((a - c) / 2) ** 2 + b * b)
    np.testing.assert_allclose(db.sigma_major(p), 1 / np.sqrt(lam_min), rtol=1e-12)


def test_tile_map_reproduces_planar_tiles():
    tile, shape = 16, (50, 70)
    yc, xc = db.tile_centers(tile, shape)
    med = 3.0 + 0.5 * yc[:, None] - 0.25 * xc[None, :]
    m = db.upsample_tiles(med, tile, shape)
    y, x = np.indices(shape, dtype=float)
    inside = (y >= yc[0]) & (y <= yc[-1]) & (x >= xc[0]) & (x <= xc[-1])
    np.testing.assert_allclose(m[inside], (3.0 + 0.5 * y - 0.25 * x)[inside], atol=1e-12)
    # clamped outside the tile centres
    np.testing.assert_allclose(m[0, 0], med[0, 0])


@pytest.mark.parametrize("ratio", [2.67, 2.0, 1.6, 1.4])
def test_deblend_known_answer_noise_free(ratio):
    img, centers = lattice(ratio)
    masks, origins = cells(img, centers, ratio)
    st = db.deblend(img, centers + 0.2, masks, origins)
    assert st["source"] in ("measured", "interpolated")
    np.testing.assert_allclose(st["background"], BACKGROUND, atol=0.01 * AMPLITUDE)
    assert (st["status"] == db.STATUS_OK).all()
    np.testing.assert_allclose(st["params"][:, 0], AMPLITUDE, rtol=0.01)
    np.testing.assert_allclose(st["params"][:, 1:3] + origins, centers, atol=0.01)


def test_model_background_logquad_known_answer():
    img, centers = lattice(2.0)
    img = np.clip(np.rint(img * 50), 0, 65535).astype(np.uint16)
    from oracle import pipeline

    seeds, _ = pipeline.detect_centers(img, area_min=1, area_max=200, pad=3, refine="logquad",
                                       logquad_variant="cpu", threshold="otsu")
    got, _, st = db.deblended_centers(img.astype(np.float64), seeds, "logquad")
    fin = np.isfinite(got).all(axis=1)
    err = cKDTree(centers).query(got[fin])[0]
    assert fin.sum() == len(seeds)
    assert np.median(err) < 0.005, np.median(err)
