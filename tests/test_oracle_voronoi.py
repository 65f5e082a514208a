"""Pins of the Voronoi-refinement oracle (oracle/voronoi_refine.py): nearest-seed labels against
SciPy's k-d tree and a planted tie; the resampling against the closed forms it must reproduce
(Keys' cubic reproduces quadratics; edge-aligned end points) and SciPy's order-0 zoom; erosion,
the 3x3 edge box and the Gaussian pre-smoothing against scipy.ndimage; the ROI fill value
against a direct border walk; both estimators against symmetry (centred blob, mirror and
transpose equivariance) and sub-pixel recovery of radially symmetric blobs; the pipeline on a
small hole page."""

import numpy as np
import pytest
from scipy import ndimage
from scipy.spatial import cKDTree

from imageprocessor_new_amd import synthetic
from oracle import voronoi_refine as vr


def test_labels_nearest_seed_and_ties():
    rng = np.random.default_rng(0)
    seeds = rng.random((40, 2)) * [97, 61]
    lab = vr.voronoi_labels(seeds, (61, 97))
    yy, xx = np.mgrid[:61, :97]
    d, i = cKDTree(seeds).query(np.stack([xx.ravel(), yy.ravel()], axis=1), k=2)
    clear = d[:, 0] < d[:, 1]
    assert np.array_equal(lab.ravel()[clear], i[clear, 0])
    # a pixel midway between two seeds belongs to the lower index
    tie = vr.voronoi_labels(np.array([[6.0, 2.0], [2.0, 2.0]]), (5, 9))
    assert tie[2, 4] == 0 and tie[2, 3] == 1
    tie = vr.voronoi_labels(np.array([[2.0, 2.0], [6.0, 2.0]]), (5, 9))
    assert tie[2, 4] == 0 and tie[2, 5] == 1
    assert (vr.voronoi_labels(np.zeros((0, 2)), (3, 4)) == -1).all()


def test_bicubic_reproduces_quadratics_and_end_points():
    Hin, Win, f = 9, 12, 4
    y, x = np.mgrid[:Hin, :Win].astype(np.float64)
    src = (0.3 * y * y - 2.0 * y + 0.1 * x * x + 1.5 * x + 40.0)[None].astype(np.float32)
    up = vr.upsample_bicubic(src, f)[0]
    Ho, Wo = Hin * f, Win * f
    sy = np.arange(Ho) * np.float32((Hin - 1.0) / (Ho - 1.0))
    sx = np.arange(Wo) * np.float32((Win - 1.0) / (Wo - 1.0))
    want = 0.3 * sy[:, None] ** 2 - 2.0 * sy[:, None] + 0.1 * sx[None, :] ** 2 + 1.5 * sx[None, :] + 40.0
    # taps that stay inside the input: floor(s) - 1 >= 0 and floor(s) + 2 <= n - 1
    iy = (np.floor(sy) >= 1) & (np.floor(sy) <= Hin - 3)
    ix = (np.floor(sx) >= 1) & (np.floor(sx) <= Win - 3)
    assert np.abs(up[np.ix_(iy, ix)] - want[np.ix_(iy, ix)]).max() < 2e-4 * np.abs(want).max()
    assert up[0, 0] == src[0, 0, 0] and up[-1, -1] == src[0, -1, -1]


def test_nearest_matches_scipy_zoom_order0():
    rng = np.random.default_rng(1)
    m = (rng.random((3, 7, 11)) < 0.5).astype(np.float32)
    got = vr.upsample_nearest(m, 4)
    want = np.stack([ndimage.zoom(s, 4, order=0, grid_mode=False) for s in m])
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("margin", [0, 1, 2, 4])
def test_erosion_matches_scipy(margin):
    rng = np.random.default_rng(2)
    m = rng.random((2, 30, 23)) < 0.8
    got = vr.erode_square(m, margin)
    if margin == 0:
        assert np.array_equal(got, m)
        return
    want = ndimage.binary_erosion(m, structure=np.ones((1, 2 * margin + 1, 2 * margin + 1), bool))
    assert np.array_equal(got, want)


def test_box3_and_gaussian_match_scipy():
    rng = np.random.default_rng(3)
    a = (rng.random((2, 17, 21)) * 1000).astype(np.float32)
    box = vr.box3_edge(a)
    want = ndimage.uniform_filter(a.astype(np.float64), size=(1, 3, 3), mode="nearest")
    assert np.abs(box - want).max() < 1e-3
    for sigma in (0.5, 1.3):
        g = vr.gaussian_filter_yx(a, sigma)
        want = ndimage.gaussian_filter(a, sigma=(0, sigma, sigma), mode="reflect")
        assert g.dtype == np.float32
        assert np.abs(g - want).max() <= 2e-4 * np.abs(want).max()
    tiny = (rng.random((1, 2, 3)) * 10).astype(np.float32)  # axes shorter than the filter
    want = ndimage.gaussian_filter(tiny, sigma=(0, 1.3, 1.3), mode="reflect")
    assert np.abs(vr.gaussian_filter_yx(tiny, 1.3) - want).max() < 1e-4


def test_roi_fill_is_the_border_percentile():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 4000, size=(40, 50)).astype(np.uint16)
    seeds = np.array([[12.3, 10.1], [30.0, 12.0], [20.2, 28.7], [40.0, 30.0]])
    lab = vr.voronoi_labels(seeds, img.shape)
    rows = np.array([[9, 7, 7, 7, 12.3, 10.1], [26, 8, 8, 8, 30.0, 12.0], [17, 25, 7, 8, 20.2, 28.7]])
    rois, masks, origins = vr.extract_rois(img, lab, rows, pad=3)
    for j, (x0, y0) in enumerate(origins):
        x1 = min(50, int(rows[j, 0] + rows[j, 2]) + 3)
        y1 = min(40, int(rows[j, 1] + rows[j, 3]) + 3)
        vals = []
        for yy in range(y0, y1):
            for xx in range(x0, x1):
                if lab[yy, xx] != j:
                    continue
                frame = yy in (y0, y1 - 1) or xx in (x0, x1 - 1)
                out = any(lab[yy + dy, xx + dx] != j for dy, dx in ((1, 0), (-1, 0), (0, 1), (0, -1))
                          if 0 <= yy + dy < 40 and 0 <= xx + dx < 50)
                if frame or out:
                    vals.append(float(img[yy, xx]))
        bg = np.percentile(vals, 10)
        h, w = y1 - y0, x1 - x0
        cell = lab[y0:y1, x0:x1] == j
        assert np.array_equal(masks[j, :h, :w], cell) and not masks[j, h:].any() and not masks[j, :, w:].any()
        assert np.array_equal(rois[j, :h, :w], np.where(cell, img[y0:y1, x0:x1], bg))
        assert (rois[j, h:] == bg).all() and (rois[j, :, w:] == bg).all()


def _blob(H, W, cx, cy, r=5.0, s=1.2, A=3000.0, B=500.0):
    from scipy.special import erfc

    y, x = np.mgrid[:H, :W].astype(np.float64)
    rho = np.hypot(x - cx, y - cy)
    return B + A * 0.5 * erfc((rho - r) / (np.sqrt(2) * s))


# Mirror / transpose tolerance: radial symmetry is equivariant to float32 rounding; the
# isophote votes divide by a curvature term whose float32 sums (a - 2c + b against b - 2c + a)
# round differently in the mirror image, and near-flat pixels amplify that.
@pytest.mark.parametrize("fn,tol", [(vr.radial_symmetry, 2e-5), (vr.isophote_curvature, 5e-3)],
                         ids=["radial", "isophote"])
def test_estimators_symmetry_and_recovery(fn, tol):
    # centred: the computation is symmetric under the half turn, so the centre is exact up to the
    # float32 rounding of the votes
    img = _blob(23, 23, 11.0, 11.0)[None]
    mask = np.ones_like(img, dtype=bool)
    c, q = fn(img, mask)
    assert np.abs(c[0] - [11.0, 11.0]).max() < 1e-5
    # off-centre blobs: mirror and transpose equivariance, recovery of the true centre
    truth = np.array([[10.37, 12.81], [12.6, 9.45], [11.2, 11.9]])
    imgs = np.stack([_blob(23, 25, x, y) for x, y in truth])
    masks = np.ones_like(imgs, dtype=bool)
    c, q = fn(imgs, masks)
    assert np.isfinite(q).all()
    assert np.abs(c - truth).max() < 0.05, c - truth
    cm, _ = fn(imgs[:, :, ::-1].copy(), masks)
    assert np.abs(cm[:, 0] - (24 - c[:, 0])).max() < tol and np.abs(cm[:, 1] - c[:, 1]).max() < tol
    ct, _ = fn(imgs.transpose(0, 2, 1).copy(), masks.transpose(0, 2, 1).copy())
    assert np.abs(ct[:, ::-1] - c).max() < tol


def test_pipeline_on_a_hole_page():
    arr = synthetic.make_workload("square2048", seed=5, H=256, W=320)
    # the isophote votes come from second differences: on these noisy flat-topped disks
    # (Poisson + N(0, 30^2)) the plateau's noise curvature spreads them by tenths of a pixel
    for refine, med, worst in (("radial_symmetry", 0.05, 0.3), ("isophote_curvature", 0.5, 1.5)):
        c, meta = vr.detect_voronoi(arr.image, refine, area_min=20, area_max=2000)
        key = "radial_symmetry_residual" if refine == "radial_symmetry" else "isophote_curvature_spread"
        assert c.shape[0] == meta[key].shape[0] == meta["voronoi_cell_area"].shape[0] > 40
        d, _ = cKDTree(arr.centers_xy).query(c)
        assert np.median(d) < med and d.max() < worst, (refine, np.median(d), d.max())
