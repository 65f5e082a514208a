"""Large-page path on the GPU: the page renderer (spx_render_page), tiled detect_centers with
erf-disk refinement and Voronoi cells, and the device pair search of the dedupe."""

import numpy as np
import pytest

from imageprocessor_new_amd import synthetic

pytestmark = pytest.mark.gpu


def test_rendered_page_statistics():
    # noise model: per-pixel mean = clean intensity, variance = clean + read_sigma^2
    holes = np.array([[40.3, 50.7, 8.0, 1.2, 1.0, 3000.0]], np.float32)
    img = synthetic.render_page(holes, 256, 256, background=500.0, read_sigma=30.0, seed=3).cpu().numpy()
    bg = img[150:, 150:].astype(np.float64)
    assert abs(bg.mean() - 500.0) < 1.0
    assert abs(bg.var() - (500.0 + 900.0 + 1.0 / 12)) < 60.0
    clean = synthetic.render_clean((256, 256), np.array([[40.3, 50.7]]), np.array([8.0]), np.array([1.2]),
                                   np.ones(1), 3000.0, 500.0)
    inside = clean > 3000
    assert abs(img[inside].mean() - clean[inside].mean()) < 5.0


def test_tiled_detect_on_rendered_mixed_page():
    from imageprocessor_new_amd.centers import detect_centers

    H = W = 4096
    holes, _, _ = synthetic.mixed_page_holes(seed=2, H=H, W=W, region=1024)
    img_t = synthetic.render_page(holes, H, W, seed=2)
    res = detect_centers(img_t, backend="gpu", threshold="otsu", tile_h=1024, overlap=128, refine="erf_disk",
                         area_min=20, area_max=5000, fit_cells=True)
    from scipy.spatial import cKDTree

    d, _ = cKDTree(res.centers_xy).query(holes[:, :2].astype(np.float64))
    assert (d <= 0.5).mean() > 0.97
    assert np.median(d[d <= 0.5]) < 0.02
    assert set(res.meta["stage_ms"]) >= {"threshold", "components", "refine", "dedupe"}
    # no duplicates survive the band rule and the dedupe
    dd, _ = cKDTree(res.centers_xy).query(res.centers_xy, k=2)
    assert dd[:, 1].min() > 1.5


def test_device_pairs_give_the_kdtree_dedupe():
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.centers import _candidate_pairs_device, _candidate_pairs_kdtree, _merge_pairs

    rng = np.random.default_rng(7)
    pts = rng.uniform(0, 3000, (20000, 2))
    dup = pts[:500] + rng.normal(0, 0.5, (500, 2))
    trip = pts[500:600] + rng.normal(0, 0.4, (100, 2))
    pts = np.concatenate([pts, dup, trip, pts[600:650] + [1.5, 0.0]])
    with gpu_device(0):
        pairs = _candidate_pairs_device(pts, 1.5, (3001, 3001))
    assert pairs is not None
    a = _merge_pairs(pts, pairs, 1.5)
    b = _merge_pairs(pts, _candidate_pairs_kdtree(pts, 1.5), 1.5)
    np.testing.assert_array_equal(a, b)
    assert a.shape[0] < pts.shape[0]


def test_device_strip_path_matches_the_row_path():
    """The device-resident erf strip (one strip, global threshold) gives the centres of the
    row-by-row path (_detect_single_tile's erf group) on the same threshold."""
    import torch

    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.centers import _detect_single_tile, _erf_strip_device
    from imageprocessor_new_amd import _lib
    from imageprocessor_new_amd.components import otsu_threshold_device

    H = W = 1024
    holes, _, _ = synthetic.mixed_page_holes(seed=4, H=H, W=W, region=512)
    img_t = synthetic.render_page(holes, H, W, seed=4)
    opts = _lib.fit_options()
    with gpu_device(0):
        thr = int(otsu_threshold_device(img_t))
        c_dev, ok = _erf_strip_device(img_t, thr, invert=False, area_min=20, area_max=5000, connectivity=8,
                                      fit_half=None, fit_options=opts, model_code=_lib.MODEL_ERF_DISK6,
                                      fit_cells=True)
        got = c_dev[ok].cpu().numpy()
        ref, _ = _detect_single_tile(img_t, threshold="otsu", invert=False, area_min=20, area_max=5000, pad=3,
                                     refine="erf_disk", connectivity=8, pre_threshold=float(thr), fit_half=None,
                                     fit_options=opts, fit_cells=True)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


def test_learned_start_in_the_device_strip_path():
    """With a (barely trained) network the strip path still returns the plain fit's centres:
    every fit either converges from the proposal to the same minimum or is refitted."""
    import torch

    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.centers import detect_centers
    from imageprocessor_new_amd import learned

    H = W = 1024
    holes, _, _ = synthetic.mixed_page_holes(seed=5, H=H, W=W, region=512)
    img_t = synthetic.render_page(holes, H, W, seed=5)
    kw = dict(backend="gpu", threshold="otsu", tile_h=512, overlap=128, refine="erf_disk", area_min=20,
              area_max=5000, fit_cells=True)
    with gpu_device(0) as dev:
        net, _ = learned.train_learned_start(n_tiles=4096, epochs=1, batch=1024, device=dev)
        plain = detect_centers(img_t, **kw)
        got = detect_centers(img_t, learned_start=net, **kw)
        single = detect_centers(img_t, **dict(kw, tile_h=None), learned_start=net)
    from scipy.spatial import cKDTree

    assert got.centers_xy.shape == plain.centers_xy.shape
    d, _ = cKDTree(plain.centers_xy).query(got.centers_xy)
    assert d.max() < 1e-3
    assert got.meta["learned_fallbacks"] is not None
    d1, _ = cKDTree(holes[:, :2].astype(np.float64)).query(single.centers_xy)
    assert (d1 < 0.5).mean() > 0.97
    with pytest.raises(ValueError):
        detect_centers(img_t, learned_start=net, **dict(kw, refine="logquad"))
