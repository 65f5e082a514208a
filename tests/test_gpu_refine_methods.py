"""GPU parity of the per-component refinements behind detect_centers(backend="gpu") against
the CPU oracle: the window log-quadratic "logquad" (spx_refine_logquad_window), the
weighted centroid of "weighted" / "none" (spx_refine_weighted), the gradient-moment edges of
"edge_gradmoment" (spx_edge_moment), the "auto" split and the group-by-group output order.

Tolerances: 1e-6 px for the window log-quadratic (its plane background solves 3x3 normal
equations in image coordinates, where LAPACK and the kernel's elimination differ by up to
~1e-7 grey levels); 1e-9 px for the weighted centroid and the gradient moments (sums and
ratios only). Accept / refuse flags must agree exactly.
"""

import numpy as np
import pytest
import torch

from imageprocessor_new_amd import synthetic
from oracle import edge_erf as ref_edge
from oracle import logquad_window as ref_lqw
from oracle import pipeline, tiling

pytestmark = pytest.mark.gpu

TOL_LQW = 1e-6
TOL = 1e-9


def _page(seed, shape=(384, 448), pitch=28.0, r=(2.5, 9.0), amp=3000.0, bg=500.0, read=30.0):
    rng = np.random.default_rng(seed)
    lat = synthetic.square_lattice(shape[0], shape[1], pitch)
    c = lat + rng.normal(0, 0.3, lat.shape)
    radius = rng.uniform(r[0], r[1], c.shape[0])
    return synthetic.render_hole_array(shape, c, radius=radius, sigma=rng.uniform(0.8, 1.6, c.shape[0]),
                                       amplitude=amp, background=bg, read_sigma=read, seed=seed)


def _dev_tensors(img, labels):
    from imageprocessor_new_amd.backends import gpu_device

    with gpu_device(0) as dev:
        return dev, torch.from_numpy(img).to(dev), torch.from_numpy(labels.astype(np.int32)).to(dev)


@pytest.mark.parametrize("seed,shape,pitch,r", [(0, (384, 448), 12.0, (1.5, 3.5)), (1, (257, 301), 28.0, (2.5, 9.0)),
                                                (2, (96, 130), 12.0, (1.5, 3.5))])
def test_logquad_window_matches_oracle(seed, shape, pitch, r):
    # small holes at about twice their diameter (the method's design case: all accepted)
    # and a mixed page whose large holes overfill the 9x9 window (many refused)
    from imageprocessor_new_amd.refine import refine_logquad_window_device

    arr = _page(seed, shape, pitch=pitch, r=r)
    t, labels, rows5, rows6 = pipeline.component_rows(arr.image, area_min=1, area_max=10**6)
    dev, img_t, lab_t = _dev_tensors(arr.image, labels)
    got, gok = refine_logquad_window_device(img_t, lab_t, torch.from_numpy(rows6).to(dev),
                                            torch.from_numpy(rows5[:, 4].astype(np.int32)).to(dev))
    want, wok = ref_lqw.refine_logquad_window(arr.image.astype(np.float64), labels, rows6, rows5[:, 4])
    np.testing.assert_array_equal(gok, wok)
    assert gok.mean() > 0.3
    np.testing.assert_allclose(got, want, atol=TOL_LQW, rtol=0)


def test_logquad_window_options_and_float_image():
    from imageprocessor_new_amd.refine import refine_logquad_window_device

    arr = _page(5, (200, 220))
    t, labels, rows5, rows6 = pipeline.component_rows(arr.image, area_min=1, area_max=10**6)
    img32 = arr.image.astype(np.float32) * np.float32(0.37)
    dev, img_t, lab_t = _dev_tensors(img32, labels)
    for kw in [dict(plane_bg=False), dict(win_rad=3, bg_win_pad=3, bg_dilate=1, fit_dilate=0),
               dict(bg_tail_frac=0.0, core_frac=0.4, weight_power=2.0), dict(win_rad=6, bg_win_pad=4)]:
        got, gok = refine_logquad_window_device(img_t, lab_t, torch.from_numpy(rows6).to(dev),
                                                torch.from_numpy(rows5[:, 4].astype(np.int32)).to(dev), pad=2, **kw)
        want, wok = ref_lqw.refine_logquad_window(img32.astype(np.float64), labels, rows6, rows5[:, 4], pad=2, **kw)
        np.testing.assert_array_equal(gok, wok, err_msg=str(kw))
        np.testing.assert_allclose(got, want, atol=TOL_LQW, rtol=0, err_msg=str(kw))


def test_logquad_window_recovers_log_quadratic_spots():
    # float page of exactly log-quadratic spots on a plane (the oracle's exact-recovery pin)
    from imageprocessor_new_amd.refine import refine_centers_logquad_gpu_match_cpu

    H, W = 96, 128
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    g = 300.0 + 0.35 * xx - 0.2 * yy
    labels = np.zeros((H, W), np.int32)
    rng = np.random.default_rng(4)
    truth = []
    for i in range(4):
        for j in range(3):
            x0, y0 = 20 + 24 * i + rng.uniform(-0.5, 0.5), 20 + 28 * j + rng.uniform(-0.5, 0.5)
            dx, dy = xx - x0, yy - y0
            spot = 900.0 * np.exp(-(0.2 * dx * dx + 0.15 * dy * dy + 0.08 * dx * dy))
            spot[np.hypot(dx, dy) > 3.2] = 0.0
            g += spot
            labels[spot > 270.0] = len(truth) + 1
            truth.append((x0, y0))
    rows = []
    for lab in range(1, len(truth) + 1):
        ys, xs = np.nonzero(labels == lab)
        rows.append([xs.min(), ys.min(), xs.max() - xs.min() + 1, ys.max() - ys.min() + 1, xs.mean(), ys.mean()])
    c, ok = refine_centers_logquad_gpu_match_cpu(g.astype(np.float32), labels, np.asarray(rows, np.float32),
                                                 np.arange(1, len(truth) + 1, dtype=np.int32))
    assert ok.all()
    # float32 grey levels: log-quadratic to ~1e-7 relative
    np.testing.assert_allclose(c, np.asarray(truth), atol=2e-5)


def test_logquad_match_cpu_validation_and_four_columns():
    from imageprocessor_new_amd.refine import refine_centers_logquad_gpu_match_cpu

    arr = _page(6, (128, 128))
    t, labels, rows5, rows6 = pipeline.component_rows(arr.image, area_min=1, area_max=10**6)
    g = arr.image.astype(np.float32)
    with pytest.raises(ValueError, match="float32"):
        refine_centers_logquad_gpu_match_cpu(arr.image, labels, rows6, rows5[:, 4])
    with pytest.raises(ValueError, match="must match"):
        refine_centers_logquad_gpu_match_cpu(g, labels, rows6, rows5[:-1, 4])
    with pytest.raises(ValueError, match="4 or 6"):
        refine_centers_logquad_gpu_match_cpu(g, labels, rows6[:, :5], rows5[:, 4])
    got, gok = refine_centers_logquad_gpu_match_cpu(g, labels, rows6[:, :4], rows5[:, 4])
    r4 = rows6.copy()
    r4[:, 4] = rows5[:, 0] + rows5[:, 2] // 2
    r4[:, 5] = rows5[:, 1] + rows5[:, 3] // 2
    want, wok = ref_lqw.refine_logquad_window(g.astype(np.float64), labels, r4, rows5[:, 4])
    np.testing.assert_array_equal(gok, wok)
    np.testing.assert_allclose(got, want, atol=TOL_LQW, rtol=0)


@pytest.mark.parametrize("seed", [0, 3])
def test_weighted_and_edge_moment_match_oracle(seed):
    from imageprocessor_new_amd.edge_erf import refine_edge_moment_device
    from imageprocessor_new_amd.refine import refine_weighted_rows_device

    arr = _page(seed)
    t, labels, rows5, rows6 = pipeline.component_rows(arr.image, area_min=1, area_max=10**6)
    dev, img_t, lab_t = _dev_tensors(arr.image, labels)
    got, gok = refine_weighted_rows_device(img_t, lab_t, rows5, pad=3)
    want, wok = ref_lqw.refine_weighted_rows(arr.image, labels, rows5, pad=3)
    np.testing.assert_array_equal(gok, wok)
    np.testing.assert_allclose(got[gok], want[wok], atol=TOL, rtol=0)
    got, gok = refine_edge_moment_device(img_t, torch.from_numpy(rows6).to(dev))
    want, wok = ref_edge.refine_edge_moment(arr.image.astype(np.float32), rows6)
    np.testing.assert_array_equal(gok, wok)
    np.testing.assert_allclose(got, want, atol=TOL, rtol=0)


@pytest.mark.parametrize("refine", ["logquad", "weighted", "none", "edge_gradmoment", "edge_moment", "auto",
                                    "edge_erf_centered"])
def test_detect_centers_methods_match_oracle_pipeline(refine):
    import imageprocessor_new_amd as ipn

    if refine == "edge_erf_centered":
        arr = _page(8, amp=150.0, bg=30.0, read=2.0)
    else:
        arr = _page(7)
    res = ipn.detect_centers(arr.image, backend="gpu", threshold="otsu", refine=refine, area_min=10, area_max=5000)
    want, t = pipeline.detect_centers(arr.image, refine=refine, area_min=10, area_max=5000)
    assert res.meta["threshold_value"] == t
    assert res.centers_xy.shape == want.shape
    np.testing.assert_allclose(res.centers_xy, want, atol=TOL_LQW, rtol=0)
    if refine == "auto":
        # both groups present: small holes through logquad, large ones through the edges
        _, _, rows5, _ = pipeline.component_rows(arr.image, area_min=10, area_max=5000)
        size = np.maximum(rows5[:, 2], rows5[:, 3])
        assert (size <= 12).any() and (size > 12).any()


def test_tiled_default_logquad_matches_oracle():
    from imageprocessor_new_amd.legacy import detect_centers_tiled

    arr = synthetic.make_workload("square2048", seed=8, H=1024, W=384)
    with pytest.warns(DeprecationWarning):
        res = detect_centers_tiled(arr.image, tile_h=300, overlap=64, backend="gpu", threshold="otsu",
                                   refine="logquad", area_min=20, area_max=2000)
    want = tiling.detect_centers_tiled_logquad(arr.image, tile_h=300, overlap=64, area_min=20, area_max=2000,
                                               logquad_variant="gpu")
    assert res.centers_xy.shape == want.shape
    np.testing.assert_allclose(res.centers_xy, want, atol=TOL_LQW, rtol=0)
