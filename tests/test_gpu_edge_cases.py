"""Empty, flat, single-feature and ragged inputs through the GPU entry points: no error, and
outputs of the right shapes (empty, or NaN with ok = False), as the reference's wrappers
return them; the component labelling of an empty mask equals the oracle's."""
import numpy as np
import pytest

from imageprocessor_new_amd import synthetic
from oracle import segment

pytestmark = pytest.mark.gpu

REFINES = ("logquad", "weighted", "none", "edge_gradmoment", "edge_erf", "erf_disk", "erf_ellipse", "auto")


@pytest.mark.parametrize("refine", REFINES)
def test_flat_image_gives_no_centres(refine):
    from imageprocessor_new_amd.centers import detect_centers

    res = detect_centers(np.full((96, 80), 500, np.uint16), backend="gpu", threshold="otsu", refine=refine)
    assert res.centers_xy.shape == (0, 2)
    tiled = detect_centers(np.full((300, 80), 500, np.uint16), backend="gpu", threshold="otsu", refine=refine,
                           tile_h=128, overlap=32)
    assert tiled.centers_xy.shape == (0, 2)


# Each method on a hole it accepts: the window log-quadratic refuses holes that overfill its
# 9 x 9 window, and the erf-edge fit's conditioning gate refuses edges of high amplitude
# (both as the reference); the others take the 3000-count, r = 8 hole.
HOLE = {"logquad": dict(radius=3.0), "edge_erf": dict(amplitude=150.0, background=30.0, read_sigma=2.0)}


@pytest.mark.parametrize("refine", REFINES)
def test_one_hole_gives_one_centre_near_the_truth(refine):
    from imageprocessor_new_amd.centers import detect_centers

    c = np.array([[30.3, 27.6]])
    kw = dict(radius=8.0, sigma=1.2, noise=True, seed=3)
    kw.update(HOLE.get(refine, {}))
    arr = synthetic.render_hole_array((64, 60), c, **kw)
    res = detect_centers(arr.image, backend="gpu", threshold="otsu", refine=refine, area_min=5, area_max=2000)
    assert res.centers_xy.shape == (1, 2)
    assert np.abs(res.centers_xy - c).max() < 0.25


def test_empty_batches_of_every_refinement():
    from imageprocessor_new_amd.edge_erf import refine_centers_edge_moment_gpu
    from imageprocessor_new_amd.fit import fit_erf_disks

    img = np.full((64, 64), 500, np.uint16)
    r = fit_erf_disks(img, np.zeros((0, 2)))
    assert r.params.shape == (0, 7) and r.centers_xy.shape == (0, 2) and r.ok.shape == (0,)
    c, ok = refine_centers_edge_moment_gpu(img, np.zeros((0, 6), np.float32))
    assert c.shape == (0, 2) and ok.shape == (0,)


def test_components_of_an_empty_and_a_full_mask_match_the_oracle():
    from imageprocessor_new_amd.components import connected_components_stats

    for m in (np.zeros((40, 33), np.uint8), np.ones((40, 33), np.uint8)):
        got = connected_components_stats(m)
        n, labels, stats, cents = segment.connected_components(m, 8)
        assert got.num_labels == n
        np.testing.assert_array_equal(got.labels, labels)
        np.testing.assert_array_equal(got.stats, stats)
        np.testing.assert_allclose(got.centroids, cents, rtol=0, atol=1e-9)


def test_patches_at_the_size_bounds_and_past_the_frame():
    """half = 2 and half = 31 patches, disks and ellipses, seeds on and outside the frame's
    corners: statuses and parameters as the oracle's."""
    from imageprocessor_new_amd.fit import fit_erf_disks
    from oracle import erfdisk

    arr = synthetic.make_workload("hex16384", seed=13, H=200, W=200)
    n = arr.centers_xy.shape[0]
    seeds = np.concatenate([np.rint(arr.centers_xy), [[0.0, 0.0], [199.0, 199.0], [-3.0, 50.0], [50.0, 205.0]]])
    K = seeds.shape[0]
    half = np.where(np.arange(K) % 2 == 0, 31, 2).astype(np.int16)
    half[:n][arr.radius > 9] = 31
    model = (np.arange(K) % 3 == 0).astype(np.uint8)
    res = fit_erf_disks(arr.image, seeds, half=half, model=model)
    ref_p, _, _, ref_st = erfdisk.fit_batch(arr.image, seeds, half, model)
    st = res.status.astype(np.int64)
    same = (st == ref_st) | (np.isin(st, (0, 1)) & np.isin(ref_st, (0, 1)))
    assert same.mean() >= 0.95, (np.bincount(st), np.bincount(ref_st))
    both = (st == 0) & (ref_st == 0) & (half == 31)
    assert both.sum() >= 5
    assert np.abs(res.centers_xy[both] - ref_p[both, :2]).max() <= 1e-3
    assert np.all(res.status[seeds[:, 0] < 0] == ref_st[seeds[:, 0] < 0])


@pytest.mark.parametrize("refine", REFINES)
@pytest.mark.parametrize("tile_h", [None, 48])
def test_a_copy_far_from_the_origin_gives_the_same_centres(refine, tile_h):
    """The same 64 x 256 patch of holes at x = 0 and at x = 32512 of a 64 x 32768 page: every
    method returns the second copy's centres as the first's moved by exactly 32512 px (to 1e-9
    px: FP32 coordinates there are 4e-3 px apart)."""
    from imageprocessor_new_amd.centers import detect_centers

    rng = np.random.default_rng(8)
    c = np.stack([20.0 + 30.0 * np.arange(8) + rng.uniform(-0.4, 0.4, 8), 32.0 + rng.uniform(-0.4, 0.4, 8)], 1)
    kw = dict(radius=8.0, sigma=1.2, noise=True, seed=4)
    kw.update(HOLE.get(refine, {}))
    small = synthetic.render_hole_array((64, 256), c, **kw).image
    bg = np.uint16(np.median(small))
    page = np.full((64, 32768), bg, np.uint16)
    off = 32768 - 256
    page[:, :256] = small
    page[:, off:] = small
    res = detect_centers(page, backend="gpu", threshold="otsu", refine=refine, area_min=5, area_max=2000,
                         tile_h=tile_h, overlap=16)
    xy = res.centers_xy
    a, b = xy[xy[:, 0] < 256], xy[xy[:, 0] >= off]
    assert a.shape[0] >= 6 and a.shape == b.shape
    a, b = a[np.lexsort(a.T)], b[np.lexsort(b.T)]
    assert np.abs(b - [off, 0] - a).max() <= 1e-9
