"""The batched erf-disk fit at full size: the 8192 x 8192 square array of 65,536 holes in
31 x 31 patches, checked through properties that hold at any size (integer-shift
equivariance, mirror symmetry, run-to-run identity) and a seeded sample against the FP64
oracle."""
import numpy as np
import pytest

from imageprocessor_new_amd import synthetic
from oracle import erfdisk

pytestmark = pytest.mark.gpu

HALF = 15


@pytest.fixture(scope="module")
def full():
    from imageprocessor_new_amd.fit import fit_erf_disks

    print("rendering the 8192 x 8192 array", flush=True)
    arr = synthetic.make_workload("square8192", seed=21)
    seeds = np.rint(arr.centers_xy)
    res = fit_erf_disks(arr.image, seeds, half=HALF, model="disk")
    return arr, seeds, res


def _fit(img, seeds):
    from imageprocessor_new_amd.fit import fit_erf_disks

    return fit_erf_disks(np.ascontiguousarray(img), seeds, half=HALF, model="disk")


def test_full_size_sample_matches_the_oracle(full):
    arr, seeds, res = full
    K = seeds.shape[0]
    assert K == 65536 and res.ok.mean() >= 0.999
    idx = np.sort(np.random.default_rng(4).choice(K, size=300, replace=False))
    ref_p, _, _, ref_st = erfdisk.fit_batch(arr.image, seeds[idx], HALF, 0)
    st = res.status[idx].astype(np.int64)
    assert (st == ref_st).mean() >= 0.99
    both = (st == 0) & (ref_st == 0)
    err = np.abs(res.centers_xy[idx][both] - ref_p[both, :2]).max()
    print("max |gpu - fp64 oracle| centre over the sample:", err)
    assert err <= 1e-3
    inside = res.ok & arr.inside
    assert np.median(np.abs(res.centers_xy[inside] - arr.centers_xy[inside])) < 0.02


def test_integer_shift_gives_the_same_fits(full):
    # the page cut by (dx, dy) = (5, 3): every patch inside both frames holds the same pixels
    arr, seeds, res = full
    dx, dy = 5, 3
    H, W = arr.image.shape
    moved = _fit(arr.image[dy:, dx:], seeds - [dx, dy])
    keep = ((seeds[:, 0] - HALF >= dx) & (seeds[:, 1] - HALF >= dy) & (seeds[:, 0] + HALF <= W - 1)
            & (seeds[:, 1] + HALF <= H - 1))
    assert keep.mean() > 0.99
    np.testing.assert_array_equal(moved.status[keep], res.status[keep])
    np.testing.assert_array_equal(moved.iterations[keep], res.iterations[keep])
    np.testing.assert_array_equal(moved.params[keep, 2:], res.params[keep, 2:])
    assert np.abs(moved.centers_xy[keep] + [dx, dy] - res.centers_xy[keep]).max() <= 1e-9


def test_mirror_image_gives_mirrored_fits(full):
    arr, seeds, res = full
    W = arr.image.shape[1]
    mirrored = _fit(arr.image[:, ::-1], np.stack([W - 1 - seeds[:, 0], seeds[:, 1]], 1))
    both = (mirrored.status == 0) & (res.status == 0)
    assert both.mean() >= 0.999
    c = mirrored.centers_xy[both]
    d = np.abs(np.stack([W - 1 - c[:, 0], c[:, 1]], 1) - res.centers_xy[both]).max()
    print("max centre difference of the mirrored page:", d)
    assert d <= 1e-4
    assert np.abs(mirrored.params[both, 2:4] - res.params[both, 2:4]).max() <= 1e-4


def test_second_run_is_identical(full):
    arr, seeds, res = full
    again = _fit(arr.image, seeds)
    np.testing.assert_array_equal(again.params, res.params)
    np.testing.assert_array_equal(again.centers_xy, res.centers_xy)
    np.testing.assert_array_equal(again.status, res.status)
