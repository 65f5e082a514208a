"""Pins of the tiling oracle and equality of the product's dedupe with the reference rule."""

import numpy as np

from imageprocessor_new_amd import synthetic
from oracle import segment, tiling


def test_dedupe_identity_and_means():
    pts = np.array([[0.0, 0.0], [10.0, 0.0], [0.5, 0.5], [20.0, 20.0], [20.0, 20.0]])
    out = tiling.dedupe_centers(pts, eps=1.5)
    np.testing.assert_allclose(out, [[0.25, 0.25], [10.0, 0.0], [20.0, 20.0]])
    far = np.array([[0.0, 0.0], [5.0, 5.0], [9.0, 1.0]])
    np.testing.assert_array_equal(tiling.dedupe_centers(far), far)


def test_product_dedupe_equals_reference_rule():
    from imageprocessor_new_amd.centers import dedupe_centers

    rng = np.random.default_rng(0)
    base = rng.uniform(0, 200, size=(300, 2))
    dup = base[rng.choice(300, 120)] + rng.normal(0, 0.6, size=(120, 2))
    chain = np.array([[50.0, 50.0], [51.4, 50.0], [52.8, 50.0], [54.2, 50.0]])  # transitive cluster
    pts = np.concatenate([base, dup, chain])[rng.permutation(424)]
    a = dedupe_centers(pts, eps=1.5)
    b = tiling.dedupe_centers(pts, eps=1.5)
    assert a.shape == b.shape
    np.testing.assert_array_equal(a, b)
    edge = np.array([[0.0, 0.0], [1.5, 0.0], [3.0, 0.0 + 1e-12]])  # exactly eps apart
    np.testing.assert_array_equal(dedupe_centers(edge), tiling.dedupe_centers(edge))


def test_tiling_with_one_strip_is_untiled():
    arr = synthetic.make_workload("square2048", seed=1, H=300, W=256)
    t = segment.otsu_u16(arr.image, 4)
    ref, _, _ = segment.detect_centers_logquad(arr.image, threshold_value=int(np.floor(t * 0.8)),
                                               area_min=20, area_max=2000)
    one = tiling.detect_centers_tiled_logquad(arr.image, tile_h=400, overlap=64, area_min=20, area_max=2000)
    np.testing.assert_allclose(np.sort(one, axis=0), np.sort(tiling.dedupe_centers(ref), axis=0))
