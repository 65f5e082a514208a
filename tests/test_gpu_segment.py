"""GPU parity of Otsu, threshold + connected components, log-quadratic refinement and the
detect_centers pipeline against the CPU oracle (bit-exact for integer outputs; 1e-8 px for
the FP64 log-quadratic centres)."""

import numpy as np
import pytest
import torch

from imageprocessor_new_amd import synthetic
from oracle import segment, tiling

pytestmark = pytest.mark.gpu


def _dev():
    from imageprocessor_new_amd.backends import gpu_device

    return gpu_device(0)


@pytest.mark.parametrize("seed,size,step", [(0, 512, 1), (1, 640, 1), (2, 700, 4), (3, 333, 3)])
def test_otsu_matches_oracle(seed, size, step):
    from imageprocessor_new_amd.components import otsu_threshold_device

    arr = synthetic.make_workload("square2048", seed=seed, H=size, W=size + 17)
    with _dev() as dev:
        t = otsu_threshold_device(torch.from_numpy(arr.image).to(dev), step)
    assert t == segment.otsu_u16(arr.image, step)


def test_otsu_small_and_constant():
    from imageprocessor_new_amd.components import otsu_threshold_device

    rng = np.random.default_rng(5)
    for shape in [(1, 1), (3, 7), (31, 2)]:
        img = rng.integers(0, 65535, size=shape).astype(np.uint16)
        with _dev() as dev:
            assert otsu_threshold_device(torch.from_numpy(img).to(dev)) == segment.otsu_u16(img)
    with _dev() as dev:
        assert otsu_threshold_device(torch.full((9, 9), 77, dtype=torch.uint16, device=dev)) == 0


@pytest.mark.parametrize("conn", [8, 4])
@pytest.mark.parametrize("seed,shape,frac", [(0, (57, 93), 0.45), (1, (64, 64), 0.3), (2, (301, 257), 0.55),
                                             (3, (1, 40), 0.5), (4, (40, 1), 0.5), (5, (600, 700), 0.6),
                                             (6, (512, 512), 0.35)])
def test_components_match_oracle(conn, seed, shape, frac):
    from imageprocessor_new_amd.components import threshold_components_device

    rng = np.random.default_rng(seed)
    img = (rng.random(shape) < frac).astype(np.uint16) * 1000
    with _dev() as dev:
        num, labels, stats, cents = threshold_components_device(torch.from_numpy(img).to(dev), 500,
                                                                connectivity=conn)
        labels, stats, cents = labels.cpu().numpy(), stats.cpu().numpy(), cents.cpu().numpy()
    rn, rl, rs, rc = segment.connected_components_fast(img > 500, conn)
    assert num == rn
    assert np.array_equal(labels, rl)
    assert np.array_equal(stats, rs)
    np.testing.assert_array_equal(cents, rc)


def test_components_invert_and_capacity_retry():
    from imageprocessor_new_amd.components import threshold_components_device

    rng = np.random.default_rng(9)
    img = rng.integers(0, 1000, size=(120, 130)).astype(np.uint16)
    with _dev() as dev:
        num, labels, stats, cents = threshold_components_device(torch.from_numpy(img).to(dev), 300,
                                                                invert=True, capacity=4)
        labels = labels.cpu().numpy()
    rn, rl, _, _ = segment.connected_components_fast(img <= 300, 8)
    assert num == rn and np.array_equal(labels, rl)


def test_logquad_and_pipeline_match_oracle_c0():
    from imageprocessor_new_amd.centers import detect_centers

    arr = synthetic.make_workload("square2048", seed=4, H=768, W=640)
    res = detect_centers(arr.image, backend="gpu", threshold="otsu", refine="logquad", area_min=20,
                         area_max=2000, pad=3, logquad_variant="cpu")
    ref, rows, t = segment.detect_centers_logquad(arr.image, area_min=20, area_max=2000, pad=3)
    assert res.meta["threshold_value"] == t
    assert res.centers_xy.shape == ref.shape
    np.testing.assert_allclose(res.centers_xy, ref, rtol=0, atol=1e-8)
    # statistical accuracy against the truth (nearest true centre)
    d = np.hypot(*(res.centers_xy[:, None, :] - arr.centers_xy[None, :, :]).transpose(2, 0, 1)).min(axis=1)
    assert np.median(d) < 0.2


def test_logquad_border_and_small_rois():
    from imageprocessor_new_amd.centers import detect_centers

    arr = synthetic.make_workload("square2048", seed=6, H=200, W=200)
    # holes cut by the border and a tiny speck; area filters off
    img = arr.image.copy()
    img[5:7, 100:102] = 4000
    res = detect_centers(img, backend="gpu", threshold="otsu", refine="logquad", area_min=1,
                         area_max=5000, pad=3, logquad_variant="cpu")
    ref, _, _ = segment.detect_centers_logquad(img, area_min=1, area_max=5000, pad=3)
    assert res.centers_xy.shape == ref.shape
    np.testing.assert_allclose(res.centers_xy, ref, rtol=0, atol=1e-8)


def test_empty_image_gives_no_centers():
    from imageprocessor_new_amd.centers import detect_centers

    res = detect_centers(np.full((64, 64), 500, np.uint16), backend="gpu", threshold="otsu")
    assert res.centers_xy.shape == (0, 2)


def test_tiled_matches_oracle():
    from imageprocessor_new_amd.legacy import detect_centers_tiled

    arr = synthetic.make_workload("square2048", seed=8, H=1024, W=384)
    with pytest.warns(DeprecationWarning):
        res = detect_centers_tiled(arr.image, tile_h=300, overlap=64, backend="gpu", threshold="otsu",
                                   refine="logquad", area_min=20, area_max=2000, logquad_variant="cpu")
    ref = tiling.detect_centers_tiled_logquad(arr.image, tile_h=300, overlap=64, area_min=20, area_max=2000)
    assert res.centers_xy.shape == ref.shape
    np.testing.assert_allclose(res.centers_xy, ref, rtol=0, atol=1e-8)


def test_tiled_overlap_error():
    from imageprocessor_new_amd.centers import detect_centers

    with pytest.raises(ValueError, match="overlap must be < tile_h"):
        detect_centers(np.zeros((64, 64), np.uint16), backend="gpu", tile_h=64, overlap=64)


def test_erf_disk_refine_through_detect_centers():
    from imageprocessor_new_amd.centers import detect_centers

    arr = synthetic.make_workload("square8192", seed=12, H=384, W=384)
    res = detect_centers(arr.image, backend="gpu", threshold="otsu", refine="erf_disk", area_min=20,
                         area_max=2000, fit_half=15)
    d = np.hypot(*(res.centers_xy[:, None, :] - arr.centers_xy[None, :, :]).transpose(2, 0, 1)).min(axis=1)
    assert res.centers_xy.shape[0] >= 0.9 * arr.centers_xy.shape[0]
    assert np.median(d) < 0.02


def _page8(seed, shape=(300, 340)):
    rng = np.random.default_rng(seed)
    lat = synthetic.square_lattice(shape[0], shape[1], 24.0)
    c = lat + rng.normal(0, 0.3, lat.shape)
    arr = synthetic.render_hole_array(shape, c, radius=rng.uniform(3, 7, c.shape[0]), sigma=1.2, amplitude=150.0,
                                      background=30.0 + 40.0 * (seed % 2), read_sigma=3.0, seed=seed)
    arr.image = np.minimum(arr.image, 255).astype(np.uint16)  # 8-bit grey levels
    return arr


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_triangle_matches_oracle(seed):
    from imageprocessor_new_amd.components import triangle_threshold_device

    arr = _page8(seed)
    img = arr.image if seed < 3 else (255 - np.minimum(arr.image, 255)).astype(np.uint16)
    assert img.max() <= 255
    with _dev() as dev:
        t = triangle_threshold_device(torch.from_numpy(np.ascontiguousarray(img)).to(dev))
    assert t == segment.triangle_u8(img)


def test_triangle_rejects_16_bit_images():
    from imageprocessor_new_amd.components import triangle_threshold_device

    arr = synthetic.make_workload("square2048", seed=1, H=128, W=128)
    with _dev() as dev, pytest.raises(ValueError, match="8-bit"):
        triangle_threshold_device(torch.from_numpy(arr.image).to(dev))


@pytest.mark.parametrize("mo,mc,invert", [(1, 0, False), (0, 2, False), (2, 1, True)])
def test_morphology_matches_oracle(mo, mc, invert):
    from imageprocessor_new_amd.components import binary_mask_device

    arr = _page8(5)
    t = segment.otsu_u16(arr.image)
    with _dev() as dev:
        m = binary_mask_device(torch.from_numpy(arr.image).to(dev), t, invert=invert, morph_open=mo,
                               morph_close=mc).cpu().numpy()
    np.testing.assert_array_equal(m.astype(bool), segment.binary_mask(arr.image, t, invert, mo, mc))


def test_detect_centers_triangle_and_morphology_match_oracle():
    from imageprocessor_new_amd.centers import detect_centers
    from oracle import pipeline

    arr = _page8(6)
    res = detect_centers(arr.image, backend="gpu", threshold="triangle", morph_open=1, morph_close=1,
                         refine="weighted", area_min=5, area_max=4000)
    want, t = pipeline.detect_centers(arr.image, threshold="triangle", morph_open=1, morph_close=1,
                                      refine="weighted", area_min=5, area_max=4000)
    assert res.meta["threshold_value"] == t
    assert res.centers_xy.shape == want.shape and want.shape[0] > 50
    np.testing.assert_allclose(res.centers_xy, want, atol=1e-9, rtol=0)
    with pytest.raises(ValueError, match="Unsupported threshold"):
        detect_centers(arr.image, backend="gpu", threshold="mean")


def test_components_without_the_label_image_give_the_same_stats():
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.components import threshold_components_device

    rng = np.random.default_rng(17)
    img = (rng.random((700, 900)) * 1000).astype(np.uint16)
    with gpu_device(0) as dev:
        t = torch.from_numpy(img.astype(np.int32)).to(torch.uint16).to(dev)
        n1, lab, st1, c1 = threshold_components_device(t, 600)
        n2, none, st2, c2 = threshold_components_device(t, 600, labels_out=False)
        assert none is None and lab is not None and n1 == n2 > 100
        assert torch.equal(st1, st2) and torch.equal(c1, c2)
