"""GPU parity of the edge-erf refinement (spx_edge_erf, spx_edge_profile_erf) against the FP64
oracle: same accept/refuse decisions, edge positions and centres within 1e-6 px. Both sides
compute in FP64; the normal equations are accepted up to cond(J^T J) = 1e8, so a rounding
difference of the solve (LAPACK LU against the kernel's elimination) or of erf reaches
~1e-9 relative, 4e-7 px at coordinates ~4000 on noisy profiles."""

import math

import numpy as np
import pytest
import torch
from scipy.special import erf

from imageprocessor_new_amd import synthetic
from oracle import edge_erf as ref
from oracle import segment

pytestmark = pytest.mark.gpu

TOL = 1e-6


def test_known_answer_at_six_offsets():
    # reference tests/test_gpu_edge_erf.py:24-33
    from imageprocessor_new_amd.edge_erf import edge_from_profile_erf

    starts = np.array([0.0, 100.0, 150.0, 500.0, 1300.0, 2000.0])
    k = np.arange(24.0)
    e = np.vectorize(math.erf)
    profile = 15.0 + 70.0 * 0.5 * (1 + e((k - 10.37) / (math.sqrt(2) * 1.3)))
    found, ok = edge_from_profile_erf(np.tile(profile, (6, 1)), starts, starts + 10.0)
    assert ok.all()
    np.testing.assert_allclose(found - starts, 10.37, atol=1e-6)


def test_profiles_match_oracle():
    # ideal profiles of the reference's diagnostics (tests/diag_erf_unit.py:106-123) plus
    # noisy ones with random seeds, polarities, widths and lengths
    from imageprocessor_new_amd.edge_erf import edge_from_profile_erf

    rng = np.random.default_rng(3)
    for L in (5, 7, 21, 31):
        B = 512
        xs = np.arange(L, dtype=np.float64)
        edge = rng.uniform(0.3 * L, 0.7 * L, B)
        sig = rng.uniform(0.3, 4.0, B)
        amp = rng.choice([-1.0, 1.0], B) * rng.uniform(5.0, 200.0, B)
        prof = 40.0 + amp[:, None] * 0.5 * (1 + erf((xs[None] - edge[:, None]) / (np.sqrt(2) * sig[:, None])))
        prof += rng.normal(0.0, 1.0, prof.shape) * rng.uniform(0.0, 3.0, (B, 1))
        base = rng.uniform(-50.0, 4000.0, B).round()
        seed = base + edge + rng.uniform(-1.5, 1.5, B)
        got, gok = edge_from_profile_erf(prof, base, seed)
        want, wok = ref.edge_erf(prof, base, seed)
        agree = gok == wok
        assert agree.mean() >= 0.998, (L, np.nonzero(~agree)[0][:5])
        both = gok & wok
        assert both.sum() > B // 4
        np.testing.assert_allclose(got[both], want[both], atol=TOL, rtol=0)
        assert np.isnan(got[~gok]).all()


def test_short_profiles_return_seed():
    from imageprocessor_new_amd.edge_erf import edge_from_profile_erf

    x, ok = edge_from_profile_erf(np.ones((3, 4)), np.zeros(3), np.array([1.0, 2.0, 3.0]))
    assert not ok.any()
    np.testing.assert_array_equal(x, [1.0, 2.0, 3.0])


def _page(seed=1, amp=150.0, bg=30.0, read=2.0, shape=(256, 256), pitch=32.0):
    lat = synthetic.square_lattice(shape[0], shape[1], pitch)
    rng = np.random.default_rng(seed)
    c = lat + rng.normal(0, 0.3, lat.shape)
    return synthetic.render_hole_array(shape, c, radius=8.0, sigma=1.2, amplitude=amp, background=bg,
                                       read_sigma=read, seed=seed)


def _rows(img, area_min=20):
    t = segment.otsu_u16(img)
    num, labels, stats, cents = segment.connected_components_fast(img > t, 8)
    keep = stats[1:, 4] >= area_min
    return np.concatenate([stats[1:, :4], cents[1:]], axis=1)[keep]


@pytest.mark.parametrize("seed,shape,pitch", [(1, (256, 256), 32.0), (2, (300, 411), 27.0), (4, (97, 130), 24.0)])
def test_refine_matches_oracle(seed, shape, pitch):
    from imageprocessor_new_amd.edge_erf import refine_centers_edge_erf_gpu

    arr = _page(seed=seed, shape=shape, pitch=pitch)
    rows = _rows(arr.image).astype(np.float32)
    got, gok = refine_centers_edge_erf_gpu(arr.image, rows)
    want, wok = ref.refine_edge_erf(arr.image.astype(np.float32), rows)
    np.testing.assert_array_equal(gok, wok)
    assert gok.mean() > 0.5
    np.testing.assert_allclose(got[gok], want[wok], atol=TOL, rtol=0)


def test_options_and_border_rows_match_oracle():
    # features cut by the image border (profiles clipped to the edge pixels), wider bands,
    # one pass and three smoothing passes
    from imageprocessor_new_amd.edge_erf import refine_centers_edge_erf_gpu

    arr = _page(seed=7, shape=(200, 200), pitch=30.0)
    rows = _rows(arr.image, area_min=5).astype(np.float32)
    kw = dict(band_rad=6, edge_rad=12, smooth_passes=3, loc_rad=4, iters=1, erf_fit_rad=6)
    got, gok = refine_centers_edge_erf_gpu(arr.image, rows, **kw)
    want, wok = ref.refine_edge_erf(arr.image.astype(np.float32), rows, **kw)
    np.testing.assert_array_equal(gok, wok)
    np.testing.assert_allclose(got[gok], want[wok], atol=TOL, rtol=0)


def test_high_amplitude_edges_fail_the_cond_gate():
    from imageprocessor_new_amd.edge_erf import refine_centers_edge_erf_gpu

    arr = _page(amp=3000.0, bg=500.0, read=30.0)
    got, gok = refine_centers_edge_erf_gpu(arr.image, _rows(arr.image))
    assert not gok.any()


def test_u16_and_f32_images_agree_and_empty_rows():
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.edge_erf import refine_centers_edge_erf_gpu, refine_edge_erf_device

    arr = _page(seed=5)
    rows = _rows(arr.image).astype(np.float32)
    with gpu_device(0) as dev:
        a, aok = refine_edge_erf_device(torch.from_numpy(arr.image).to(dev), torch.from_numpy(rows).to(dev))
        b, bok = refine_edge_erf_device(torch.from_numpy(arr.image.astype(np.float32)).to(dev),
                                        torch.from_numpy(rows).to(dev))
    np.testing.assert_array_equal(aok, bok)
    np.testing.assert_array_equal(a, b)
    c, ok = refine_centers_edge_erf_gpu(arr.image, np.zeros((0, 6), np.float32))
    assert c.shape == (0, 2) and ok.shape == (0,)
    with pytest.raises(ValueError):
        refine_centers_edge_erf_gpu(arr.image, np.zeros((3, 4), np.float32))


def test_bad_options_are_errors():
    # argument errors of the C-ABI (SPX_E_ARG) surface as ValueError
    from imageprocessor_new_amd.edge_erf import refine_centers_edge_erf_gpu

    arr = _page(seed=5, shape=(64, 64))
    rows = np.array([[20, 20, 16, 16, 28.0, 28.0]], np.float32)
    with pytest.raises(ValueError, match="edge_rad"):
        refine_centers_edge_erf_gpu(arr.image, rows, edge_rad=16)
    with pytest.raises(ValueError, match="band_rad"):
        refine_centers_edge_erf_gpu(arr.image, rows, band_rad=-1)
    with pytest.raises(TypeError):
        refine_centers_edge_erf_gpu(arr.image, rows, edge_radius=3)


def test_detect_centers_edge_erf_matches_oracle_pipeline():
    import imageprocessor_new_amd as ipn

    arr = _page(seed=11, shape=(320, 320))
    res = ipn.detect_centers(arr.image, backend="gpu", threshold="otsu", refine="edge_erf", area_min=20,
                             area_max=10**6)
    rows = _rows(arr.image, area_min=20).astype(np.float32)
    want, wok = ref.refine_edge_erf(arr.image.astype(np.float32), rows)
    want = want[wok & np.isfinite(want).all(axis=1)]
    assert "refine=edge_erf" in res.method and res.backend == "gpu"
    assert res.centers_xy.shape == want.shape
    np.testing.assert_allclose(res.centers_xy, want, atol=TOL, rtol=0)
