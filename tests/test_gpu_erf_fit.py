"""GPU parity of the batched erf-disk / erf-ellipse LM against the FP64 oracle.

Bar: centres, r and sigma within 1e-3 px of the oracle; A, B within 1e-4 relative;
identical status codes; bit-identical results when run twice.
"""

import numpy as np
import pytest

from imageprocessor_new_amd import synthetic
from oracle import erfdisk

pytestmark = pytest.mark.gpu


def _fit_gpu(img, seeds, half, model, **kw):
    from imageprocessor_new_amd.fit import fit_erf_disks

    return fit_erf_disks(img, seeds, half=half, model=model, **kw)


def _compare(res, ref_params, ref_status, tol_px=1e-3, status_agree=0.99):
    """Statuses agree on >= status_agree of the features; where both fits converged the
    parameters agree with the FP64 oracle (centres, r, sigma to tol_px; A, B relative 1e-4).
    A fit that stops at the iteration cap in one precision and converges in the other is
    the only tolerated status difference."""
    st = res.status.astype(np.int64)
    assert (st == ref_status).mean() >= status_agree, (np.bincount(st), np.bincount(ref_status))
    differ = st != ref_status
    assert np.all(np.isin(st[differ], (0, 1)) & np.isin(ref_status[differ], (0, 1)))
    both = (st == 0) & (ref_status == 0)
    assert both.mean() > 0.9
    d = np.abs(res.params[both].astype(np.float64) - ref_params[both])
    d[:, :2] = np.abs(res.centers_xy[both] - ref_params[both, :2])
    print("max |gpu - fp64 oracle| per output x0 y0 r s A B q:", d.max(axis=0))
    assert d[:, 0].max() <= tol_px and d[:, 1].max() <= tol_px
    assert d[:, 2].max() <= tol_px and d[:, 3].max() <= tol_px
    rel = d[:, 4:6] / np.abs(ref_params[both][:, 4:6])
    assert rel.max() <= 1e-4
    assert d[:, 6].max() <= 1e-4


def test_erf_disk_matches_fp64_oracle_c1_shape():
    arr = synthetic.make_workload("square8192", seed=3, H=512, W=512)
    seeds = np.rint(arr.centers_xy)
    res = _fit_gpu(arr.image, seeds, 15, "disk")
    ref_p, ref_rms, ref_it, ref_st = erfdisk.fit_batch(arr.image, seeds, 15, 0)
    _compare(res, ref_p, ref_st)
    ok = res.ok & arr.inside
    err = np.abs(res.params[ok, :2] - arr.centers_xy[ok])
    assert np.median(err) < 0.02  # statistical accuracy against the truth


def test_noise_free_recovery():
    c = np.array([[40.3, 41.7], [100.45, 37.2], [70.1, 90.9], [20.6, 110.2]])
    arr = synthetic.render_hole_array((140, 140), c, radius=[8, 7.3, 9.1, 6.2], sigma=[1.2, 0.9, 1.5, 1.0],
                                      noise=False)
    res = _fit_gpu(arr.image, c + 0.4, 15, "disk", step_tol=1e-7)
    assert (res.status == 0).all(), res.status
    # integer rounding of the image is the only noise: errors well below 1e-3 px
    assert np.abs(res.params[:, :2] - c).max() < 2e-4
    assert np.abs(res.params[:, 2] - arr.radius).max() < 2e-4


def test_mixed_disks_and_ellipses():
    arr = synthetic.make_workload("hex16384", seed=5, H=600, W=600)
    seeds = np.rint(arr.centers_xy)
    model = arr.ellipse.astype(np.uint8)
    half = np.clip(np.ceil(arr.radius + 3 * arr.sigma).astype(int) + 2, 7, 31)
    res = _fit_gpu(arr.image, seeds, half, model)
    ref_p, _, _, ref_st = erfdisk.fit_batch(arr.image, seeds, half, model)
    # patches up to 63 x 63 at a 39 px pitch hold parts of neighbouring holes: a few fits
    # run to the iteration cap
    _compare(res, ref_p, ref_st, status_agree=0.97)


def test_border_and_degenerate_patches():
    arr = synthetic.make_workload("square8192", seed=7, H=256, W=256)
    seeds = np.concatenate([np.rint(arr.centers_xy), [[0.0, 0.0], [255.0, 3.0], [128.0, 255.0]]])
    res = _fit_gpu(arr.image, seeds, 15, "disk")
    ref_p, _, _, ref_st = erfdisk.fit_batch(arr.image, seeds, 15, 0)
    assert (res.status.astype(np.int64) == ref_st).mean() >= 0.97
    flat = np.full((64, 64), 700, dtype=np.uint16)
    res2 = _fit_gpu(flat, [[32.0, 32.0]], 10, "disk")
    assert res2.status[0] == 6 and not res2.ok[0]  # flat patch: no start


def test_deterministic():
    arr = synthetic.make_workload("square8192", seed=11, H=384, W=384)
    seeds = np.rint(arr.centers_xy)
    a = _fit_gpu(arr.image, seeds, 15, "disk")
    b = _fit_gpu(arr.image, seeds, 15, "disk")
    assert np.array_equal(a.params, b.params, equal_nan=True)
    assert np.array_equal(a.status, b.status)


def test_init_start_is_used():
    arr = synthetic.make_workload("square8192", seed=2, H=256, W=256)
    seeds = np.rint(arr.centers_xy)
    first = _fit_gpu(arr.image, seeds, 15, "disk")
    again = _fit_gpu(arr.image, seeds, 15, "disk", init=first.params)
    ok = first.ok & again.ok
    assert np.abs(again.params[ok, :2] - first.params[ok, :2]).max() < 1e-3
    assert again.iterations[ok].mean() < first.iterations[ok].mean()


def test_seed_neighbours_match_oracle():
    import torch

    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.fit import seed_neighbours_device

    rng = np.random.default_rng(11)
    seeds = rng.uniform(-3, 2003, (4000, 2)).astype(np.float32)
    seeds[:40] = seeds[40:80] + np.float32(0.5)  # close pairs
    seeds[100:130] = np.nan  # seeds that take no part
    seeds[130:140, 1] = np.inf
    for radius, nb in [(60.0, 8), (25.0, 3), (200.0, 1)]:
        with gpu_device(0) as dev:
            got = seed_neighbours_device(torch.from_numpy(seeds).to(dev), (2000, 2000), radius, nb).cpu().numpy()
        np.testing.assert_array_equal(got, erfdisk.seed_neighbours(seeds, radius, nb))
        assert np.all(got[100:140] == -1) and not np.isin(np.arange(100, 140), got).any()


@pytest.mark.parametrize("seed", [5, 6])
def test_cell_masked_fits_match_oracle_c2_shape(seed):
    # staggered array, r 4-20 px, patches 15-63 px at a 39 px pitch, 10 % ellipses: every
    # patch restricted to its Voronoi cell among the 8 nearest seeds
    from imageprocessor_new_amd.fit import cell_radius

    arr = synthetic.make_workload("hex16384", seed=seed, H=500, W=500)
    seeds = np.rint(arr.centers_xy).astype(np.float32)
    model = arr.ellipse.astype(np.uint8)
    half = np.clip(np.ceil(arr.radius + 3 * arr.sigma).astype(int) + 2, 7, 31)
    res = _fit_gpu(arr.image, seeds, half, model, cells=True)
    nbr = erfdisk.seed_neighbours(seeds, cell_radius(int(half.max())), 8)
    ref_p, _, _, ref_st = erfdisk.fit_batch(arr.image, seeds, half, model, neighbours=nbr)
    _compare(res, ref_p, ref_st, status_agree=0.97)
    # the cell removes the neighbours' light: more fits converge, and the radius (the CD)
    # is no longer pulled by it (centres of a symmetric array are little biased either way)
    plain = _fit_gpu(arr.image, seeds, half, model)
    inside = arr.inside
    m_ok = (res.status == 0) & inside
    p_ok = (plain.status == 0) & inside
    both = m_ok & p_ok
    m_r = np.median(np.abs(res.params[both, 2] - arr.radius[both]))
    p_r = np.median(np.abs(plain.params[both, 2] - arr.radius[both]))
    print("converged masked / plain:", m_ok.sum(), p_ok.sum(), "median radius error:", m_r, p_r)
    assert m_ok.sum() >= p_ok.sum()
    assert m_r < p_r


def test_cells_are_a_no_op_when_patches_sit_inside_them():
    # 31 x 31 patches on the 32 px square array lie inside their cells, except where two
    # rounded seeds end up 30 px apart (a bisector row of ties goes to the lower index)
    from imageprocessor_new_amd.fit import cell_radius

    arr = synthetic.make_workload("square8192", seed=4, H=320, W=320)
    seeds = np.rint(arr.centers_xy).astype(np.float32)
    nbr = erfdisk.seed_neighbours(seeds, cell_radius(15), 8)
    inside = np.zeros(len(seeds), bool)
    for k in range(len(seeds)):
        gy, gx = np.meshgrid(np.arange(31) + (seeds[k, 1] - 15), np.arange(31) + (seeds[k, 0] - 15), indexing="ij")
        inside[k] = erfdisk.cell_mask(gx, gy, seeds, k, nbr[k]).all()
    assert inside.mean() > 0.95
    a = _fit_gpu(arr.image, seeds, 15, "disk")
    b = _fit_gpu(arr.image, seeds, 15, "disk", cells=True)
    np.testing.assert_array_equal(a.status[inside], b.status[inside])
    np.testing.assert_array_equal(a.params[inside], b.params[inside])


def test_launch_configurations_give_identical_fits_and_autotune_picks_one():
    import torch

    from imageprocessor_new_amd import _lib
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.fit import LAUNCH_CANDIDATES, _TUNED, autotune_launch, fit_erf_disks_device

    arr = synthetic.make_workload("hex16384", seed=9, H=600, W=600)
    K = arr.centers_xy.shape[0]
    hs = np.clip(np.ceil(arr.radius + 3 * arr.sigma).astype(int) + 2, 7, 31).astype(np.int16)
    with gpu_device(0) as dev:
        img = torch.from_numpy(arr.image).to(dev)
        seeds = torch.from_numpy(np.rint(arr.centers_xy).astype(np.float32)).to(dev)
        half = torch.from_numpy(hs).to(dev)
        model = torch.from_numpy(arr.ellipse.astype(np.uint8)).to(dev)
        outs = []
        for tw, pers in LAUNCH_CANDIDATES:
            o = _lib.fit_options(tile_waves=tw, persistent=pers)
            r = fit_erf_disks_device(img, seeds, half, model, options=o)
            outs.append({k: v.cpu().numpy() for k, v in r.items()})
        # every bucket kernel at 2, 3 and 4 waves per SIMD (register allocation only)
        for code in (1, 2, 3):
            o = _lib.fit_options(tile_waves=4, bucket_waves=sum(code << (2 * b) for b in range(10)))
            r = fit_erf_disks_device(img, seeds, half, model, options=o)
            outs.append({k: v.cpu().numpy() for k, v in r.items()})
        tuned = autotune_launch(img, seeds, half, model, repeats=2)
        assert tuned["best"][:2] in LAUNCH_CANDIDATES and len(tuned["ms"]) == len(LAUNCH_CANDIDATES)
        mask = _lib.bucket_mask_of(hs, arr.ellipse.astype(np.uint8))
        assert sorted(tuned["bucket_ms"]) == [b for b in range(10) if (mask >> b) & 1]
        for b, per in tuned["bucket_ms"].items():
            code = (tuned["best"][2] >> (2 * b)) & 3
            assert code in (1, 2, 3) and per[(2, 3, 4)[code - 1]] == min(per.values())
        assert _TUNED
        again = {k: v.cpu().numpy() for k, v in fit_erf_disks_device(img, seeds, half, model).items()}
    for o in outs[1:] + [again]:
        for k in ("params", "centers", "rms", "iters", "status"):
            np.testing.assert_array_equal(o[k], outs[0][k])


def test_bucket_mask_skips_only_the_masked_classes():
    import torch

    from imageprocessor_new_amd import _lib
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.fit import fit_erf_disks_device

    arr = synthetic.make_workload("hex16384", seed=10, H=600, W=600)
    hs = np.clip(np.ceil(arr.radius + 3 * arr.sigma).astype(int) + 2, 7, 31).astype(np.int16)
    md = arr.ellipse.astype(np.uint8)
    full = _lib.bucket_mask_of(hs, md)
    assert full != 0
    # a feature's bucket: the patch class of its in-image patch pixels (no neighbours here)
    seeds_np = np.rint(arr.centers_xy).astype(np.int64)
    side = 2 * hs.astype(np.int64) + 1
    lo = seeds_np - hs[:, None]
    cols = np.clip(lo[:, 0] + side, 0, 600) - np.clip(lo[:, 0], 0, 600)
    rows = np.clip(lo[:, 1] + side, 0, 600) - np.clip(lo[:, 1], 0, 600)
    bucket = np.array([2 * _lib.patch_class(int(n)) + int(m) for n, m in zip(cols * rows, md)])
    assert np.all(((full >> bucket) & 1) == 1)
    low = 1 << int(bucket.min())  # only the lowest occupied bucket
    inside = bucket == int(bucket.min())
    with gpu_device(0) as dev:
        args = (torch.from_numpy(arr.image).to(dev), torch.from_numpy(np.rint(arr.centers_xy).astype(np.float32)).to(dev),
                torch.from_numpy(hs).to(dev), torch.from_numpy(md).to(dev))
        ref = {k: v.cpu().numpy() for k, v in fit_erf_disks_device(*args).items()}
        got = {k: v.cpu().numpy() for k, v in fit_erf_disks_device(*args, options=_lib.fit_options(bucket_mask=full)).items()}
        part = {k: v.cpu().numpy() for k, v in fit_erf_disks_device(*args, options=_lib.fit_options(bucket_mask=low)).items()}
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k])
    assert inside.any() and (~inside).any()
    for k in ref:
        np.testing.assert_array_equal(part[k][inside], ref[k][inside])
    assert np.all(part["status"][~inside] == _lib.FIT_TOO_FEW)
    assert np.all(np.isnan(part["params"][~inside])) and np.all(np.isnan(part["centers"][~inside]))


def test_fp64_centres_keep_their_precision_far_from_the_origin():
    """Near x = 16400 an FP32 coordinate is spaced 2e-3 px apart; the FP64 centres agree with
    the oracle as closely as near the origin, and the FP32 parameters are those centres
    rounded to FP32."""
    rng = np.random.default_rng(17)
    n = 8
    c = np.stack([16390.0 + 40.0 * np.arange(n) + rng.uniform(-0.5, 0.5, n), 32.0 + rng.uniform(-0.5, 0.5, n)], 1)
    arr = synthetic.render_hole_array((64, 16448 + 320), c, radius=rng.uniform(6, 10, n),
                                      sigma=rng.uniform(0.8, 1.6, n), noise=True, seed=5)
    seeds = np.rint(c)
    res = _fit_gpu(arr.image, seeds, 15, "disk")
    ref_p, _, _, ref_st = erfdisk.fit_batch(arr.image, seeds, 15, 0)
    both = (res.status == 0) & (ref_st == 0)
    assert both.all(), (res.status, ref_st)
    err = np.abs(res.centers_xy - ref_p[:, :2]).max()
    print("max |FP64 centre - oracle| near x = 16400:", err)
    assert err <= 1e-4
    np.testing.assert_array_equal(res.params[:, :2], res.centers_xy.astype(np.float32))


@pytest.mark.parametrize("budget", [1, 4])
def test_budgeted_fit_and_fallback_match_the_oracle(budget):
    # proposals near the plain fit's minimum (what a good learned start gives), some far off
    # (which the judge refuses): the kernel's budget, residual judge and in-launch fallback
    # against the oracle's restatement of the same rule
    arr = synthetic.make_workload("hex16384", seed=15, H=500, W=500)
    seeds = np.rint(arr.centers_xy)
    half = np.clip(np.ceil(arr.radius + 3 * arr.sigma), 7, 31).astype(np.int16)
    model = arr.ellipse.astype(np.uint8)
    ref0, _, _, st0 = erfdisk.fit_batch(arr.image, seeds, half, model)
    rng = np.random.default_rng(budget)
    K = seeds.shape[0]
    init = np.nan_to_num(ref0, nan=0.0).copy()
    init[:, :2] = np.where(np.isfinite(ref0[:, :2]), ref0[:, :2], seeds)
    init[:, 2:4] = np.where(np.isfinite(ref0[:, 2:4]), ref0[:, 2:4], [6.0, 1.2])
    init[:, 4:6] = np.where(np.isfinite(ref0[:, 4:6]), ref0[:, 4:6], [3000.0, 500.0])
    init[:, 6] = np.where(np.isfinite(ref0[:, 6]), ref0[:, 6], 1.0)
    init[:, :2] += rng.normal(0, 0.03, (K, 2))
    init[:, 2] *= 1 + rng.normal(0, 0.01, K)
    init[:, 4] *= 1 + rng.normal(0, 0.02, K)
    far = rng.random(K) < 0.2
    init[far, 0] += 1.5
    init = init.astype(np.float32)
    ref, _, ref_it, ref_st = erfdisk.fit_batch(arr.image, seeds, half, model, init=init, learned_iters=budget)
    res = _fit_gpu(arr.image, seeds, half, model, init=init, learned_iters=budget)
    _compare(res, ref, ref_st, status_agree=0.98)
    # the far starts: with one attempt the judge refuses almost all and the wave refits them
    # (iterations beyond the budget, as in the oracle); with four, LM recovers many in budget
    conv = far & (res.status == 0) & (ref_st == 0)
    fb_gpu, fb_ref = res.iterations[conv] > budget, ref_it[conv] > budget
    assert (fb_gpu == fb_ref).mean() > 0.95
    if budget == 1:
        assert fb_gpu.mean() > 0.9
    # the accepted fits sit at the plain fit's minimum within the metric's 1e-3 px
    both = (res.status == 0) & (st0 == 0)
    assert np.abs(res.centers_xy[both] - ref0[both, :2]).max() < 1e-3
