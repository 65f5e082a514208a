"""GPU parity of the one-pass outline fit (fit_outline_fast) against the oracle (oracle/onepass.py),
step by step on the same inputs, then the whole fit:

- seeds: the same maxima in the same order, refined positions to 1e-9 px;
- geometry: the same nearest distances, spacing, window, stack and box sizes;
- plain stack: bit-equal pixel counts, means to 1e-12 relative; each hole against it to 1e-9;
- pair maps to 1e-9 of their largest value (fixed-point sums against FP64 sums);
- page model: cost, scale, constant, gradient and Gauss-Newton matrix to 1e-8 relative, the grid of
  starts' table to 1e-9; band templates to 1e-12 of their peak;
- steps A to C: the page CD to 1e-4 px and the centres to 1e-4 px of the oracle's (the hole pass
  samples FP32 templates in both; the reductions differ in order), and within 0.01 px of the CD the
  page was drawn with;
- the public entry: result fields, the same answer twice (bit for bit), dark holes read as bright
  ones (the reference's own checks, tests/test_outline_fit_fast.py:236-252).
"""
import numpy as np
import pytest
import torch

from imageprocessor_new_amd import onepass as P
from imageprocessor_new_amd.backends import gpu_device
from imageprocessor_new_amd.outline import fit_outline_fast
from oracle import onepass as OP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def drawn():
    return {blur: OP.drawn_page(blur) for blur in ("gauss2", "lens")}


def _blobs(seed=8, h=80, w=100):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(10, w - 6, 11.0), np.arange(10, h - 6, 11.0))
    c = np.stack([xs.ravel(), ys.ravel()], 1) + rng.uniform(-0.5, 0.5, (xs.size, 2))
    yy, xx = np.indices((h, w), dtype=float)
    img = np.full((h, w), 20.0)
    for x, y in c:
        img += 200.0 * np.exp(-0.5 * ((xx - x) ** 2 + (yy - y) ** 2) / 1.6 ** 2)
    return img.astype(np.float32)


def test_seeds_match_the_oracle(drawn):
    for img in (_blobs(), drawn["gauss2"][3].astype(np.float32), _blobs(seed=9, h=64, w=300)):
        want = OP.find_seeds(img)
        with gpu_device(0) as dev:
            got = P.find_seeds_device(torch.from_numpy(img).to(dev)).cpu().numpy()
        print("seeds", img.shape, len(want), flush=True)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)


def test_geometry_stack_and_versus_match_the_oracle(drawn):
    _, _, _, page, seeds = drawn["gauss2"]
    ref = OP.Page(page, seeds)
    pg = P.OnePassPage(page, seeds)
    np.testing.assert_allclose(pg.nearest_px, ref.nn, rtol=1e-15)
    assert (pg.spacing, pg.R, pg.Ms, pg.reach, pg.L) == (ref.spacing, ref.R, ref.Ms, ref.reach, ref.L)
    inside = ref.interior()
    assert np.array_equal(pg.interior(), inside)
    with gpu_device(0):
        stack, wsum = pg.stack_of(inside)
    want_s, want_w = ref.stack_of(inside)
    stack, wsum = stack.cpu().numpy(), wsum.cpu().numpy()
    assert np.array_equal(wsum, want_w) and np.array_equal(np.isnan(stack), np.isnan(want_s))
    fin = np.isfinite(want_s)
    np.testing.assert_allclose(stack[fin], want_s[fin], rtol=1e-12)
    with gpu_device(0):
        o = pg.versus(torch.from_numpy(want_s).to("cuda"), 5.0)
    want = ref.versus(want_s, 5.0)
    for k in range(6):
        np.testing.assert_allclose(o[:, k], want[:, k], rtol=1e-9, atol=1e-9 * np.abs(want[:, k]).max(),
                                   err_msg=f"column {k}")
    ref.build_stack()
    pg.build_stack()
    assert np.array_equal(pg.use, ref.use)
    assert pg.noise == pytest.approx(ref.noise, rel=1e-9)
    np.testing.assert_allclose(pg.a_rel, ref.a_rel, rtol=1e-9)


@pytest.mark.parametrize("blur", ["gauss2", "lens"])
def test_pair_map_and_page_model_match_the_oracle(drawn, blur):
    _, _, _, page, seeds = drawn[blur]
    ref = OP.Page(page, seeds, blur=blur)
    ref.build_stack()
    pg = P.OnePassPage(page, seeds, blur=blur)
    pg.build_stack()
    sref = OP.ShapeFitter(ref.t, box_px=ref.L, f_max=ref.f_max, U=ref.U, R=ref.R, n_poly=OP.N_POLY)
    with gpu_device(0) as dev:
        sf = P.Fitter(pg.t, box_px=pg.L, R=pg.R, device=dev)
        assert (sf.F, sf.npts, sf.nky, sf.nkx) == (sref.F, sref.npts, sref.nky, sref.nkx)
        gam = pg.pair_map(sf)
        want_g = ref.pair_map(sref)
        np.testing.assert_allclose(gam.cpu().numpy().reshape(want_g.shape), want_g, rtol=0,
                                   atol=1e-9 * np.abs(want_g).max())
        y, w = pg.stack_data(sf)
        yr, wr = ref.stack_data(sref)
        np.testing.assert_allclose(y.cpu().numpy(), yr, rtol=1e-12)
        np.testing.assert_allclose(w.cpu().numpy(), wr, rtol=1e-12)
        th = pg.start_theta()
        np.testing.assert_allclose(th, ref.start_theta(), rtol=1e-12)
        th = np.asarray(th) + np.r_[0.05, -0.03, 0.2, 0.1, 0.01, -0.02, 0.01, 0.0, np.zeros(sf.P - 8) + 0.05]
        free, tm = P.stage_map(pg.t, "full")
        data = sf.prepare(gam[None], y[None], w[None])
        got = sf.evaluate(th[None], data, tm)
        want = sref.evaluate(th[None], yr[None], wr[None], want_g.reshape(1, *want_g.shape), tm)
        for k, (a, b) in enumerate(zip(got, want)):
            np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10 * np.abs(b).max(), err_msg=str(k))
        base = np.asarray(th)[None].copy()
        bd = sf.grid_start(base, data, keep=(tab := []))
        br = sref.grid_start(base, yr[None], wr[None], want_g.reshape(1, *want_g.shape))
        np.testing.assert_allclose(tab[0]["cost"], sref.last_grid[0], rtol=1e-9)
        np.testing.assert_array_equal(bd, br)
        # the hole pass's templates: the band synthesis against the oracle's
        grids = pg.band_grids(th, 32, 0.75, 8, 256, (0.0, 0.1)).cpu().numpy()
        band = OP.Band(box_px=32, f_max=0.75, upsample=8)
        for k, d in enumerate((0.0, 0.1)):
            want_t = band.to_grid(OP.analytic_hat(ref.t, th, band.fx, band.fy, 256, d))
            np.testing.assert_allclose(grids[k], want_t, rtol=0, atol=1e-12 * np.abs(want_t).max())


@pytest.mark.parametrize("blur", ["gauss2", "lens"])
def test_one_pass_fit_matches_the_oracle(drawn, blur):
    t, theta, c, page, seeds = drawn[blur]
    want = OP.fit_images([page], [seeds], blur=blur)[0]
    got = P.fit_images([page], [seeds], blur=blur)[0]
    print(blur, "cd", t.cd_px(got["theta"]), t.cd_px(want["theta"]), "truth", t.cd_px(theta),
          "first", t.cd_px(got["theta_first"]), t.cd_px(want["theta_first"]), flush=True)
    assert t.cd_px(got["theta_first"]) == pytest.approx(t.cd_px(want["theta_first"]), abs=1e-6)
    assert t.cd_px(got["theta"]) == pytest.approx(t.cd_px(want["theta"]), abs=1e-4)
    assert t.cd_px(got["theta"]) == pytest.approx(t.cd_px(theta), abs=0.01)
    assert (got["good"] == want["good"]).mean() > 0.98
    np.testing.assert_allclose(got["centers_xy"], want["centers_xy"], rtol=0, atol=1e-4)
    assert got["chi2"] == pytest.approx(want["chi2"], rel=1e-3)
    assert got["exchanged"] == want["exchanged"] and got["window_px"] == want["window_px"]


def test_fit_outline_fast_entry(drawn):
    t, theta, c, page, seeds = drawn["gauss2"]
    a = fit_outline_fast(page, seeds, pixel_size=21.0, unit="nm")
    b = fit_outline_fast(page, seeds, pixel_size=21.0, unit="nm")
    assert np.array_equal(a.theta, b.theta) and np.array_equal(a.centers_xy, b.centers_xy)
    assert a.rays_px.shape == (128,) and a.summary["per_hole_size"] is False
    assert a.outline_cd_px == pytest.approx(t.cd_px(theta), abs=0.01)
    assert a.cd_x_px * a.cd_y_px == pytest.approx(a.outline_cd_px ** 2, rel=1e-9)
    assert a.summary["holes_stacked"] == int(a.good.sum()) and 0.5 < a.summary["leftover_chi2"] < 2
    truth = c + np.array(t.centroid_px(theta))
    d = np.hypot(*(a.centers_xy[a.good] - truth[a.good]).T)
    assert np.sqrt(np.mean(d ** 2)) < 0.02
    # the method's own seeds must not decide the page CD
    own = fit_outline_fast(page, None)
    assert abs(own.outline_cd_px - a.outline_cd_px) * 21.0 < 0.05
    # dark holes on a bright background read the same; the level comes back in their grey levels
    dark = fit_outline_fast(5000.0 - page, seeds, invert=True)
    assert dark.outline_cd_px == pytest.approx(a.outline_cd_px, abs=1e-3)
    both = dark.good & a.good
    assert both.mean() > 0.6
    np.testing.assert_allclose(dark.level[both], 5000.0 - a.level[both], atol=0.05)


def test_sparse_page_takes_the_wide_paths():
    """Pitch 20 px: spacing > 15 px (the wider cell list), window R = 15 (the stack's bins no
    longer fit in LDS: the global fixed-point path), and the fit still matches the oracle."""
    t, theta, c, page, seeds = OP.drawn_page("gauss2", shape=(200, 260), pitch=20.0, seed=11)
    ref = OP.Page(page, seeds)
    pg = P.OnePassPage(page, seeds)
    assert pg.spacing > 15.0 and (pg.R, pg.Ms, pg.L) == (ref.R, ref.Ms, ref.L) and pg.Ms ** 2 * 12 > 64 * 1024
    inside = ref.interior()
    with gpu_device(0):
        stack, wsum = pg.stack_of(inside)
    want_s, want_w = ref.stack_of(inside)
    assert np.array_equal(wsum.cpu().numpy(), want_w)
    fin = np.isfinite(want_s)
    np.testing.assert_allclose(stack.cpu().numpy()[fin], want_s[fin], rtol=1e-12)
    want = OP.fit_images([page], [seeds])[0]
    got = P.fit_images([page], [seeds])[0]
    print("sparse cd", t.cd_px(got["theta"]), t.cd_px(want["theta"]), t.cd_px(theta), flush=True)
    assert t.cd_px(got["theta"]) == pytest.approx(t.cd_px(want["theta"]), abs=1e-4)
    assert t.cd_px(got["theta"]) == pytest.approx(t.cd_px(theta), abs=0.02)


def test_pages_in_lockstep_read_as_alone(drawn):
    """Pages fitted together (one group in lockstep, and a second window size in its own group)
    give each page the theta and centres it gets alone."""
    pages, seeds = [], []
    for seed in (5, 6):
        _, _, _, page, s = OP.drawn_page("gauss2", seed=seed)
        pages.append(page)
        seeds.append(s)
    _, _, _, sparse, s = OP.drawn_page("gauss2", shape=(200, 260), pitch=20.0, seed=11)
    pages.append(sparse)
    seeds.append(s)
    together = P.fit_images(pages, seeds)
    for k in range(3):
        alone = P.fit_images([pages[k]], [seeds[k]])[0]
        np.testing.assert_array_equal(together[k]["theta"], alone["theta"])
        np.testing.assert_array_equal(together[k]["centers_xy"], alone["centers_xy"])
