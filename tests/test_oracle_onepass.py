"""Pins of the one-pass outline fit oracle (oracle/onepass.py).

- the reference's own checks (tests/test_outline_fit_fast.py:27-99): the analytic template against
  the rendered one, the model's derivatives against central differences, a known outline recovered
  by the page fit from the grid of starts and from the truth, the gauss2 exchange;
- closed forms: a rectangle's transform, one band coefficient's cosine on the grid, the pair map
  against a direct sum over the pairs, the plain stack and the hole-vs-stack fit on a page of
  identical holes, the smoothing against scipy's Gaussian filter, the cell-list nearest distance
  against brute force;
- the whole fit on a page drawn from a known outline and blur (steps A to C).
"""
import math

import numpy as np
import pytest

from oracle import holes as OH
from oracle import onepass as OP
from oracle.outline import Template

PIXEL_NM = 21.0


@pytest.mark.parametrize("blur, box", [("gauss2", 24), ("lens", 48)])
def test_analytic_template_matches_the_rendered_one(blur, box):
    t = Template(upsample=8, size_px=box, blur=blur)
    band = OP.Band(box_px=box, f_max=1.0, upsample=8)
    rng = np.random.default_rng(1)
    for _ in range(2):
        a, b = rng.uniform(1.6, 3.2, 2)
        bp = t.blur_pack(s1x=rng.uniform(1.0, 1.5), s1y=rng.uniform(1.0, 1.5), s2x=rng.uniform(2, 3),
                         s2y=rng.uniform(2, 3), w2=rng.uniform(0.02, 0.2), sx=rng.uniform(0.6, 1.2),
                         sy=rng.uniform(0.6, 1.2), cutoff=rng.uniform(0.35, 0.6))
        theta = t.pack(a, b, rng.uniform(1.7, 3.5), rng.uniform(-0.3, 0.3), rng.uniform(-0.03, 0.03, (2, 2)), bp)
        for delta, bin_px in ((0.0, 0.0), (0.1, 0.0), (-0.1, 0.125)):
            rendered = t.render(theta, delta, box_px=bin_px)
            analytic = band.to_grid(OP.analytic_hat(t, theta, band.fx, band.fy, 512, delta, box_px=bin_px))
            assert np.abs(analytic - rendered).max() < 3e-4 * rendered.max()


def test_polygon_transform_of_a_rectangle_and_band_synthesis():
    # rectangle [-a, a] x [-b, b] (counter-clockwise): 4 a b sinc(2 a fx) sinc(2 b fy)
    a, b = 1.7, 0.6
    x, y = np.array([a, -a, -a, a]), np.array([b, b, -b, -b])
    fx, fy = np.meshgrid(np.linspace(-0.9, 0.9, 13), np.linspace(-0.7, 0.8, 11))
    np.testing.assert_allclose(OP.polygon_hat(x, y, fx, fy), 4 * a * b * np.sinc(2 * a * fx) * np.sinc(2 * b * fy),
                               atol=1e-12)
    # one coefficient c at f = (kx, ky) / L: the grid is 2 Re(c e^{2 pi i f u}) / L^2 (its mirror is implied)
    band = OP.Band(box_px=16, f_max=0.5, upsample=4)
    coeff = np.zeros(band.fx.shape, complex)
    r = int(np.flatnonzero(band.rows == band.N - 3)[0])  # ky = -3
    coeff[r, 2] = 0.7 - 0.4j
    u = (np.arange(band.N) - band.N // 2) / band.U
    want = 2 * np.real((0.7 - 0.4j) * np.exp(2j * np.pi * (2 / 16 * u[None, :] - 3 / 16 * u[:, None]))) / 16 ** 2
    np.testing.assert_allclose(band.to_grid(coeff), want, atol=1e-14)


@pytest.mark.parametrize("blur", ["gauss2", "lens"])
def test_model_derivatives_match_central_differences(blur):
    t = Template(upsample=8, size_px=24, blur=blur)
    sf = OP.ShapeFitter(t, box_px=48, f_max=0.6, U=4, R=7, n_poly=256)
    theta = t.pack(2.9, 2.1, 2.6, 0.3, [[0.03, -0.02], [0.02, 0.015]],
                   t.blur_pack(s1x=1.2, s1y=1.35, s2x=2.4, s2y=2.1, w2=0.08, sx=1.1, sy=1.25, cutoff=0.45))
    gamma = np.ones((1, sf.U, sf.U, sf.nky, sf.nkx), dtype=complex)
    model = sf.synth(sf.template_hat(theta[None], np.eye(sf.P)), gamma)[0]
    for k in range(sf.P):
        step = np.zeros(sf.P)
        step[k] = 1e-5
        numeric = (sf.synth(sf.template_hat((theta + step)[None]), gamma)[0, 0]
                   - sf.synth(sf.template_hat((theta - step)[None]), gamma)[0, 0]) / 2e-5
        assert np.abs(model[1 + k] - numeric).max() < 1e-3 * np.abs(numeric).max(), k


@pytest.mark.parametrize("blur", ["gauss2", "lens"])
def test_page_fit_recovers_a_known_outline(blur):
    t = Template(upsample=8, size_px=24, blur=blur)
    sf = OP.ShapeFitter(t, box_px=48, f_max=0.6, U=4, R=7, n_poly=256)
    truth = t.pack(2.4, 2.3, 2.2, 0.2, np.zeros((2, 2)),
                   t.blur_pack(s1x=1.27, s1y=1.27, s2x=2.5, s2y=2.5, w2=0.05, sx=1.2, sy=1.2, cutoff=0.45))
    gamma = np.ones((1, sf.U, sf.U, sf.nky, sf.nkx), dtype=complex)
    mask = sf.pad[:, None, :, None] & sf.pad[None, :, None, :]
    sd = 0.16
    clean = 180.0 * sf.synth(sf.template_hat(truth[None]), gamma)[0, 0] + 40.0
    y = np.where(mask, clean + np.random.default_rng(0).normal(0, sd, clean.shape), 0.0)
    w = np.where(mask, 1.0 / sd ** 2, 0.0)
    true_cd = t.cd_px(truth) * PIXEL_NM
    found = []
    for start, grid in ((t.start(1.7, 1.0), True), (truth, False)):
        theta, scale, const, cost, info = sf.fit(start[None], gamma, y[None], w[None], grid=grid)
        found.append(t.cd_px(theta[0]) * PIXEL_NM)
        assert abs(found[-1] - true_cd) < 0.05
        assert 0.9 < cost[0] / mask.sum() < 1.1 and scale[0] == pytest.approx(180, rel=0.02)
        assert info[-1][0] == "full" and info[-1][2].shape == (1,)
    assert found[0] == pytest.approx(found[1], abs=1e-3)


def test_gauss2_exchange_is_the_same_blur_under_the_other_labels():
    t = Template(upsample=8, size_px=24, blur="gauss2")
    sf = OP.ShapeFitter(t, box_px=56, f_max=0.6, U=4, R=8, n_poly=256)
    theta = t.pack(2.4, 2.3, 2.2, 0.1, [[0.01, -0.02], [0.015, 0.005]],
                   t.blur_pack(s1x=1.6, s1y=1.5, s2x=1.2, s2y=1.1, w2=0.45))
    other = OP.exchanged(t, theta)
    assert OP.second_is_narrower(t, theta) and not OP.second_is_narrower(t, other)
    np.testing.assert_allclose(sf.template_hat(other[None]), sf.template_hat(theta[None]), rtol=0, atol=1e-12)
    np.testing.assert_allclose(OP.exchanged(t, other), theta, atol=1e-12)
    edge = OP.exchanged(t, t.pack(2.4, 2.3, 2.2, 0.1, np.zeros((2, 2)), t.blur_pack(w2=0.05)))
    assert 1 / (1 + np.exp(-edge[-1])) == pytest.approx(0.6, abs=1e-3)


def test_smoothing_and_nearest_against_library_and_brute_force():
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(3)
    img = rng.normal(100, 20, (37, 53))
    for sigma in (1.0, 2.3):
        np.testing.assert_allclose(OP.smooth(img, sigma), gaussian_filter(img, sigma, mode="nearest", truncate=2.5),
                                   rtol=0, atol=1e-12)
    c = rng.uniform(0, 60, (80, 2))
    cells = OP.CellList(c, 60, 60, 16.0)
    nn = OP.nearest(c, cells)
    brute = OH.nearest_distance(c)
    near = brute <= 16.0  # within one cell side the 3 x 3 cells always hold the nearest
    np.testing.assert_allclose(nn[near], brute[near], rtol=1e-15)
    assert (nn >= brute - 1e-12).all()


def _identical_holes(shape=(64, 80), pitch=12, frac=(0.25, 0.5), amp=None, lev=None):
    """A page of identical linear-ramp holes (no overlap) at a common sub-pixel position."""
    h, w = shape
    xs, ys = np.meshgrid(np.arange(8, w - 8, pitch) + frac[0], np.arange(8, h - 8, pitch) + frac[1])
    c = np.stack([xs.ravel(), ys.ravel()], 1)
    n = len(c)
    amp = np.ones(n) if amp is None else amp
    lev = np.zeros(n) if lev is None else lev
    img = np.zeros(shape)
    R = 4
    for (x, y), ai, li in zip(c, amp, lev):
        for py in range(int(np.rint(y)) - R - 1, int(np.rint(y)) + R + 2):
            for px in range(int(np.rint(x)) - R - 1, int(np.rint(x)) + R + 2):
                img[py, px] = li + ai * (50.0 + 3.0 * (px - x) - 2.0 * (py - y))
    return img, c


def test_plain_stack_and_versus_on_identical_holes():
    rng = np.random.default_rng(4)
    img0, c = _identical_holes()
    pg = OP.Page(img0, c)
    assert pg.R == 9 and pg.spacing == pytest.approx(12.0)
    pg.R = 4  # windows inside each hole's drawn square
    pg.Ms = 2 * int(math.ceil((pg.R + 0.5) * pg.U)) + 1
    stack, wsum = pg.stack_of(np.ones(pg.n, bool))
    n = (np.arange(pg.Ms) - pg.Ms // 2) / pg.U
    fin = np.isfinite(stack)
    want = 50.0 + 3.0 * n[None, :] - 2.0 * n[:, None]
    np.testing.assert_allclose(stack[fin], want[fin], atol=1e-12)
    assert fin.sum() == 81 and set(np.unique(wsum[fin])) == {float(pg.n)}
    amp, lev = rng.uniform(0.5, 2.0, pg.n), rng.uniform(-5, 5, pg.n)
    img, _ = _identical_holes(amp=amp, lev=lev)
    pg.img = img.astype(np.float32)
    o = pg.versus(stack, 3.0)
    np.testing.assert_allclose(o[:, 0], amp, rtol=1e-5)
    np.testing.assert_allclose(o[:, 1], lev, atol=1e-3)
    assert (o[:, 2] < 1e-3).all() and (o[:, 5] == 81).all()


def test_pair_map_is_the_sum_over_pairs():
    rng = np.random.default_rng(6)
    h, w = 70, 90
    xs, ys = np.meshgrid(np.arange(12, w - 8, 9.3), np.arange(12, h - 8, 9.7))
    c = np.stack([xs.ravel(), ys.ravel()], 1) + rng.uniform(-0.4, 0.4, (xs.size, 2))
    pg = OP.Page(np.zeros((h, w)), c)
    pg.use = pg.interior()
    pg.a = rng.uniform(0.8, 1.2, pg.n)
    sf = OP.ShapeFitter(pg.t, box_px=pg.L, f_max=pg.f_max, U=pg.U, R=pg.R)
    gam = pg.pair_map(sf).reshape(pg.U * pg.U, sf.nky, sf.nkx)
    ky, kx = np.meshgrid(np.arange(-sf.kmax, sf.kmax + 1), np.arange(sf.kmax + 1), indexing="ij")
    fx, fy = kx / pg.L, ky / pg.L
    want = np.zeros_like(gam)
    U = pg.U
    v = -pg.c * U
    s = (v - np.rint(v)) / U
    q = np.mod(np.rint(v), U).astype(int)
    cls = q[:, 1] * U + q[:, 0]
    amp = pg.a / pg.a[pg.use].mean()
    count = np.bincount(cls[pg.use], minlength=U * U)
    for i in np.flatnonzero(pg.use):
        d = pg.c - pg.c[i] - s[i]
        near = (np.abs(pg.c[:, 0] - pg.c[i, 0]) <= pg.reach) & (np.abs(pg.c[:, 1] - pg.c[i, 1]) <= pg.reach)
        for j in np.flatnonzero(near):
            want[cls[i]] += amp[j] / count[cls[i]] * np.exp(-2j * np.pi * (fx * d[j, 0] + fy * d[j, 1]))
    assert np.abs(gam - want).max() < 2e-4 * np.abs(want).max()


def test_find_seeds_on_gaussian_blobs():
    rng = np.random.default_rng(8)
    h, w = 80, 100
    xs, ys = np.meshgrid(np.arange(10, w - 6, 11.0), np.arange(10, h - 6, 11.0))
    c = np.stack([xs.ravel(), ys.ravel()], 1) + rng.uniform(-0.5, 0.5, (xs.size, 2))
    yy, xx = np.indices((h, w), dtype=float)
    img = np.full((h, w), 20.0)
    for x, y in c:
        img += 200.0 * np.exp(-0.5 * ((xx - x) ** 2 + (yy - y) ** 2) / 1.6 ** 2)
    seeds = OP.find_seeds(img.astype(np.float32))
    assert len(seeds) == len(c)
    d = np.hypot(*(seeds[:, None, :] - c[None]).transpose(2, 0, 1)).min(axis=1)
    assert d.max() < 0.1


@pytest.mark.parametrize("blur", ["gauss2", "lens"])
def test_one_pass_fit_recovers_the_drawn_outline(blur):
    t, theta, c, page, seeds = OP.drawn_page(blur)
    res = OP.fit_images([page], [seeds], blur=blur)[0]
    cd = t.cd_px(res["theta"])
    assert cd == pytest.approx(t.cd_px(theta), abs=0.01)
    ok = res["good"]
    assert ok.sum() > 0.6 * len(c)
    truth = c + np.array(t.centroid_px(theta))
    d = np.hypot(*(res["centers_xy"][ok] + np.array(t.centroid_px(res["theta"])) - truth[ok]).T)
    assert np.sqrt(np.mean(d ** 2)) < 0.02
    assert 0.5 < res["chi2"] / res["bins"] < 2.0
