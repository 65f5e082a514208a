"""The page outline fit's oracle (oracle/outline.py) against known answers: the reference's own
geometry tests (ellipse widths are the axes, rays end on the outline), closed forms (superellipse
area, perimeter = d area / d offset, blur widths), finite differences, and a noise-free shape
fit and a noisy page fit that recover the outline they were drawn with. Also the public entry's
argument checks (CPU: they run before any device work)."""
import math

import numpy as np
import pytest

from oracle import outline as O


def _theta(t, a, b, p=2.0, phi=0.0, coeffs=None, **blur):
    coeffs = np.zeros((len(t.ks), 2)) if coeffs is None else np.asarray(coeffs, float)
    return t.pack(a, b, p, phi, coeffs, t.blur_pack(**(blur or dict(s1x=1.0, s1y=1.0))))


@pytest.fixture(scope="module")
def t24():
    return O.Template(upsample=8, size_px=24)


def test_ellipse_widths_are_the_axes_and_the_area_is_exact(t24):
    # reference tests/test_outline_fit.py:26-35
    th = _theta(t24, 4.0, 2.0)
    cx, cy = O.ellipse_widths(t24, th, np.array([0.0]))
    assert cx[0] == pytest.approx(8.0, rel=1e-6) and cy[0] == pytest.approx(4.0, rel=1e-6)
    d = np.array([-0.3, 0.2])
    cx, cy = O.ellipse_widths(t24, th, d)
    cd = np.array([t24.cd_px(th, float(v)) for v in d])
    np.testing.assert_allclose(cx * cy, cd ** 2, rtol=1e-9)


def test_rays_end_on_the_outline_about_its_centroid(t24):
    # reference tests/test_outline_fit.py:38-54
    th = _theta(t24, 3.0, 2.5, phi=0.3, coeffs=[[0.08, -0.05], [0.02, 0.0]])
    rays = O.outline_rays(t24, th, np.array([0.0, 0.1]), n_rays=64)
    assert rays.shape == (2, 64) and (rays > 0).all()
    cx, cy = t24.centroid_px(th)
    ang = 2 * np.pi * np.arange(64) / 64
    for j in (0, 13, 40):
        x, y = cx + rays[0, j] * math.cos(ang[j]), cy + rays[0, j] * math.sin(ang[j])
        r, _ = t24.radius_at(th, np.array([math.atan2(y, x)]))
        assert math.hypot(x, y) == pytest.approx(r[0], abs=1e-9)
    assert np.all(rays[1] - rays[0] > 0.099) and np.all(rays[1] - rays[0] < 0.13)


@pytest.mark.parametrize("a,b,p", [(4.0, 2.0, 2.0), (3.0, 3.0, 2.0), (3.5, 2.5, 4.0)])
def test_coverage_area_is_the_superellipse_area(t24, a, b, p):
    th = _theta(t24, a, b, p=p)
    exact = 4 * a * b * math.gamma(1 + 1 / p) ** 2 / math.gamma(1 + 2 / p)
    cover = t24.indicator(th).sum() / t24.U ** 2
    assert cover == pytest.approx(exact, rel=2e-4)
    assert t24.outline_area(th) == pytest.approx(exact, rel=1e-6)
    # blur and pixel box keep the integral; so does the extra box of the shape stage
    assert t24.render(th).sum() == pytest.approx(t24.indicator(th).sum(), rel=1e-10)
    assert t24.render(th, box_px=1 / 8).sum() == pytest.approx(t24.indicator(th).sum(), rel=1e-10)


def test_edge_offset_terms_follow_the_perimeter(t24):
    r = 3.0
    th = _theta(t24, r, r)
    t0, t1, t2 = t24.render_quad(th)
    # d area / d delta = perimeter; d^2 area / d delta^2 / 2 = pi for a circle (the second
    # difference over +-0.1 px sees the coverage's 1/32 px sub-sample steps: ~10 % off)
    assert t1.sum() / t24.U ** 2 == pytest.approx(2 * math.pi * r, rel=2e-3)
    assert t2.sum() / t24.U ** 2 == pytest.approx(math.pi, rel=0.12)
    # the offset outline of a circle is the circle of radius r + delta
    assert t24.render(th, 0.2).sum() == pytest.approx(t24.render(_theta(t24, r + 0.2, r + 0.2)).sum(), rel=2e-4)


def test_blur_widths_have_their_closed_forms(t24):
    th = _theta(t24, 2.0, 2.0, s1x=1.1, s1y=0.9, s2x=2.4, s2y=2.0, w2=0.1)
    sx, sy = t24.blur_sigma_px(th)
    assert sx == pytest.approx(math.sqrt(0.9 * 1.1 ** 2 + 0.1 * 2.4 ** 2), rel=1e-3)
    assert sy == pytest.approx(math.sqrt(0.9 * 0.9 ** 2 + 0.1 * 2.0 ** 2), rel=1e-3)
    t1 = O.Template(upsample=8, size_px=24, blur="gauss1")
    fw = t1.blur_fwhm_px(t1.pack(2.0, 2.0, 2.0, 0.0, np.zeros((2, 2)), t1.blur_pack(sx=1.3, sy=0.8)))
    assert fw[0] == pytest.approx(2 * math.sqrt(2 * math.log(2)) * 1.3, rel=1e-2)
    assert fw[1] == pytest.approx(2 * math.sqrt(2 * math.log(2)) * 0.8, rel=1e-2)
    tl = O.Template(upsample=8, size_px=48, blur="lens")
    otf = tl.otf(tl.pack(2.0, 2.0, 2.0, 0.0, np.zeros((2, 2)), tl.blur_pack(sx=0.01, sy=0.01, cutoff=0.5)))
    assert otf[0, 0] == pytest.approx(1.0) and np.all(otf[tl.nu >= 0.5] < 1e-12)


def test_spatial_derivatives_match_finite_differences(t24):
    th = _theta(t24, 3.0, 2.2, phi=0.2, s1x=1.2, s1y=1.0)
    t0 = t24.render(th)
    dx, dy = t24.derivatives_xy(t0)
    U = t24.U
    fdx = (np.roll(t0, -1, axis=1) - np.roll(t0, 1, axis=1)) * U / 2
    fdy = (np.roll(t0, -1, axis=0) - np.roll(t0, 1, axis=0)) * U / 2
    assert np.abs(dx - fdx).max() < 5e-3 * np.abs(dx).max()
    assert np.abs(dy - fdy).max() < 5e-3 * np.abs(dy).max()
    # a planted sign error would fail: the derivative of a centred bump is odd, negative to the right
    c = t24.N // 2
    assert dx[c, c + 3 * U] < 0 and dx[c, c - 3 * U] > 0


def test_greedy_colours_separate_neighbours():
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 100, (300, 2))
    col = O.greedy_colours(c, 12.0)
    d = np.hypot(c[:, None, 0] - c[None, :, 0], c[:, None, 1] - c[None, :, 1])
    i, j = np.nonzero((d <= 12.0) & (d > 0))
    assert np.all(col[i] != col[j])
    # each hole takes the smallest colour its earlier neighbours leave free
    for k in range(len(c)):
        used = set(col[(d[k] <= 12.0) & (np.arange(len(c)) < k)].tolist())
        assert col[k] == min(set(range(len(used) + 1)) - used)


def test_shape_stage_recovers_a_noise_free_round_outline():
    t = O.Template(upsample=4, size_px=16)
    truth = _theta(t, 2.6, 2.6, s1x=1.1, s1y=1.1, s2x=2.2, s2y=2.2, w2=0.05)
    stack = 7.0 * t.render(truth, box_px=1 / 4) + 0.3
    ok = np.ones(stack.shape, bool)
    theta, scale, const, chi2 = O.fit_stages(t, lambda th: t.render(th, box_px=1 / 4)[ok], stack[ok],
                                             np.ones(ok.sum()), t.start(2.0, 1.0), stages=("round",))
    assert t.cd_px(theta) == pytest.approx(t.cd_px(truth), abs=1e-4)
    assert scale == pytest.approx(7.0, rel=1e-4) and const == pytest.approx(0.3, abs=1e-4)
    assert chi2 < 1e-8


def _lattice_page(seed=5, noise=4.0):
    rng = np.random.default_rng(seed)
    t = O.Template(upsample=8, size_px=24)
    th = t.pack(2.4, 2.2, 2.6, 0.1, [[0.03, -0.02], [0.01, 0.0]],
                t.blur_pack(s1x=1.2, s1y=1.1, s2x=2.5, s2y=2.5, w2=0.08))
    X, Y = np.meshgrid(np.arange(6, 170, 10.0), np.arange(6, 120, 10.0))
    c = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.4, 0.4, (X.size, 2))
    a = rng.uniform(900, 1100, len(c))
    de = rng.normal(0, 0.08, len(c))
    page = O.render_model_page(t, th, c, a, 100.0, de, (120, 170), noise=noise, seed=seed)
    seeds = c + rng.uniform(-0.5, 0.5, c.shape)
    return dict(t=t, theta=th, c=c, a=a, delta=de, page=page, seeds=seeds)


def test_page_fit_recovers_the_outline_it_was_drawn_with():
    K = _lattice_page()
    t, th = K["t"], K["theta"]
    res = O.fit_outline(K["page"], K["seeds"], rounds=3)
    assert res["per_hole_size"]
    assert res["cd_outline_px"] == pytest.approx(t.cd_px(th), abs=0.05)
    truth_xy = K["c"] + np.array(t.centroid_px(th))
    ok = res["good"]
    assert ok.mean() > 0.5
    assert np.sqrt(np.mean(np.sum((res["centers_xy"] - truth_xy) ** 2, axis=1)[ok])) < 0.02
    true_cd = np.array([t.cd_px(th, float(d)) for d in K["delta"]])
    assert 0.85 < np.polyfit(true_cd[ok], res["cd_px"][ok], 1)[0] < 1.15
    assert abs(np.mean(res["cd_px"][ok] - true_cd[ok])) < 0.05


def test_fit_outline_checks_its_input():
    from imageprocessor_new_amd.outline import fit_outline

    img = np.zeros((40, 40))
    seeds = np.stack(np.meshgrid(np.arange(4, 40, 8.0), np.arange(4, 40, 8.0)), -1).reshape(-1, 2)
    with pytest.raises(ValueError, match="at least"):
        fit_outline(img, seeds[:3])
    with pytest.raises(ValueError, match="blur"):
        fit_outline(img, seeds, blur="airy")
    with pytest.raises(ValueError, match="rounds"):
        fit_outline(img, seeds, rounds=0)
    with pytest.raises(ValueError, match="n_rays"):
        fit_outline(img, seeds, n_rays=8)
    with pytest.raises(ValueError, match="image"):
        fit_outline(np.zeros((0, 4)), seeds)
    with pytest.raises(ValueError, match="pixel_size"):
        fit_outline(img, seeds, pixel_size=0.0)


def test_product_greedy_colours_equal_the_oracle_rule():
    from imageprocessor_new_amd.outline import greedy_colours

    rng = np.random.default_rng(12)
    for n, span, reach in ((400, 120.0, 15.0), (1000, 300.0, 12.5)):
        c = rng.uniform(0, span, (n, 2))
        assert np.array_equal(greedy_colours(c, reach), O.greedy_colours(c, reach))
    X, Y = np.meshgrid(np.arange(20) * 8.5, np.arange(15) * 8.5)
    c = np.stack([X.ravel(), Y.ravel()], 1) + rng.normal(0, 0.05, (300, 2))
    assert np.array_equal(greedy_colours(c, 1.5 * 8.5), O.greedy_colours(c, 1.5 * 8.5))


def test_fit_outline_fast_checks_its_arguments():
    """The reference's checks and error texts (tests/test_outline_fit_fast.py:115-126); without a
    GPU (or with backend='cpu') the method says it needs one."""
    from imageprocessor_new_amd import OutlineFastResult, fit_outline_fast  # noqa: F401

    img = np.ones((32, 32))
    seeds = np.stack([np.arange(12.0), np.arange(12.0)], 1)
    with pytest.raises(ValueError, match="nonempty 2D"):
        fit_outline_fast(np.zeros((0, 4)), seeds)
    with pytest.raises(ValueError, match="at least 10 finite seeds"):
        fit_outline_fast(img, seeds[:3])
    with pytest.raises(ValueError, match="blur must be one of"):
        fit_outline_fast(img, seeds, blur="box")
    with pytest.raises(ValueError, match="n_rays"):
        fit_outline_fast(img, seeds, n_rays=8)
    with pytest.raises(RuntimeError, match="needs the GPU"):
        fit_outline_fast(img, seeds, backend="cpu")
