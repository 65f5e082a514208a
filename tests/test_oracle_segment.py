"""Pins of the front-end oracle (oracle/segment.py) against closed forms and brute force."""

import numpy as np
import pytest

from oracle import segment


def _bimodal(seed, shape=(24, 31), lo=(200, 20), hi=(900, 60), frac=0.3):
    rng = np.random.default_rng(seed)
    fg = rng.random(shape) < frac
    img = np.where(fg, rng.normal(hi[0], hi[1], shape), rng.normal(lo[0], lo[1], shape))
    return np.clip(np.rint(img), 0, 65535).astype(np.uint16)


@pytest.mark.parametrize("seed", range(8))
def test_otsu_matches_brute_force(seed):
    img = _bimodal(seed, frac=0.1 + 0.08 * seed)
    assert segment.otsu_u16(img) == segment.otsu_brute_force(img)


def test_otsu_two_levels_and_constant():
    img = np.array([[10, 10, 10, 50], [10, 50, 50, 50]], dtype=np.uint16)
    assert segment.otsu_u16(img) == 10  # lowest t separating the classes
    assert segment.otsu_u16(np.full((5, 5), 7, np.uint16)) == 0  # no split: OpenCV returns 0


def test_otsu_wrong_rule_would_fail():
    # Planted mistake: taking the last maximum instead of the first changes the answer
    # on a histogram with an empty gap between the classes.
    img = np.array([[100] * 6 + [200] * 6], dtype=np.uint16)
    t = segment.otsu_u16(img)
    assert t == 100 and segment.otsu_brute_force(img) == 100


def test_threshold_polarity():
    img = np.array([[1, 5, 9]], dtype=np.uint16)
    assert segment.threshold(img, 5).tolist() == [[False, False, True]]
    assert segment.threshold(img, 5, invert=True).tolist() == [[True, True, False]]


def _brute_cc(mask, conn):
    H, W = mask.shape
    lab = np.zeros((H, W), np.int64)
    nxt = 0
    nb = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)] if conn == 8 else \
        [(-1, 0), (0, -1), (0, 1), (1, 0)]
    for y in range(H):
        for x in range(W):
            if mask[y, x] and not lab[y, x]:
                nxt += 1
                stack = [(y, x)]
                lab[y, x] = nxt
                while stack:
                    cy, cx = stack.pop()
                    for dy, dx in nb:
                        yy, xx = cy + dy, cx + dx
                        if 0 <= yy < H and 0 <= xx < W and mask[yy, xx] and not lab[yy, xx]:
                            lab[yy, xx] = nxt
                            stack.append((yy, xx))
    return lab, nxt


@pytest.mark.parametrize("conn", [8, 4])
@pytest.mark.parametrize("seed", range(4))
def test_cc_partition_stats_and_order(conn, seed):
    rng = np.random.default_rng(100 + seed)
    mask = rng.random((23, 37)) < 0.35
    num, labels, stats, cents = segment.connected_components(mask, conn)
    num2, labels2, stats2, cents2 = segment.connected_components_fast(mask, conn)
    assert num == num2 and np.array_equal(labels, labels2) and np.array_equal(stats, stats2)
    np.testing.assert_array_equal(cents, cents2)
    blab, bn = _brute_cc(mask, conn)
    assert num == bn + 1
    # same partition: a bijection between label sets
    pairs = set(zip(labels[mask].tolist(), blab[mask].tolist()))
    assert len(pairs) == bn
    W = mask.shape[1]
    # numbering: first 2x2 block (8-conn) or first pixel (4-conn) in raster order
    keys = []
    for k in range(1, num):
        ys, xs = np.nonzero(labels == k)
        key = (ys // 2) * ((W + 1) // 2) + xs // 2 if conn == 8 else ys * W + xs
        keys.append(int(key.min()))
        assert stats[k, 4] == ys.size
        assert stats[k, 0] == xs.min() and stats[k, 2] == xs.max() - xs.min() + 1
        assert cents[k, 0] == xs.sum() / ys.size
    assert keys == sorted(keys)
    assert stats[0, 4] == (~mask).sum()


def test_logquad_exact_on_sampled_gaussian():
    # log of a sampled Gaussian is exactly quadratic: the fit returns the centre.
    yy, xx = np.indices((23, 23)).astype(np.float64)
    cx, cy = 11.37, 10.81
    g = 50.0 + 2000.0 * np.exp(-((xx - cx) ** 2 / (2 * 2.1 ** 2) + (yy - cy) ** 2 / (2 * 1.7 ** 2)))
    ring = np.concatenate([g[:2].ravel(), g[-2:].ravel(), g[:, :2].ravel(), g[:, -2:].ravel()])
    g = g - np.median(ring) + 50.0  # make the border median exactly the offset
    mask = (g - 50.0) > 30.0
    px, py = segment.refine_logquadratic(g, mask)
    # the border median is not exactly the pedestal: compare with the closed form on the
    # values the method actually fits
    bg = segment.bg_from_border(g)
    v = g - bg
    assert np.all(v[mask] > 0)
    assert abs(px - cx) < 0.05 and abs(py - cy) < 0.05
    # exact when the pedestal is removed exactly
    g2 = 2000.0 * np.exp(-((xx - cx) ** 2 / (2 * 2.1 ** 2) + (yy - cy) ** 2 / (2 * 1.7 ** 2)))
    g2[:2, :] = g2[-2:, :] = g2[:, :2] = g2[:, -2:] = 0.0
    px, py = segment.refine_logquadratic(g2, g2 > 1.0)
    assert abs(px - cx) < 1e-9 and abs(py - cy) < 1e-9


def test_logquad_translation_equivariant_and_gates():
    rng = np.random.default_rng(3)
    yy, xx = np.indices((30, 30))

    def blob(cx, cy):
        g = 1500 * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / 8.0) * (1 + 0.01 * rng.standard_normal((30, 30)))
        g[:2, :] = g[-2:, :] = g[:, :2] = g[:, -2:] = 0.0  # border median exactly 0
        return g

    rng = np.random.default_rng(3)
    g = blob(12.3, 13.2)
    m = g > 300
    a = segment.refine_logquadratic(g, m)
    g2 = np.zeros_like(g)
    g2[2:, 3:] = g[:-2, :-3]
    g2[:2, :] = g2[-2:, :] = g2[:, :2] = g2[:, -2:] = 0.0
    b = segment.refine_logquadratic(g2, g2 > 300)
    assert abs(b[0] - a[0] - 3) < 1e-9 and abs(b[1] - a[1] - 2) < 1e-9
    tiny = np.zeros((10, 10), bool)
    tiny[4, 4:8] = True
    assert np.isnan(segment.refine_logquadratic(g[:10, :10], tiny)[0])  # < 6 pixels


def test_triangle_is_the_farthest_bin_from_the_peak_line():
    # geometric restatement: perpendicular distance from (i, h_i) to the line through
    # (lo, 0) and (peak, h_peak), maximised over lo < i <= peak; OpenCV then steps one bin
    # towards the tail. Random long-tailed histograms on both sides of the peak.
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = 20000
        vals = np.concatenate([rng.normal(200 if trial % 2 else 40, 8, n), rng.uniform(0, 255, n // 4)])
        img = np.clip(np.rint(vals), 0, 255).astype(np.uint8).reshape(-1, 50)
        t = segment.triangle_u8(img)
        h = np.bincount(img.ravel(), minlength=256).astype(np.float64)
        nz = np.nonzero(h)[0]
        lo, hi, peak = max(nz[0] - 1, 0), min(nz[-1] + 1, 255), int(np.argmax(h))
        flip = peak - lo < hi - peak
        if flip:
            h, lo, peak = h[::-1], 255 - hi, 255 - peak
        i = np.arange(lo + 1, peak + 1)
        p0, p1 = np.array([lo, 0.0]), np.array([peak, h[peak]])
        dvec = (p1 - p0) / np.hypot(*(p1 - p0))
        rel = np.stack([i - p0[0], h[i] - p0[1]], axis=1)
        dist = rel[:, 0] * dvec[1] - rel[:, 1] * dvec[0]
        want = int(i[np.argmax(dist)]) - 1
        assert t == (255 - want if flip else want)


def test_triangle_rejects_16_bit_levels():
    with pytest.raises(ValueError):
        segment.triangle_u8(np.array([[0, 300]], np.uint16))


def test_morphology_matches_scipy():
    from scipy import ndimage

    cross = ndimage.generate_binary_structure(2, 1)
    rng = np.random.default_rng(4)
    m = rng.random((37, 41)) < 0.55
    np.testing.assert_array_equal(segment.morph_cross(m, "erode"),
                                  ndimage.binary_erosion(m, cross, border_value=1))
    np.testing.assert_array_equal(segment.morph_cross(m, "dilate"),
                                  ndimage.binary_dilation(m, cross, border_value=0))
    img = (rng.random((37, 41)) * 100).astype(np.uint16)
    got = segment.binary_mask(img, 40, morph_open=2, morph_close=1)
    ref = img > 40
    ref = ndimage.binary_erosion(ref, cross, iterations=2, border_value=1)
    ref = ndimage.binary_dilation(ref, cross, iterations=2, border_value=0)
    ref = ndimage.binary_dilation(ref, cross, iterations=1, border_value=0)
    ref = ndimage.binary_erosion(ref, cross, iterations=1, border_value=1)
    np.testing.assert_array_equal(got, ref)
