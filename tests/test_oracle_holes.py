"""Pins of the hole-stage oracle (oracle/holes.py) by what the mathematics fixes: Keys
interpolation reproduces quadratics, the disk templates against finite differences of the
erf profile, the colour rule's separation, the bounded Gauss-Newton step against a direct
solve, the far light of one hole against the far kernel, and the hole pass recovering a
noise-free page's holes from perturbed starts (a sign or index slip in the Jacobian, the
weights or the render bookkeeping stops that)."""
import math

import numpy as np
import pytest
from scipy.special import erfc

from oracle import holes as O


def _lattice(n_side, spacing, margin, rng, jitter=0.2):
    g = margin + spacing * np.arange(n_side)
    X, Y = np.meshgrid(g, g)
    c = np.stack([X.ravel(), Y.ravel()], axis=1).astype(np.float64)
    return c + rng.uniform(-jitter, jitter, c.shape)


def test_keys_partition_and_quadratic_reproduction():
    rng = np.random.default_rng(0)
    t = rng.uniform(0, 1, 500)
    w = O.keys_weights(t)
    assert np.allclose(w.sum(axis=1), 1.0, atol=1e-14)
    taps = np.array([-1.0, 0.0, 1.0, 2.0])
    # sum_j w_j (tap_j - t)^m = 0 for m = 1, 2 (quadratics reproduced)
    for m in (1, 2):
        assert np.allclose((w * (taps[None, :] - t[:, None]) ** m).sum(axis=1), 0.0, atol=1e-13)
    assert np.allclose(O.keys_weights(np.zeros(1)), [[0.0, 1.0, 0.0, 0.0]])


def test_sample_reproduces_quadratic_grid():
    N = 40
    j, i = np.meshgrid(np.arange(N, dtype=float), np.arange(N, dtype=float))
    f = lambda x, y: 0.3 * x * x - 0.2 * y * y + 0.1 * x * y + 2.0 * x - y + 5.0 + 0.01 * x * x * y * y
    grids = np.stack([f(j, i), 2.0 * f(j, i)], axis=-1)
    rng = np.random.default_rng(1)
    sx, sy = rng.uniform(2, N - 4, 300), rng.uniform(2, N - 4, 300)
    v = O.sample(grids, sx, sy)
    assert np.allclose(v[:, 0], f(sx, sy), rtol=1e-12, atol=1e-9)
    assert np.allclose(v[:, 1], 2.0 * f(sx, sy), rtol=1e-12, atol=1e-9)
    # taps off the grid read 0: a point beyond the last column sees only in-grid taps
    assert O.sample(grids, np.array([N + 5.0]), np.array([10.0]))[0, 0] == 0.0


def test_disk_templates_against_finite_differences():
    radius, sigma, R, U = 3.5, 1.1, 9, 8
    T = O.disk_templates(radius, sigma, R, U)
    N = T["N"]
    assert N == O.template_box(R) * U
    xg = (np.arange(N) - N // 2) / U
    X, Y = np.meshgrid(xg, xg)
    rho = np.hypot(X, Y)
    core = rho < T["r0"] - 0.5  # inside the near window, w = 1
    prof = lambda d: 0.5 * erfc((rho - radius - d) / (math.sqrt(2.0) * sigma))
    h = 1e-4
    assert np.allclose(T["g5"][..., 0][core], prof(0.0)[core], atol=1e-15)
    assert np.allclose(T["g5"][..., 1][core], ((prof(h) - prof(-h)) / (2 * h))[core], atol=1e-7)
    assert np.allclose(T["g5"][..., 4][core], ((prof(h) + prof(-h) - 2 * prof(0.0)) / (2 * h * h))[core], atol=1e-5)
    # spatial derivatives against central differences on the fine grid (spacing 1/U px)
    t0w = T["g5"][..., 0]
    dx = (t0w[:, 2:] - t0w[:, :-2]) * U / 2
    dy = (t0w[2:, :] - t0w[:-2, :]) * U / 2
    assert np.max(np.abs(T["g5"][:, 1:-1, 2] - dx)) < 2e-3
    assert np.max(np.abs(T["g5"][1:-1, :, 3] - dy)) < 2e-3
    # the far kernel is the complement of the near window: zero inside r0
    B = T["far"].shape[0]
    xb = np.arange(B) - B // 2
    rb = np.hypot(*np.meshgrid(xb, xb))
    assert np.all(T["far"][rb <= T["r0"]] == 0.0)
    assert T["Rr"] >= T["r0"] and T["r_light"] == T["r0"] + 2.0


def test_colours_separate_holes_by_one_and_a_half_spacings():
    rng = np.random.default_rng(2)
    c = _lattice(14, 11.0, 8.0, rng, jitter=1.5)
    sp = float(np.median(O.nearest_distance(c)))
    chunks = O.colours(c, 170, 170, sp)
    allidx = np.sort(np.concatenate(chunks))
    assert np.array_equal(allidx, np.arange(len(c)))
    for ix in chunks:
        if len(ix) > 1:
            assert O.nearest_distance(c[ix]).min() >= 1.5 * sp - 1e-9
    assert O.window_half(sp) == max(3, math.ceil(0.75 * sp))


def test_hole_update_is_the_bounded_direct_solve():
    rng = np.random.default_rng(3)
    n = 20
    c = rng.uniform(10, 50, (n, 2))
    a, level, delta = rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n), rng.uniform(-0.1, 0.1, n)
    acc = np.zeros((n, 22))
    mats, rhss = [], []
    for i in range(n):
        M = rng.normal(size=(8, 5))
        A = M.T @ M
        b = rng.normal(size=5) * 0.05
        q = 0
        for r in range(5):
            for s in range(r, 5):
                acc[i, q] = A[r, s]
                q += 1
        acc[i, 15:20] = b
        acc[i, 20], acc[i, 21] = 4.0 * 9, 9
        mats.append(A)
        rhss.append(b)
    idx = np.arange(n)
    c1, a1, l1, d1 = c.copy(), a.copy(), level.copy(), delta.copy()
    stat, c_old = O.hole_update(acc, idx, c1, a1, l1, d1, True, 0.25, 0.3, 1.0, 1.0)
    for i in range(n):
        A = mats[i] + np.eye(5) * 1e-9 * np.trace(mats[i]) / 5
        s = np.linalg.solve(A, rhss[i])
        assert np.allclose(c1[i] - c[i], np.clip(s[:2], -0.25, 0.25), atol=1e-12)
        assert np.isclose(a1[i], min(max(min(max(a[i] + s[2], 0.5 * a[i]), 2 * a[i]), 1e-3), 5.0), atol=1e-12)
        assert np.isclose(l1[i], level[i] + s[3], atol=1e-12)
        assert np.isclose(d1[i], np.clip(delta[i] + s[4], -0.3, 0.3), atol=1e-12)
        assert np.isclose(stat[i, 2], 2.0)
        assert np.isclose(stat[i, 3], math.sqrt(np.linalg.inv(A)[4, 4]), rtol=1e-9)
    assert np.array_equal(c_old, c)


def test_far_field_of_one_hole_is_the_far_kernel():
    rng = np.random.default_rng(4)
    B, h, w = 32, 70, 80
    far = rng.normal(size=(B, B))
    c = np.array([[30.0, 40.0], [30.25, 40.5]])
    geom = O.far_geometry(c, h, w, B)
    out = O.far_field(c[:1], np.array([2.0]), np.zeros(1), 7.0, far, h, w, geom)
    ys, xs = np.mgrid[0:h, 0:w]
    ky, kx = ys - 40 + B // 2, xs - 30 + B // 2
    inb = (ky >= 0) & (ky < B) & (kx >= 0) & (kx < B)
    exp = np.zeros((h, w))
    exp[inb] = 2.0 * far.astype(np.float32)[ky[inb], kx[inb]]
    assert np.allclose(out, exp, atol=1e-5)
    # a fractional position is the Keys interpolation of the kernel; delta scales by 1 + 4 delta / cd0
    out2 = O.far_field(c[1:], np.array([1.0]), np.array([0.35]), 7.0, far, h, w, geom)
    wx, wy = O.keys_weights(np.array(0.25)), O.keys_weights(np.array(0.5))
    exp2 = np.zeros((h, w))
    for q in range(4):
        for p in range(4):
            sh_y, sh_x = 40 + q - 1, 30 + p - 1
            ky, kx = ys - sh_y + B // 2, xs - sh_x + B // 2
            inb = (ky >= 0) & (ky < B) & (kx >= 0) & (kx < B)
            exp2[inb] += wy[q] * wx[p] * far.astype(np.float32)[ky[inb], kx[inb]]
    assert np.allclose(out2, exp2 * (1 + 4 * 0.35 / 7.0), atol=1e-5)


def _known_page(rng, size=True):
    radius, sigma, spacing = 3.5, 1.0, 12.0
    c = _lattice(8, spacing, 14.0, rng, jitter=0.3)
    h = w = 14 + 7 * 12 + 14
    R = O.window_half(spacing)
    T = O.disk_templates(radius, sigma, R)
    a = rng.uniform(900, 1100, len(c))
    delta = rng.uniform(-0.2, 0.2, len(c)) if size else np.zeros(len(c))
    bg = 100.0
    page = O.render_page(c, a, bg, delta, T, h, w, far=True, cd0=2 * radius)
    return dict(c=c, a=a, delta=delta, bg=bg, page=page, T=T, R=R, cd0=2 * radius, h=h, w=w)


def test_hole_pass_keeps_the_truth_of_a_noise_free_page():
    rng = np.random.default_rng(5)
    K = _known_page(rng)
    n = len(K["c"])
    res = O.fit_holes(K["page"].astype(np.float32), K["c"], K["a"], np.full(n, K["bg"]), K["delta"], K["T"],
                      R=K["R"], noise=1.0, sweeps=1, size=True, cd0=K["cd0"])
    assert np.abs(res["c"] - K["c"]).max() < 1e-4
    assert np.abs(res["delta"] - K["delta"]).max() < 1e-4
    assert np.abs(res["a"] / K["a"] - 1).max() < 1e-5


@pytest.mark.parametrize("size", [False, True])
def test_hole_pass_recovers_perturbed_holes(size):
    rng = np.random.default_rng(6)
    K = _known_page(rng, size=size)
    n = len(K["c"])
    c0 = K["c"] + rng.uniform(-0.15, 0.15, K["c"].shape)
    a0 = K["a"] * rng.uniform(0.93, 1.07, n)
    lev0 = K["bg"] + rng.uniform(-5, 5, n)
    res = O.fit_holes(K["page"].astype(np.float32), c0, a0, lev0, np.zeros(n), K["T"], R=K["R"], noise=1.0,
                      sweeps=6, size=size, cd0=K["cd0"])
    err = np.abs(res["c"] - K["c"]).max()
    assert err < 2e-3, err
    assert np.abs(res["a"] / K["a"] - 1).max() < 5e-3
    if size:
        assert np.abs(res["delta"] - K["delta"]).max() < 5e-3
    moves = [entry["move99"] for entry in res["log"]]
    assert moves[-1] < moves[0]


@pytest.mark.parametrize("d0", [0.0, 0.3])
def test_normal_equations_use_the_models_jacobian(d0):
    """hole_normal's sums equal those of central differences of the hole's own model
    a (T0 + d T1 + d^2 T2)(p - c) + level in (x, y, a, level, d), with unit weights. The
    centre columns are -a Tx, -a Ty (the derivative of T0 alone, as the reference's
    hole_normal): exact at d = 0, so at d = 0.3 only the (a, level, d) block is compared."""
    radius, sigma, R = 3.0, 1.0, 7
    T = O.disk_templates(radius, sigma, R)
    g5, U, N = T["g5"], T["U"], T["N"]
    h, w = 40, 40
    rng = np.random.default_rng(7)
    image = rng.normal(50, 5, (h, w)).astype(np.float32)
    page = np.zeros((h, w))
    c = np.array([[19.3, 20.6]])
    a, level, delta = np.array([800.0]), np.array([40.0]), np.array([d0])
    off = np.arange(-R, R + 1)
    px, py = np.meshgrid(19 + off, 21 + off)

    def model(p):
        x, y, aa, lev, de = p
        v = O.sample(g5, (px - x) * U + N // 2, (py - y) * U + N // 2)
        return aa * (v[..., 0] + de * (v[..., 1] + de * v[..., 4])) + lev

    p0 = np.array([c[0, 0], c[0, 1], a[0], level[0], delta[0]])
    # the data: the same hole moved by (0.3, -0.2) px and 5 % brighter, plus noise, so that the
    # right-hand side carries every column's sign
    image[py, px] = (model(p0 + np.array([0.3, -0.2, 40.0, 0.0, 0.0])) + image[py, px] - 50.0).astype(np.float32)
    acc = O.hole_normal(image, page, g5, np.array([0]), c, a, level, delta, U, R, 1.0, 0.0, 0.0, True)[0]
    steps = np.array([1e-4, 1e-4, 1e-3, 1e-3, 1e-4])
    J = []
    for k in range(5):
        e = np.zeros(5)
        e[k] = steps[k]
        J.append(((model(p0 + e) - model(p0 - e)) / (2 * steps[k])).ravel())
    J = np.array(J)
    resid = (image[py, px] - (page[py, px] + level[0])).ravel()
    normal = J @ J.T
    cols = range(5) if d0 == 0.0 else range(2, 5)
    q = 0
    for r in range(5):
        for s in range(r, 5):
            if r in cols and s in cols:
                scale = math.sqrt(normal[r, r] * normal[s, s])
                assert abs(acc[q] - normal[r, s]) <= 2e-3 * scale, (r, s, acc[q], normal[r, s])
            q += 1
        if r in cols:
            assert abs(acc[15 + r] - J[r] @ resid) <= 2e-3 * math.sqrt(normal[r, r] * (resid @ resid))
