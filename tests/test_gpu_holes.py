"""Hole stage on the GPU (spx_hole_templates / spx_hole_step / spx_hole_render / spx_far_splat)
against the FP64 oracle (oracle/holes.py) on the same inputs.

Tolerances: the device samples FP32 templates with FP32 Keys weights and sums J^T W J in FP32
per lane before an FP64 group reduction (the reference also samples FP32 templates and keeps
FP32 row sums); the oracle is FP64 throughout on the same FP32 template values. Centres and
edge offsets agree to 2e-5 px, brightness to 1e-5 relative, levels to 1e-3 grey levels after
three sweeps."""
import math

import numpy as np
import pytest
import torch

from oracle import holes as O

pytestmark = pytest.mark.gpu


def _lattice(n_side, spacing, margin, rng, jitter=0.3):
    g = margin + spacing * np.arange(n_side)
    X, Y = np.meshgrid(g, g)
    c = np.stack([X.ravel(), Y.ravel()], axis=1).astype(np.float64)
    return c + rng.uniform(-jitter, jitter, c.shape)


def _page(rng, n_side=10, spacing=12.0, radius=3.5, sigma=1.0, noise=0.0, size=True):
    c = _lattice(n_side, spacing, 14.0, rng)
    h = w = int(28 + spacing * (n_side - 1))
    R = O.window_half(spacing)
    T = O.disk_templates(radius, sigma, R)
    a = rng.uniform(900, 1100, len(c))
    delta = rng.uniform(-0.2, 0.2, len(c)) if size else np.zeros(len(c))
    page = O.render_page(c, a, 100.0, delta, T, h, w, far=True, cd0=2 * radius)
    if noise:
        page = page + rng.normal(0, noise, page.shape)
    return dict(c=c, a=a, delta=delta, bg=100.0, image=page.astype(np.float32), T=T, R=R, radius=radius,
                sigma=sigma, h=h, w=w)


def _device_templates(pg, K):
    N = pg.tN
    tm = pg.tmpl.cpu().numpy().reshape(N, N, 8).astype(np.float64)
    T = dict(K["T"])
    T["g5"] = np.ascontiguousarray(tm[..., [0, 1, 3, 4, 2]])
    T["g3"] = np.ascontiguousarray(tm[..., [0, 1, 2]])
    T["far"] = pg.far["kernel"].cpu().numpy().astype(np.float64)
    return T


def test_templates_match_oracle():
    from imageprocessor_new_amd import _lib, holes
    from imageprocessor_new_amd.backends import gpu_device, stream_handle

    for radius, sigma, R in ((3.5, 1.0, 9), (2.4, 1.3, 7), (10.0, 1.6, 24)):
        assert holes.near_reach(radius, sigma, R) == O.near_reach(radius, sigma, R)
        N, r0, Rr, rl = holes.near_reach(radius, sigma, R)
        T = O.disk_templates(radius, sigma, R)
        with gpu_device(0) as dev:
            tm = torch.empty(N * N * 8, dtype=torch.float32, device=dev)
            far = torch.empty(128 * 128, dtype=torch.float32, device=dev)
            _lib.call("spx_hole_templates", radius, sigma, N, 8, r0, 2.0, 128, tm.data_ptr(), far.data_ptr(),
                      stream_handle())
            tm = tm.cpu().numpy().reshape(N, N, 8)
            far = far.cpu().numpy().reshape(128, 128)
        ref = T["g5"][..., [0, 1, 4, 2, 3]]
        assert np.abs(tm[..., :5] - ref).max() < 2e-7 * max(1.0, np.abs(ref).max())
        assert np.all(tm[..., 5:] == 0.0)
        assert np.abs(far - T["far"]).max() < 1e-7


def test_colours_and_spacing_match_oracle():
    from imageprocessor_new_amd import holes

    rng = np.random.default_rng(11)
    K = _page(rng, n_side=12)
    pg = holes.HolePage(K["image"], K["c"], noise=1.0)
    sp = float(np.median(O.nearest_distance(K["c"])))
    assert abs(pg.spacing - sp) < 1e-6
    assert pg.R == O.window_half(sp)
    mine = [ix.cpu().numpy() for ix in pg.colours()]
    ref = O.colours(K["c"], K["h"], K["w"], pg.spacing)
    assert len(mine) == len(ref)
    for x, y in zip(mine, ref):
        assert np.array_equal(np.sort(x), y)


@pytest.mark.parametrize("size,far", [(False, False), (True, True)])
def test_hole_pass_matches_oracle(size, far):
    from imageprocessor_new_amd import holes

    rng = np.random.default_rng(12 + int(size))
    K = _page(rng, noise=3.0, size=size)
    n = len(K["c"])
    c0 = K["c"] + rng.uniform(-0.1, 0.1, K["c"].shape)
    a0 = K["a"] * rng.uniform(0.95, 1.05, n)
    lev0 = K["bg"] + rng.uniform(-3, 3, n)
    pg = holes.HolePage(K["image"], c0, noise=3.0, brightness=a0, level=lev0, window_px=K["R"])
    pg.set_template(K["radius"], K["sigma"])
    T = _device_templates(pg, K)
    chunks = [ix.cpu().numpy().astype(np.int64) for ix in pg.colours()]
    sweeps = 3
    keep = {}
    log = pg.fit_holes(sweeps=sweeps, size=size, far=far, keep=keep)
    ref = O.fit_holes(K["image"], c0, a0, lev0, np.zeros(n), T, R=K["R"], noise=3.0, sweeps=sweeps, size=size,
                      far=far, cd0=2 * K["radius"], chunks=chunks)
    res = pg.result(log)
    print(f"size={size} far={far}: max |dc| {np.abs(res.centers_xy - ref['c']).max():.2e} px, "
          f"|da/a| {np.abs(res.brightness / ref['a'] - 1).max():.2e}, |dlev| {np.abs(res.level - ref['level']).max():.2e}")
    assert np.abs(res.centers_xy - ref["c"]).max() < 2e-5
    assert np.abs(res.brightness / ref["a"] - 1).max() < 1e-5
    assert np.abs(res.level - ref["level"]).max() < 1e-3
    assert np.abs(res.edge_offset_px - ref["delta"]).max() < 2e-5
    for a, b in zip(log, ref["log"]):
        assert abs(a["rms"] - b["rms"]) < 1e-4 * b["rms"]
    light = keep["light"].cpu().numpy()
    assert np.abs(light - ref["page"]).max() < 1e-3
    # every hole's leftover against the oracle's, in hole order
    order = np.concatenate(chunks)
    ref_rms = np.empty(n)
    ref_rms[order] = ref["stat"][:, 2]
    assert np.allclose(res.hole_rms, ref_rms, rtol=1e-4)


def test_hole_pass_recovers_noise_free_holes():
    from imageprocessor_new_amd import holes

    rng = np.random.default_rng(14)
    K = _page(rng, size=True)
    n = len(K["c"])
    c0 = K["c"] + rng.uniform(-0.15, 0.15, K["c"].shape)
    res = holes.fit_holes(K["image"], c0, radius=K["radius"], sigma=K["sigma"], noise=1.0,
                          brightness=K["a"] * rng.uniform(0.93, 1.07, n), level=K["bg"] + rng.uniform(-5, 5, n),
                          sweeps=6, size=True, window_px=K["R"])
    assert np.abs(res.centers_xy - K["c"]).max() < 2e-3
    assert np.abs(res.edge_offset_px - K["delta"]).max() < 5e-3
    assert res.log[-1]["move99"] < res.log[0]["move99"]


def test_bad_arguments_raise():
    from imageprocessor_new_amd import _lib, holes

    with pytest.raises(ValueError):
        holes.HolePage(np.zeros((10, 10), np.float32), np.zeros((1, 2)), noise=1.0)
    with pytest.raises(ValueError):
        holes.HolePage(np.zeros((10, 10), np.float32), np.zeros((3, 2)), noise=0.0)
    with pytest.raises(ValueError):
        _lib.call("spx_hole_templates", -1.0, 1.0, 256, 8, 8.0, 2.0, 128, 1, None, None)
    opt = _lib.SpxHoleStepOptions(nv0=0.0)
    import ctypes
    with pytest.raises(ValueError):
        _lib.call("spx_hole_step", 1, 10, 10, 10, 1, 1, 256, 8, 3, 1, 1, 1, 1, 1, 1, 1, ctypes.byref(opt), 1, None)


def test_select_kth_matches_sorting():
    import ctypes

    from imageprocessor_new_amd import _lib
    from imageprocessor_new_amd.backends import gpu_device, stream_handle

    rng = np.random.default_rng(21)
    x = rng.normal(size=(5001, 4))
    x[::7, 1] = 0.5
    x[::3, 2] = -0.0
    x[1::3, 2] = 0.0
    x[:, 3] = np.abs(x[:, 3]) * 1e-300
    with gpu_device(0) as dev:
        xt = torch.from_numpy(x).to(dev)
        out = torch.empty(8, dtype=torch.float64, device=dev)
        for col in range(4):
            ks = [0, 2500, 5000, 1234, 4950, 17, 2499, 3]
            _lib.call("spx_select_kth", int(xt.data_ptr()), 5001, 4, (ctypes.c_int64 * 8)(*([col] * 8)),
                      (ctypes.c_int64 * 8)(*ks), 8, out.data_ptr(), stream_handle())
            got = out.cpu().numpy()
            want = np.sort(x[:, col])[ks]
            assert np.array_equal(got, want), (col, got, want)
    pg_stats = None
    from imageprocessor_new_amd import holes

    pg = holes.HolePage.__new__(holes.HolePage)
    pg.n = 5001
    with gpu_device(0) as dev:
        st = torch.from_numpy(np.abs(x)).to(dev)
        pg_stats = pg._sweep_stats(st)
    assert pg_stats[0] == np.median(np.abs(x[:, 2]))
    assert pg_stats[1] == np.percentile(np.abs(x[:, 0]), 99)
    assert pg_stats[2] == np.percentile(np.abs(x[:, 1]), 99)


def test_render_covers_every_offset_of_the_reach():
    """One hole's light over every offset inside its reach (ox^2 + oy^2 <= rl2, the reach's edge
    rows included) against the oracle's render on the same FP32 template values: the same
    offsets are lit, and the cached light agrees."""
    from imageprocessor_new_amd import _lib, holes
    from imageprocessor_new_amd.backends import gpu_device, stream_handle

    h = w = 64
    for x, y in ((31.49, 30.51), (31.51, 30.49), (32.0, 32.0), (31.3, 31.7)):
        c = np.array([[x, y], [5.0, 5.0]])
        pg = holes.HolePage(np.zeros((h, w), np.float32), c, noise=1.0, brightness=[1000.0, 1000.0])
        pg.set_template(2.0, 2.5)
        N, U, Rr = pg.tN, pg.tU, pg.Rr
        rl2 = float((pg.r_light + 0.75) ** 2)
        P = (2 * Rr + 1) ** 2
        with gpu_device(0) as dev:
            page = torch.zeros(h * w, dtype=torch.float64, device=dev)
            cache = torch.zeros(P, dtype=torch.float32, device=dev)
            idx = torch.zeros(1, dtype=torch.int32, device=dev)
            c_old = pg.c[:1].clone()
            _lib.call("spx_hole_render", int(page.data_ptr()), h, w, int(cache.data_ptr()), int(pg.tmpl.data_ptr()), N,
                      U, Rr, rl2, int(idx.data_ptr()), 1, int(c_old.data_ptr()), int(pg.c.data_ptr()),
                      int(pg.a.data_ptr()), int(pg.delta.data_ptr()), stream_handle())
            got_cache = cache.cpu().numpy()
            got_page = page.cpu().numpy().reshape(h, w)
        g3 = pg.tmpl.cpu().numpy().reshape(N, N, 8).astype(np.float64)[..., [0, 1, 2]]
        ref_page = np.zeros((h, w))
        ref_cache = np.zeros((1, P), dtype=np.float32)
        O.render_scatter(ref_page, ref_cache, g3, np.array([0]), c[:1].copy(), c.copy(), np.array([1000.0, 1000.0]),
                         np.zeros(2), U, Rr, rl2)
        ref_cache = ref_cache[0]
        assert np.array_equal(got_cache != 0, ref_cache != 0), (x, y)
        assert np.all(np.abs(got_cache - ref_cache) <= 1e-6 * np.abs(ref_cache).max() + 1e-3 * np.abs(ref_cache))
        assert np.abs(got_page - ref_page).max() <= 1e-6 * np.abs(ref_page).max()


def test_hole_pass_is_bit_reproducible():
    """The page model and the far light are summed as fixed-point integers, so a second run of
    the same sweeps gives the same page, centres, brightnesses and sweep log bit for bit (holes
    packed densely enough that their render windows overlap)."""
    from imageprocessor_new_amd import holes

    rng = np.random.default_rng(21)
    K = _page(rng, n_side=14, spacing=8.5, radius=2.4, sigma=1.2, noise=3.0)
    n = len(K["c"])
    c0 = K["c"] + rng.uniform(-0.1, 0.1, K["c"].shape)
    runs = []
    for _ in range(2):
        pg = holes.HolePage(K["image"], c0, noise=3.0, brightness=K["a"], level=np.full(n, K["bg"]))
        pg.set_template(K["radius"], K["sigma"])
        keep = {}
        log = pg.fit_holes(sweeps=4, size=True, far=True, keep=keep)
        res = pg.result(log)
        runs.append((keep["light"].cpu().numpy(), res, log))
    (l0, r0, g0), (l1, r1, g1) = runs
    assert np.array_equal(l0, l1)
    for f in ("centers_xy", "brightness", "level", "edge_offset_px"):
        assert np.array_equal(getattr(r0, f), getattr(r1, f)), f
    assert [e["rms"] for e in g0] == [e["rms"] for e in g1]


def test_outline_fit_is_bit_reproducible():
    """Two fit_outline runs on one page: the same rounds, outline and per-hole results bit for bit
    (every device sum of the fit is order-independent)."""
    from imageprocessor_new_amd import outline as G
    from imageprocessor_new_amd.backends import gpu_device

    sp, nx, ny = 8.5, 16, 12
    rng = np.random.default_rng(3)
    X, Y = np.meshgrid(10.0 + sp * np.arange(nx), 10.0 + sp * np.arange(ny))
    truth = np.stack([X.ravel(), Y.ravel()], 1) + rng.normal(0, 0.05, (nx * ny, 2))
    with gpu_device(0):
        t = G.OutlineTemplate()
        theta = t.pack(2.0, 1.9, 2.4, 0.1, [[0.02, -0.01], [0.01, 0.0]],
                       t.blur_pack(s1x=1.1, s1y=1.05, s2x=2.2, s2y=2.2, w2=0.08))
        page = G.render_outline_page(theta, truth, 3000.0 * rng.uniform(0.95, 1.05, len(truth)),
                                     (int(20 + sp * ny), int(20 + sp * nx)), level=500.0, noise=30.0,
                                     edge_offset=rng.normal(0, 0.05, len(truth)), seed=3).cpu().numpy()
    seeds = truth + rng.uniform(-0.3, 0.3, truth.shape)
    a = G.fit_outline(page, seeds, rounds=4)
    b = G.fit_outline(page, seeds, rounds=4)
    assert a.rounds == b.rounds and a.converged == b.converged
    assert np.array_equal(a.theta, b.theta)
    for f in ("centers_xy", "cd_px", "brightness", "level", "edge_offset_px"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
