"""GPU parity of the deblending behind detect_centers(background="model") against the CPU
oracle (oracle/deblend.py), through the C-ABI:

* each kernel on the oracle's own inputs: the first fit's data and start (the border percentile
  exactly), the Gaussian LM fits (FP64; LU against LAPACK rounding: 1e-7 px), the fixed-point
  render (2^-24 counts per term), the tile medians (exact), the tile map (1e-12), the cleaned
  cells, the batch refinements of cleaned cells (1e-9 px);
* a whole band against the oracle's deblend on the reference's noise-free lattices (same
  iteration count, parameters to 1e-6), and the reference's own known answers on the GPU path
  (background within 1 % of the amplitude, amplitudes within 1 %, centres within 0.01 px);
* detect_centers(background="model") against the oracle's pipeline and the reference's known
  answer (median centre error below 0.005 px), with the Voronoi estimators too, and in bands.
"""

import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree

from imageprocessor_new_amd import _lib
from imageprocessor_new_amd.backends import gpu_device, stream_handle
from imageprocessor_new_amd.centers import detect_centers
from imageprocessor_new_amd.deblend import deblend_band, sigma_major
from oracle import deblend as db
from oracle import voronoi_refine as vr

pytestmark = pytest.mark.gpu

FWHM, BACKGROUND, AMPLITUDE = 4.0, 20.0, 50.0


def lattice(ratio, shape=(120, 240)):
    pitch = ratio * FWHM
    sigma = FWHM / (2 * np.sqrt(2 * np.log(2)))
    xs = np.arange(15, shape[1] - 15, pitch)
    ys = np.arange(15, shape[0] - 15, pitch)
    xx, yy = np.meshgrid(xs, ys)
    xx[1::2] += pitch / 2
    c = np.column_stack((xx.ravel(), yy.ravel()))
    c = c[c[:, 0] < shape[1] - 15]
    y, x = np.indices(shape, dtype=float)
    img = np.full(shape, BACKGROUND)
    for cx, cy in c:
        img += AMPLITUDE * np.exp(-0.5 * ((x - cx) ** 2 + (y - cy) ** 2) / sigma ** 2)
    return img, c


def _cells(img, centers, ratio):
    labels = vr.voronoi_labels(centers, img.shape)
    half = int(np.ceil(0.6 * ratio * FWHM))
    rows = np.array([[x - half, y - half, 2 * half + 1, 2 * half + 1, x, y] for x, y in centers])
    _, masks, origins = vr.extract_rois(img, labels, rows, pad=3)
    return masks, np.asarray(origins, dtype=np.int64)


def _t(a, dev, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a).to(dev)


def p(t):
    return None if t is None else int(t.data_ptr())


def test_kernels_on_oracle_inputs():
    img, centers = lattice(2.0)
    img = img + np.random.default_rng(0).normal(0, 1.0, img.shape)
    masks, origins = _cells(img, centers, 2.0)
    seeds = centers + 0.2
    N, Hm, Wm = masks.shape
    H, W = img.shape
    with gpu_device(0) as dev:
        img_t, m_t, o_t = _t(img, dev), _t(masks, dev, np.uint8), _t(origins, dev, np.int32)
        # first data and start
        data = torch.empty((N, Hm, Wm), dtype=torch.float64, device=dev)
        p0 = torch.empty((N, 6), dtype=torch.float64, device=dev)
        # every buffer a call reads is held by a name until the call's kernels have run: a temporary
        # freed while its pointer is in flight can be handed to the next allocation
        seeds_t = _t(seeds, dev)
        _lib.call("spx_deblend_init", p(img_t), H, W, p(m_t), p(o_t), p(seeds_t), N, Hm, Wm, p(data), p(p0),
                  stream_handle())
        raw = db.windows(img, origins, (Hm, Wm))
        from scipy.ndimage import binary_erosion

        border = masks & ~binary_erosion(masks, structure=np.ones((1, 3, 3)), border_value=0)
        bg0 = np.array([np.percentile(r[b], 10) if b.any() else 0.0 for r, b in zip(raw, border)])
        want_data = np.where(masks, raw - bg0[:, None, None], 0.0)
        want_p0 = db.initial_params(want_data, masks, seeds - origins)
        assert np.array_equal(data.cpu().numpy(), want_data)
        np.testing.assert_allclose(p0.cpu().numpy(), want_p0, rtol=1e-12, atol=1e-12)
        # fits
        out = torch.empty((N, 6), dtype=torch.float64, device=dev)
        init_t = _t(want_p0, dev)
        _lib.call("spx_gauss_fit", p(data), p(m_t), N, Hm, Wm, p(init_t), 30, p(out), stream_handle())
        got = out.cpu().numpy()
        want = db.fit_gaussians(want_data, masks, want_p0)
        print(f"{N} fits: max |dcentre| {np.abs(got[:, 1:3] - want[:, 1:3]).max():.2e}", flush=True)
        np.testing.assert_allclose(got[:, 1:3], want[:, 1:3], atol=1e-7)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-7)
        # render
        pg = want.copy()
        pg[:, 1:3] += origins
        rad = np.clip(np.ceil(4.0 * db.sigma_major(pg)), 2, max(Hm, Wm))
        model = torch.empty((H, W), dtype=torch.float64, device=dev)
        pg_t, rad_t = _t(pg, dev), _t(rad, dev)
        _lib.call("spx_gauss_render", p(pg_t), p(rad_t), N, H, W, p(model), stream_handle())
        want_model = db.render(pg, (H, W), max(Hm, Wm))
        assert np.abs(model.cpu().numpy() - want_model).max() < 1e-5
        # tile medians (on the oracle's model), the map, cleaned cells
        model_t = _t(want_model, dev)
        dist = np.full((H, W), 5.0)
        dist[::7, ::5] = 0.5
        tile = db.auto_tile(seeds, (H, W))[0]
        ny, nx = -(-H // tile), -(-W // tile)
        med = torch.empty((ny, nx), dtype=torch.float64, device=dev)
        cnt = torch.empty((ny, nx), dtype=torch.int64, device=dev)
        thr = 0.05 * AMPLITUDE
        dist_t = _t(dist, dev)
        _lib.call("spx_tile_medians", p(img_t), p(model_t), p(dist_t), H, W, tile, thr, 1.0, p(med), p(cnt),
                  stream_handle())
        free = (want_model < thr) & (dist >= 1.0)
        wm, wc = db.tile_medians(img - want_model, free, tile)
        assert np.array_equal(cnt.cpu().numpy(), wc)
        assert np.array_equal(med.cpu().numpy(), wm, equal_nan=True)
        fill = np.where(np.isnan(wm), np.nanmean(wm), wm)
        yc, xc = db.tile_centers(tile, (H, W))
        bmap = torch.empty((H, W), dtype=torch.float64, device=dev)
        fill_t, yc_t, xc_t = _t(fill, dev), _t(yc, dev), _t(xc, dev)
        _lib.call("spx_tile_map", p(fill_t), p(yc_t), p(xc_t), ny, nx, H, W, p(bmap), stream_handle())
        want_map = db.upsample_tiles(fill, tile, (H, W))
        assert np.abs(bmap.cpu().numpy() - want_map).max() < 1e-12
        cl = torch.empty((N, Hm, Wm), dtype=torch.float64, device=dev)
        _lib.call("spx_deblend_clean", p(img_t), p(bmap), p(model_t), p(pg_t), p(rad_t), p(m_t), p(o_t),
                  N, Hm, Wm, H, W, p(cl), stream_handle())
        want_cl = db.clean_cells(raw, want_map, want_model, pg, masks, origins, max(Hm, Wm))
        assert np.abs(cl.cpu().numpy() - want_cl).max() < 1e-9
        # refinement of cleaned cells on their bright core
        peak = np.where(np.isfinite(want[:, 0]) & (want[:, 0] > 0), want[:, 0],
                        np.where(masks, want_cl, -np.inf).reshape(N, -1).max(axis=1))
        support = masks & (want_cl > 0.2 * peak[:, None, None])
        # the log-quadratic's normal equations in ROI coordinates carry a condition number near 1e8:
        # LU and LAPACK agree to ~1e-8 relative there
        for method, ref, tol in ((0, db.weighted_batch, 1e-9), (1, db.logquad_batch, 1e-6)):
            c = torch.empty((N, 2), dtype=torch.float64, device=dev)
            cl_t, amp_t = _t(want_cl, dev), _t(want[:, 0], dev)
            _lib.call("spx_cell_refine", p(cl_t), p(m_t), N, Hm, Wm, method, p(amp_t), 0.2, p(c), stream_handle())
            w = ref(want_cl, support)
            g = c.cpu().numpy()
            assert np.array_equal(np.isfinite(g), np.isfinite(w))
            fin = np.isfinite(w).all(axis=1)
            assert fin.mean() > 0.9
            assert np.abs(g[fin] - w[fin]).max() < tol, method


@pytest.mark.parametrize("ratio", [2.67, 2.0, 1.6, 1.4])
def test_band_matches_oracle_and_known_answer(ratio):
    img, centers = lattice(ratio)
    seeds = centers + 0.2
    with gpu_device(0) as dev:
        img16 = torch.zeros(img.shape, dtype=torch.uint16, device=dev)  # only the cells' geometry reads it
        cells, st = deblend_band(img16, _t(img, dev), seeds)
        masks = cells.masks_host
        origins = cells.origins_host
        want = db.deblend(img, seeds, masks, origins)
    print(f"ratio {ratio}: {len(seeds)} holes, iterations {st['iterations']} (oracle {want['iterations']}), "
          f"source {st['background_info']['source']}", flush=True)
    assert st["iterations"] == want["iterations"]
    np.testing.assert_allclose(st["params"], want["params"], rtol=1e-6, atol=1e-6)
    assert np.array_equal(st["status"], want["status"])
    assert st["background_info"]["source"] in ("measured", "interpolated")
    assert (st["status"] == 0).all()
    np.testing.assert_allclose(st["params"][:, 0], AMPLITUDE, rtol=0.01)
    np.testing.assert_allclose(st["params"][:, 1:3] + origins, centers, atol=0.01)


def test_detect_centers_model_background():
    img, centers = lattice(2.0)
    img16 = np.clip(np.rint(img * 50), 0, 65535).astype(np.uint16)
    kw = dict(backend="gpu", threshold="otsu", area_max=200, logquad_variant="cpu")
    plain = detect_centers(img16, refine="logquad", **kw)
    model = detect_centers(img16, refine="logquad", background="model", **kw)
    assert "background" not in plain.meta and model.meta["background"] == "model"
    err = cKDTree(centers).query(model.centers_xy)[0]
    print(f"model background: {len(model.centers_xy)} centres, median error {np.median(err):.2e} px, "
          f"iterations {model.meta['deblend_iterations_used']}, {model.meta['deblend_status_counts']}", flush=True)
    assert len(model.centers_xy) == len(plain.centers_xy)
    assert np.median(err) < 0.005
    want, _, _ = db.deblended_centers(img16.astype(np.float64), plain.centers_xy, "logquad")
    fin = np.isfinite(want).all(axis=1)
    assert np.abs(model.centers_xy - want[fin]).max() < 1e-6
    with pytest.raises(NotImplementedError, match="edge methods"):
        detect_centers(img16, refine="edge_gradmoment", background="model", **kw)


@pytest.mark.parametrize("refine", ["weighted", "radial_symmetry"])
def test_detect_centers_model_background_other_methods_and_bands(refine):
    img, centers = lattice(2.67, shape=(200, 240))
    img16 = np.clip(np.rint(img * 50), 0, 65535).astype(np.uint16)
    kw = dict(backend="gpu", threshold="otsu", area_max=400, logquad_variant="cpu")
    first = detect_centers(img16, refine=refine, **kw)
    model = detect_centers(img16, refine=refine, background="model", **kw)
    want, quality, _ = db.deblended_centers(img16.astype(np.float64), first.centers_xy, refine)
    fin = np.isfinite(want).all(axis=1)
    print(f"{refine}: {len(model.centers_xy)} centres (oracle {fin.sum()})", flush=True)
    assert len(model.centers_xy) == fin.sum()
    assert np.abs(model.centers_xy - want[fin]).max() < (1e-5 if refine == "radial_symmetry" else 1e-8)
    if refine == "radial_symmetry":
        np.testing.assert_allclose(model.meta["radial_symmetry_residual"], quality[fin], rtol=1e-5, atol=1e-9)
    banded = detect_centers(img16, refine=refine, background="model", tile_h=100, overlap=40, **kw)
    d = cKDTree(centers).query(banded.centers_xy)[0]
    assert len(banded.centers_xy) >= len(model.centers_xy) - 2 and np.median(d) < 0.05
