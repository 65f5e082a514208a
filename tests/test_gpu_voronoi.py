"""GPU parity of the Voronoi-cell refinements (refine="radial_symmetry" / "isophote_curvature")
against the CPU oracle (oracle/voronoi_refine.py), through the C-ABI:

* nearest-seed labels and cell areas: bit-exact (few seeds, many seeds through the ring search,
  planted ties, seeds outside the image);
* cell ROIs, masks and origins: bit-exact (the fill is the float64 percentile cast to float32);
* the estimators on the oracle's ROI stacks: the kernels keep the reference's float32 / float64
  split and operation order, so only the order of the float64 sums differs: centres within
  1e-5 px, quality within 1e-6 relative; ROI stacks too large for LDS take the global-scratch
  path;
* detect_centers end to end against the oracle's pipeline (same centres to 1e-5 px, same meta),
  the strip path, and the reference's batch functions.
"""

import numpy as np
import pytest
import torch

from imageprocessor_new_amd import _lib, synthetic
from imageprocessor_new_amd.backends import gpu_device, stream_handle
from imageprocessor_new_amd.centers import detect_centers
from imageprocessor_new_amd.voronoi import (_isophote_curvature_batch_gpu, _radial_symmetry_batch_gpu,
                                            compute_voronoi_labels_gpu, voronoi_labels_device)
from oracle import voronoi_refine as vr

pytestmark = pytest.mark.gpu

TOL_C = 1e-5


@pytest.mark.parametrize("n,shape,seed", [(7, (90, 130), 0), (300, (400, 520), 1), (2000, (600, 700), 2)])
def test_labels_and_areas_match_oracle(n, shape, seed):
    rng = np.random.default_rng(seed)
    seeds = rng.random((n, 2)) * [shape[1] + 20, shape[0] + 20] - 10  # some outside the image
    seeds[: n // 3] = np.round(seeds[: n // 3])  # integer seeds: exact ties on midlines
    got = compute_voronoi_labels_gpu(seeds, shape)
    want = vr.voronoi_labels(seeds, shape)
    print(f"{n} seeds on {shape}: {np.mean(got == want):.6f} equal", flush=True)
    assert np.array_equal(got, want)
    with gpu_device(0) as dev:
        _, areas = voronoi_labels_device(torch.from_numpy(seeds).to(dev), shape, labels=False)
        assert np.array_equal(areas.cpu().numpy(), np.bincount(want.ravel(), minlength=n))


def _page(seed, H=256, W=320):
    return synthetic.make_workload("square2048", seed=seed, H=H, W=W)


def _rows(img):
    from oracle import segment

    t = segment.otsu_u16(img)
    num, _, stats, cents = segment.connected_components_fast(segment.binary_mask(img, t, False, 0, 0), 8)
    keep = [k for k in range(1, num) if 20 <= stats[k, 4] <= 2000]
    rows = np.concatenate([stats[keep, :4].astype(np.float64), np.asarray(cents)[keep]], axis=1)
    return vr.voronoi_rows(rows, img.shape)


def test_rois_match_oracle():
    arr = _page(3)
    rows = _rows(arr.image)
    labels = vr.voronoi_labels(rows[:, 4:6], arr.image.shape)
    want_r, want_m, want_o = vr.extract_rois(arr.image, labels, rows, pad=3)
    K, Hm, Wm = want_r.shape
    H, W = arr.image.shape
    with gpu_device(0) as dev:
        img_t = torch.from_numpy(arr.image).to(dev)
        lab_t = torch.from_numpy(labels).to(dev)
        rows_t = torch.from_numpy(rows).to(dev)
        rois = torch.empty((K, Hm, Wm), dtype=torch.float32, device=dev)
        masks = torch.empty((K, Hm, Wm), dtype=torch.uint8, device=dev)
        org = torch.empty((K, 2), dtype=torch.int32, device=dev)
        p = lambda t: int(t.data_ptr())
        _lib.call("spx_voronoi_rois", p(img_t), H, W, W, p(lab_t), p(rows_t), K, 3, Hm, Wm, p(rois), p(masks),
                  p(org), stream_handle())
        torch.cuda.synchronize()
        print(f"{K} ROIs of {Hm} x {Wm}", flush=True)
        assert np.array_equal(rois.cpu().numpy(), want_r.astype(np.float32))
        assert np.array_equal(masks.cpu().numpy().astype(bool), want_m)
        assert np.array_equal(org.cpu().numpy(), np.asarray(want_o, dtype=np.int32))


def _stack(seed, pad=3, H=256, W=320):
    arr = _page(seed, H, W)
    rows = _rows(arr.image)
    labels = vr.voronoi_labels(rows[:, 4:6], arr.image.shape)
    return vr.extract_rois(arr.image, labels, rows, pad=pad)


@pytest.mark.parametrize("pad", [3, 9])  # 9: the isophote stack leaves LDS for global scratch
def test_estimators_match_oracle(pad):
    rois, masks, _ = _stack(4, pad=pad)
    for name, gpu, ref in (("radial", _radial_symmetry_batch_gpu, vr.radial_symmetry),
                           ("isophote", _isophote_curvature_batch_gpu, vr.isophote_curvature)):
        gc, gq = gpu(rois, masks)
        wc, wq = ref(rois, masks)
        fin = np.isfinite(wc).all(axis=1)
        print(f"{name} pad {pad}: {rois.shape}, finite {fin.sum()}, max |dc| "
              f"{np.nanmax(np.abs(gc - wc)):.2e}", flush=True)
        assert np.array_equal(np.isfinite(gc).all(axis=1), fin)
        assert np.abs(gc[fin] - wc[fin]).max() <= TOL_C
        np.testing.assert_allclose(gq[fin], wq[fin], rtol=1e-6, atol=1e-9)


def test_estimator_factors_and_margins():
    rois, masks, _ = _stack(5)
    for f, m in ((1, 0), (2, 1), (3, 2)):
        gc, gq = _radial_symmetry_batch_gpu(rois, masks, upsample_factor=f, boundary_margin=m)
        wc, wq = vr.radial_symmetry(rois, masks, upsample_factor=f, boundary_margin=m)
        fin = np.isfinite(wc).all(axis=1)
        assert np.array_equal(np.isfinite(gc).all(axis=1), fin)
        assert np.abs(gc[fin] - wc[fin]).max() <= TOL_C, (f, m)
        gc, _ = _isophote_curvature_batch_gpu(rois, masks, upsample_factor=f, boundary_margin=m, pre_smooth_sigma=0.0)
        wc, _ = vr.isophote_curvature(rois, masks, upsample_factor=f, boundary_margin=m, pre_smooth_sigma=0.0)
        fin = np.isfinite(wc).all(axis=1)
        assert np.abs(gc[fin] - wc[fin]).max() <= TOL_C, (f, m)
    gc, gr = _radial_symmetry_batch_gpu(rois, masks, residual_max=0.0)
    assert np.isnan(gc).all() and np.isinf(gr).all()


@pytest.mark.parametrize("refine", ["radial_symmetry", "isophote_curvature"])
def test_detect_centers_matches_oracle_pipeline(refine):
    arr = _page(6, 384, 448)
    res = detect_centers(arr.image, backend="gpu", threshold="otsu", refine=refine, area_min=20, area_max=2000)
    want, meta = vr.detect_voronoi(arr.image, refine, area_min=20, area_max=2000)
    key = "radial_symmetry_residual" if refine == "radial_symmetry" else "isophote_curvature_spread"
    print(f"{refine}: {res.centers_xy.shape[0]} centres (oracle {want.shape[0]})", flush=True)
    assert res.centers_xy.shape == want.shape
    assert np.abs(res.centers_xy - want).max() <= TOL_C
    np.testing.assert_allclose(res.meta[key], meta[key], rtol=1e-6, atol=1e-9)
    assert np.array_equal(res.meta["voronoi_cell_area"], meta["voronoi_cell_area"])


def test_strip_path_finds_every_hole():
    arr = _page(7, 512, 384)
    res = detect_centers(arr.image, backend="gpu", threshold="otsu", refine="radial_symmetry", area_min=20,
                         area_max=2000, tile_h=200, overlap=64)
    from scipy.spatial import cKDTree

    t = arr.centers_xy
    inner = t[(t[:, 0] > 14) & (t[:, 0] < 384 - 14) & (t[:, 1] > 14) & (t[:, 1] < 512 - 14)]
    d, _ = cKDTree(res.centers_xy).query(inner)
    print(f"strips: {res.centers_xy.shape[0]} centres, max distance to truth {d.max():.3f}", flush=True)
    assert d.max() < 0.3
