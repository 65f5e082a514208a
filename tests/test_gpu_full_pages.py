"""The uneven batch and the tiled pipeline at full size.

* The 16384 x 16384 staggered array (~200k holes, r U[4, 20], 10 % ellipses, patches 15..63 px,
  Voronoi cells): a seeded oracle sample per (patch class, model) bucket, integer-shift
  equivariance, mirror symmetry and run-to-run identity.
* The 32768 x 32768 mixed page (square / staggered blocks, pitch 24..48 px, ~1.28M holes) through
  detect_centers(tile_h=8192, overlap=128) with the CNN start, <= 4 LM attempts and the
  fallback per strip: every hole found exactly once within 0.5 px, none lost where the strips
  meet, and a seeded sample of each strip's fits against the FP64 oracle's plain fit of the
  same patches.
"""
import numpy as np
import pytest
import torch

from imageprocessor_new_amd import synthetic
from oracle import erfdisk

pytestmark = pytest.mark.gpu


def _buckets(half, model):
    """(patch class of the whole patch, model): the strata of the oracle sample."""
    npl = ((2 * half.astype(np.int64) + 1) ** 2 + 63) // 64
    cls = np.select([npl <= 4, npl <= 8, npl <= 16, npl <= 32], [0, 1, 2, 3], 4)
    return 2 * cls + (model == 1)


@pytest.fixture(scope="module")
def uneven():
    from imageprocessor_new_amd import learned
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.fit import fit_erf_disks_device

    print("rendering the 16384 x 16384 staggered array", flush=True)
    with gpu_device(0) as dev:
        page = synthetic.render_workload_device("hex16384", seed=22)
        f = learned.page_features(page, dev)
        out = fit_erf_disks_device(page.image, f.seeds, f.half, f.model, neighbours=f.nbr)
        torch.cuda.synchronize()
        res = {k: v.cpu().numpy() for k, v in out.items()}
        res["counters"] = _counters()
        host = {k: getattr(f, k).cpu().numpy() for k in ("seeds", "half", "model", "nbr")}
    return page, host, res


def _counters():
    from imageprocessor_new_amd import _lib
    from imageprocessor_new_amd.backends import stream_handle

    c = np.zeros(8, dtype=np.uint64)
    _lib.call("spx_fit_counters", int(c.ctypes.data), 8, stream_handle())
    return c


def test_uneven_page_patches_fit_their_buckets(uneven):
    """Every feature's fitted pixels fit the bucket its pixel bound chose (the overflow counter
    stays 0), and the counted pixel evaluations are those of the compacted patches."""
    _, f, res = uneven
    c = res["counters"]
    assert int(c[5]) == 0
    assert int(c[2]) > 0 and int(c[0]) + int(c[1]) > 0


def _refit(img_t, seeds, half, model):
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.fit import cell_radius, fit_erf_disks_device, seed_neighbours_device

    with gpu_device(0) as dev:
        s = torch.from_numpy(np.ascontiguousarray(seeds, dtype=np.float32)).to(dev)
        h = torch.from_numpy(np.ascontiguousarray(half)).to(dev)
        m = torch.from_numpy(np.ascontiguousarray(model)).to(dev)
        nbr = seed_neighbours_device(s, img_t.shape, cell_radius(int(half.max())))
        out = fit_erf_disks_device(img_t, s, h, m, neighbours=nbr)
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in out.items()}


def test_uneven_page_sample_per_bucket_matches_the_oracle(uneven):
    page, f, res = uneven
    K = f["seeds"].shape[0]
    assert K > 199_000 and (res["status"] <= 1).mean() >= 0.99
    bid = _buckets(f["half"], f["model"])
    rng = np.random.default_rng(5)
    idx = np.sort(np.concatenate([rng.choice(np.nonzero(bid == b)[0], size=min(30, int((bid == b).sum())),
                                             replace=False) for b in np.unique(bid)]))
    img = page.image.cpu().numpy()
    ref, _, _, ref_st = erfdisk.fit_batch(img, f["seeds"], f["half"], f["model"], neighbours=f["nbr"], only=idx)
    st = res["status"][idx]
    assert (st == ref_st[idx]).mean() >= 0.97
    both = idx[(st == 0) & (ref_st[idx] == 0)]
    err = np.abs(res["centers"][both] - ref[both, :2]).max(axis=1)
    for b in np.unique(bid[both]):
        print(f"bucket {b}: {int((bid[both] == b).sum())} fits, max centre error {err[bid[both] == b].max():.2e} px")
    assert err.max() <= 1e-3
    rs = np.abs(res["params"][both, 2:4].astype(np.float64) - ref[both, 2:4]).max()
    assert rs <= 1e-3


def test_uneven_page_integer_shift_gives_the_same_fits(uneven):
    page, f, res = uneven
    dx, dy = 5, 3
    H, W = page.image.shape
    moved = _refit(page.image[dy:, dx:], f["seeds"] - np.float32([dx, dy]), f["half"], f["model"])
    from imageprocessor_new_amd.fit import cell_radius

    reach = cell_radius(int(f["half"].max())) + 2
    x, y = f["seeds"][:, 0], f["seeds"][:, 1]
    keep = (x - reach >= dx) & (y - reach >= dy) & (x + reach <= W - 1) & (y + reach <= H - 1)
    assert keep.mean() > 0.95
    np.testing.assert_array_equal(moved["status"][keep], res["status"][keep])
    np.testing.assert_array_equal(moved["iters"][keep], res["iters"][keep])
    np.testing.assert_array_equal(moved["params"][keep, 2:], res["params"][keep, 2:])
    assert np.abs(moved["centers"][keep] + [dx, dy] - res["centers"][keep]).max() <= 1e-9


def test_uneven_page_mirror_and_second_run(uneven):
    page, f, res = uneven
    W = page.image.shape[1]
    mirrored = _refit(torch.flip(page.image, dims=[1]).contiguous(),
                      np.stack([W - 1 - f["seeds"][:, 0], f["seeds"][:, 1]], 1), f["half"], f["model"])
    both = (mirrored["status"] == 0) & (res["status"] == 0)
    assert both.mean() >= 0.98  # ~1 % of this page's fits stop at the iteration cap
    c = mirrored["centers"][both]
    d = np.abs(np.stack([W - 1 - c[:, 0], c[:, 1]], 1) - res["centers"][both]).max()
    print("max centre difference of the mirrored page:", d)
    assert d <= 1e-4
    again = _refit(page.image, f["seeds"], f["half"], f["model"])
    for k in ("params", "centers", "status", "iters"):
        np.testing.assert_array_equal(again[k], res[k])


@pytest.fixture(scope="module")
def tiled():
    from imageprocessor_new_amd import learned
    from imageprocessor_new_amd.backends import gpu_device
    from imageprocessor_new_amd.centers import detect_centers

    H = W = 32768
    print("rendering the 32768 x 32768 mixed page", flush=True)
    holes, _, _ = synthetic.mixed_page_holes(seed=3, H=H, W=W)
    img_t = synthetic.render_page(holes, H, W, seed=3)
    with gpu_device(0) as dev:
        print("training a small network", flush=True)
        net, _ = learned.train_on_pages(n_patches=100_000, epochs=2, batch=4096, seed=10_000, device=dev)
        print("tiled detection", flush=True)
        strips = []
        res = detect_centers(img_t, backend="gpu", threshold="otsu", tile_h=8192, overlap=128, refine="erf_disk",
                             area_min=20, area_max=5000, fit_cells=True, learned_start=net, learned_iters=4,
                             fit_inputs_out=strips)
    return holes, img_t, res, strips


def test_tiled_page_finds_every_hole_once(tiled):
    from scipy.spatial import cKDTree

    holes, img_t, res, strips = tiled
    truth = holes[:, :2].astype(np.float64)
    assert truth.shape[0] > 1_200_000 and len(strips) == 5
    found = res.centers_xy
    d, j = cKDTree(found).query(truth)
    hit = d <= 0.5
    print(f"{truth.shape[0]} holes, {found.shape[0]} centres, {hit.mean():.6f} within 0.5 px, "
          f"median error {np.median(d[hit]):.4f} px, fallbacks {res.meta['learned_fallbacks']}")
    assert hit.mean() >= 0.999
    # exactly once: no two centres share a hole, no centre without a hole
    assert np.unique(j[hit]).size == hit.sum()
    dd, _ = cKDTree(truth).query(found)
    assert (dd <= 0.5).mean() >= 0.999
    # the holes where strips meet (within 64 rows of a cut) are all found
    cuts = np.array([8064 + 64, 16128 + 64, 24192 + 64, 32256 + 64], dtype=np.float64)
    near = np.abs(truth[:, 1][:, None] - cuts[None, :]).min(axis=1) <= 64
    assert near.sum() > 1000 and hit[near].all()


def test_tiled_page_fits_match_the_oracle(tiled):
    holes, img_t, res, strips = tiled
    rng = np.random.default_rng(11)
    worst = 0.0
    for s in strips:
        y0 = s["y0"]
        ok = np.nonzero(s["ok"])[0]
        idx = np.sort(rng.choice(ok, size=40, replace=False))
        y1 = min(img_t.shape[0], y0 + 8192)
        img = img_t[y0:y1].cpu().numpy()
        ref, _, _, ref_st = erfdisk.fit_batch(img, s["seeds"], s["half"], s["model"], neighbours=s["nbr"], only=idx)
        good = idx[ref_st[idx] == 0]
        assert good.size >= 38
        err = np.abs(s["centers"][good] - ref[good, :2]).max()
        worst = max(worst, float(err))
        assert err <= 1e-3, (y0, err)
    print("max |tiled learned fit - fp64 oracle| centre over the strip samples:", worst)
