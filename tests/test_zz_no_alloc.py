"""The allocation-free fit entry point (spx_fit_batch_noalloc) against spx_fit_batch_cells,
through the C-ABI, on two calls in a row (the second with fewer features), with and without
neighbour lists; the capacity limit is refused on the host."""

import ctypes

import numpy as np
import pytest
import torch

from imageprocessor_new_amd import _lib, synthetic
from imageprocessor_new_amd.backends import gpu_device, stream_handle
from imageprocessor_new_amd.fit import cell_radius, seed_neighbours_device

pytestmark = pytest.mark.gpu


def _inputs(seed, H, W, dev):
    arr = synthetic.make_workload("hex16384", seed=seed, H=H, W=W)
    seeds = torch.from_numpy(np.rint(arr.centers_xy).astype(np.float32)).to(dev)
    half = torch.from_numpy(np.clip(np.ceil(arr.radius + 3 * arr.sigma), 7, 31).astype(np.int16)).to(dev)
    model = torch.from_numpy(arr.ellipse.astype(np.uint8)).to(dev)
    img = torch.from_numpy(arr.image).to(dev)
    return img, seeds, half, model


def _call(name, img, seeds, half, model, nbr, opts):
    K = int(seeds.shape[0])
    dev = img.device
    out = {"params": torch.full((K, 7), -7.0, dtype=torch.float32, device=dev),
           "centers": torch.full((K, 2), -7.0, dtype=torch.float64, device=dev),
           "rms": torch.full((K,), -7.0, dtype=torch.float32, device=dev),
           "iters": torch.full((K,), -7, dtype=torch.int16, device=dev),
           "status": torch.full((K,), 99, dtype=torch.uint8, device=dev)}
    H, W = img.shape
    p = lambda t: None if t is None else int(t.data_ptr())
    _lib.call(name, p(img), H, W, img.stride(0), p(seeds), p(half), p(model), None, ctypes.byref(opts), K, p(nbr),
              0 if nbr is None else int(nbr.shape[1]), p(out["params"]), p(out["centers"]), p(out["rms"]),
              p(out["iters"]), p(out["status"]), stream_handle())
    torch.cuda.synchronize()
    stats = np.zeros(5, dtype=np.uint64)
    _lib.call("spx_fit_counters", int(stats.ctypes.data), 5, stream_handle())
    return {k: v.cpu().numpy() for k, v in out.items()}, stats


def _same(a, b):
    for k in ("params", "centers", "rms"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    for k in ("iters", "status"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("cells", [False, True])
def test_noalloc_matches_the_allocating_fit_twice(cells):
    with gpu_device(0) as dev:
        for seed, (H, W) in ((3, (600, 600)), (4, (300, 420))):  # the second call has fewer features
            img, seeds, half, model = _inputs(seed, H, W, dev)
            nbr = seed_neighbours_device(seeds, img.shape, cell_radius(int(half.max()))) if cells else None
            opts = _lib.fit_options()
            ref, ref_stats = _call("spx_fit_batch_cells", img, seeds, half, model, nbr, opts)
            got, got_stats = _call("spx_fit_batch_noalloc", img, seeds, half, model, nbr, opts)
            print(f"seed {seed}: {seeds.shape[0]} features, statuses {np.bincount(got['status'])}", flush=True)
            assert (got["status"] <= 1).mean() > 0.9
            _same(got, ref)
            assert np.array_equal(got_stats, ref_stats)


def test_noalloc_refuses_more_than_its_capacity():
    with gpu_device(0) as dev:
        img, seeds, half, model = _inputs(5, 128, 128, dev)
        K = (1 << 20) + 1
        dummy = torch.zeros((K, 2), dtype=torch.float32, device=dev)
        opts = _lib.fit_options()
        p = lambda t: int(t.data_ptr())
        with pytest.raises(ValueError, match="capacity"):
            _lib.call("spx_fit_batch_noalloc", p(img), 128, 128, 128, p(dummy), p(dummy), p(dummy), None,
                      ctypes.byref(opts), K, None, 0, p(dummy), None, p(dummy), p(dummy), p(dummy), stream_handle())
