"""The graph-recorded batched fit (spx_fit_graph_create / _launch / _destroy) against the
eager spx_fit_batch_cells through the C-ABI: every replay is checked bit for bit against the
eager result, with the program's own eager work and synchronisations between replays; new
contents written into the bound buffers are fitted by the next replay; the learned-start
budget and the counters follow the graph."""

import ctypes

import numpy as np
import pytest
import torch

from imageprocessor_new_amd import _lib, synthetic
from imageprocessor_new_amd.backends import gpu_device, stream_handle
from imageprocessor_new_amd.fit import FitGraph, cell_radius, fit_erf_disks_device, seed_neighbours_device

pytestmark = pytest.mark.gpu


def _inputs(seed, H, W, dev):
    arr = synthetic.make_workload("hex16384", seed=seed, H=H, W=W)
    seeds = torch.from_numpy(np.rint(arr.centers_xy).astype(np.float32)).to(dev)
    half = torch.from_numpy(np.clip(np.ceil(arr.radius + 3 * arr.sigma), 7, 31).astype(np.int16)).to(dev)
    model = torch.from_numpy(arr.ellipse.astype(np.uint8)).to(dev)
    img = torch.from_numpy(arr.image).to(dev)
    return img, seeds, half, model


def _host(out):
    return {k: v.cpu().numpy() for k, v in out.items()}


def _same(a, b):
    for k in ("params", "centers", "rms"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    for k in ("iters", "status"):
        assert np.array_equal(a[k], b[k]), k


def _poison(out):
    for k, v in out.items():
        v.fill_(99 if v.dtype == torch.uint8 else -7)


def _counters():
    stats = np.zeros(5, dtype=np.uint64)
    _lib.call("spx_fit_counters", int(stats.ctypes.data), 5, stream_handle())
    return stats


@pytest.mark.parametrize("cells", [False, True])
def test_graph_replays_equal_the_eager_fit(cells):
    with gpu_device(0) as dev:
        img, seeds, half, model = _inputs(11, 640, 600, dev)
        nbr = seed_neighbours_device(seeds, img.shape, cell_radius(int(half.max()))) if cells else None
        opts = _lib.fit_options(tile_waves=4)
        ref = _host(fit_erf_disks_device(img, seeds, half, model, options=opts, neighbours=nbr))
        torch.cuda.synchronize()
        ref_stats = _counters()
        graph = FitGraph(img, seeds, half, model, options=opts, neighbours=nbr)
        try:
            for replay in range(3):
                _poison(graph.out)
                scratch = torch.randn(1 << 20, device=dev).sum()  # eager work between replays
                graph.launch()
                torch.cuda.synchronize()
                print(f"cells={cells} replay {replay}: {seeds.shape[0]} features, scratch {float(scratch):.3f}",
                      flush=True)
                _same(_host(graph.out), ref)
                assert np.array_equal(_counters(), ref_stats)
            assert (ref["status"] <= 1).mean() > 0.9
        finally:
            graph.close()


def test_graph_fits_new_contents_of_its_buffers():
    with gpu_device(0) as dev:
        img, seeds, half, model = _inputs(12, 512, 512, dev)
        K = int(seeds.shape[0])
        img2, seeds2, half2, model2 = _inputs(13, 512, 512, dev)
        k = min(K, int(seeds2.shape[0]))
        opts = _lib.fit_options(tile_waves=2)
        graph = FitGraph(img, seeds, half, model, options=opts)
        try:
            graph.launch()
            torch.cuda.synchronize()
            first = _host(graph.out)
            _same(first, _host(fit_erf_disks_device(img, seeds, half, model, options=opts)))
            # the next batch of the same size written into the bound buffers
            img.copy_(img2)
            seeds[:k].copy_(seeds2[:k])
            half[:k].copy_(half2[:k])
            model[:k].copy_(model2[:k])
            graph.launch()
            torch.cuda.synchronize()
            second = _host(graph.out)
            print(f"{K} features, {k} replaced", flush=True)
            _same(second, _host(fit_erf_disks_device(img, seeds, half, model, options=opts)))
            assert not np.array_equal(first["params"], second["params"], equal_nan=True)
        finally:
            graph.close()


def test_graph_learned_budget_and_launch_on_a_side_stream():
    with gpu_device(0) as dev:
        img, seeds, half, model = _inputs(14, 480, 520, dev)
        K = int(seeds.shape[0])
        plain = fit_erf_disks_device(img, seeds, half, model, options=_lib.fit_options(tile_waves=4))
        init = plain["params"].clone()
        init[:, :2] += 0.3  # a proposal near the answer
        opts = _lib.fit_options(tile_waves=4, learned_iters=4)
        ref = _host(fit_erf_disks_device(img, seeds, half, model, init, options=opts))
        graph = FitGraph(img, seeds, half, model, init, options=opts)
        side = torch.cuda.Stream(device=dev)
        try:
            side.wait_stream(torch.cuda.current_stream())
            graph.launch(stream=side)
            side.synchronize()
            print(f"{K} features from a proposal on a side stream", flush=True)
            _same(_host(graph.out), ref)
        finally:
            graph.close()


def test_graph_create_refuses_bad_arguments():
    with gpu_device(0) as dev:
        img, seeds, half, model = _inputs(15, 128, 128, dev)
        with pytest.raises(ValueError, match="at least one feature"):
            FitGraph(img, seeds[:0], half[:0], model[:0])
        handle = ctypes.c_void_p()
        opts = _lib.fit_options()
        p = lambda t: int(t.data_ptr())
        with pytest.raises(ValueError, match="nb must lie"):
            _lib.call("spx_fit_graph_create", p(img), 128, 128, 128, p(seeds), p(half), p(model), None,
                      ctypes.byref(opts), int(seeds.shape[0]), p(seeds), 9, p(seeds), None, p(seeds), p(half),
                      p(model), ctypes.byref(handle))
        assert not handle
        _lib.call("spx_fit_graph_destroy", None)
