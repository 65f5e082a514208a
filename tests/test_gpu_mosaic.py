"""On-disk stripe pages through the GPU path: PageStream hands over device tensors equal to the
pages Pillow reads back, in canonical order, skipping the pages the cell mask leaves out;
process_stripe_board equals detect_centers page by page (on the pages as NumPy arrays) +
filter_centers + dedupe_centers + map_centers_to_global, and finds every rendered hole."""

import numpy as np
import pytest
import torch
from PIL import Image

from imageprocessor_new_amd import synthetic
from imageprocessor_new_amd.centers import dedupe_centers, detect_centers, filter_centers
from imageprocessor_new_amd.mosaic import PageStream, map_centers_to_global, process_stripe_board

pytestmark = pytest.mark.gpu

# the CPU backend's per-label log-quadratic: the GPU backend's window fit refuses 17 px flat-topped
# disks at its gates, as the reference's GPU algorithm does
DETECT = dict(backend="gpu", threshold="otsu", refine="logquad", area_min=20, area_max=2000, pad=3,
              logquad_variant="cpu")


def _stack(tmp_path, n_files=3, per_file=2, H=384, W=448):
    pages, truth, files = [], [], []
    for f in range(n_files):
        frames = []
        for j in range(per_file):
            arr = synthetic.make_workload("square2048", seed=40 + f * per_file + j, H=H, W=W)
            pages.append(arr.image)
            truth.append(arr.centers_xy)
            frames.append(Image.fromarray(arr.image))
        path = tmp_path / f"2_{f}_5.tif"
        frames[0].save(path, save_all=True, append_images=frames[1:],
                       compression="tiff_adobe_deflate" if f % 2 else "packbits")
        files.append(path)
    return files, pages, truth


@pytest.mark.parametrize("reverse", [False, True])
def test_page_stream_hands_over_the_pages(tmp_path, reverse):
    files, pages, _ = _stack(tmp_path)
    order = pages[::-1] if reverse else pages
    keep = [True, False, True, True, False, True]
    seen = []
    for kc, img in PageStream(files, reverse_pages=reverse, keep=lambda k: keep[k]):
        assert img.is_cuda and img.dtype == torch.uint16
        torch.cuda.current_stream().synchronize()
        assert np.array_equal(img.cpu().numpy(), order[kc]), kc
        seen.append(kc)
    print(f"reverse={reverse}: pages {seen}", flush=True)
    assert seen == [k for k in range(6) if keep[k]]


def test_process_stripe_board_equals_the_page_loop(tmp_path):
    files, pages, truth = _stack(tmp_path)
    H = pages[0].shape[0]
    stripe = {"Hprime": 5 * H + 17, "x0": 1000.0, "y0": 20000.0}
    board = {"x0": 250.0, "width": 448.0}
    mask = [True, True, False, True, True, True]
    stats = {}
    got = process_stripe_board(files, stripe, board, mask, H=H, crop_top=11, flipud=True, detect_kwargs=DETECT,
                               stats=stats)
    want = []
    for k, page in enumerate(pages):
        if not mask[k]:
            continue
        c = detect_centers(page, **DETECT).centers_xy
        c = dedupe_centers(filter_centers(c, page.shape, margin=5.0), eps=1.5)
        g = map_centers_to_global(c, stripe, board, k_canon=k, H=H, crop_top=11, flipud=True)
        if g.size:
            want.append(g)
    want = np.vstack(want)
    print(f"{len(got)} centres over {stats['pages']} pages in {stats['seconds']:.3f}s "
          f"(reader {stats['read_seconds']:.3f}s)", flush=True)
    assert stats["pages"] == 5
    assert np.array_equal(got, want)
    # every hole of a kept page away from the frame margin and inside the stripe is found
    for k, t in enumerate(truth):
        if not mask[k]:
            continue
        inner = t[(t[:, 0] >= 12) & (t[:, 0] < pages[k].shape[1] - 12) & (t[:, 1] >= 12)
                  & (t[:, 1] < H - 12)]
        g = map_centers_to_global(inner, stripe, board, k_canon=k, H=H, crop_top=11, flipud=True)
        d = np.sqrt(((g[:, None, :] - got[None, :, :]) ** 2).sum(-1)).min(axis=1)
        assert d.max() < 0.5, (k, d.max())
