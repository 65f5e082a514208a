"""On-disk stripe/board pages into the centre path.

Keeps the reference's mosaic interface (subpx/mosaic.py:16-162): file names
``<stripe>_<vertical>_<board>.tif[f]`` group into (stripe, board) stacks ordered by vertical
index; the pages of a stack are numbered in file order (``k_canon``, optionally reversed);
``process_stripe_board`` detects the centres of every page its cell mask keeps, filters the
frame margin, merges duplicates and maps them into the stripe's global frame.

What differs is where the pages go. The reference decodes each page with tifffile into NumPy
and hands it to ``detect_centers``, which copies it to the GPU before anything runs. Here the
pages are decoded by the library's native reader (spx_tiff_*) straight into pinned host
buffers by a reader thread, uploaded on a copy stream and handed over as device tensors, so
page k + 1 is read and uploaded while page k is detected (``PageStream``).
"""

from __future__ import annotations

import ctypes
import queue
import re
import threading
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib


@dataclass(frozen=True)
class TiffKey:
    stripe: int
    vertical: int
    board: int


_KEY_PATTERN = re.compile(r"(\d+)_(\d+)_(\d+)\.tiff?$", re.IGNORECASE)


def parse_tiff_key(path) -> TiffKey:
    """(stripe, vertical, board) from a file name ``<s>_<v>_<b>.tif[f]`` (reference
    subpx/mosaic.py:26-31: the whole name must match, else ValueError)."""
    name = Path(path).name
    m = _KEY_PATTERN.match(name)
    if m is None:
        raise ValueError(f"Bad filename: {name}")
    s, v, b = (int(g) for g in m.groups())
    return TiffKey(s, v, b)


def build_tiff_index(tiff_paths) -> dict:
    """{(stripe, board): [paths in increasing vertical index]} (reference subpx/mosaic.py:34-43)."""
    groups: dict = {}
    for p in map(Path, tiff_paths):
        k = parse_tiff_key(p)
        groups.setdefault((k.stripe, k.board), []).append((k.vertical, p))
    # a stable sort on the vertical index alone: equal indices keep the listing order
    return {sb: [p for _, p in sorted(items, key=lambda it: it[0])] for sb, items in groups.items()}


class TiffFile:
    """One TIFF opened by the library's reader (spx_tiff_open ... spx_tiff_close)."""

    def __init__(self, path):
        self.path = str(path)
        self._h = ctypes.c_void_p()
        _lib.call("spx_tiff_open", self.path.encode(), ctypes.byref(self._h))
        n = ctypes.c_int32()
        _lib.call("spx_tiff_page_count", self._h, ctypes.byref(n))
        self.pages = []
        for k in range(n.value):
            H, W, bits = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
            _lib.call("spx_tiff_page_info", self._h, k, ctypes.byref(H), ctypes.byref(W), ctypes.byref(bits))
            self.pages.append((int(H.value), int(W.value), np.dtype(np.uint8 if bits.value == 8 else np.uint16)))

    def __len__(self) -> int:
        return len(self.pages)

    def read_into(self, k: int, ptr: int, pitch_bytes: int) -> None:
        """Decode page k into host memory at ``ptr`` (rows ``pitch_bytes`` apart)."""
        _lib.call("spx_tiff_read_page", self._h, int(k), ctypes.c_void_p(int(ptr)), int(pitch_bytes))

    def asarray(self, k: int) -> np.ndarray:
        H, W, dt = self.pages[k]
        out = np.empty((H, W), dtype=dt)
        self.read_into(k, out.ctypes.data, out.strides[0])
        return out

    def close(self) -> None:
        if self._h:
            _lib.call("spx_tiff_close", self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _page_list(files_v, reverse_pages: bool):
    """[(file index, page index, (H, W, dtype))] in canonical order."""
    out = []
    for i, fp in enumerate(files_v):
        with TiffFile(fp) as tf:
            out += [(i, k, info) for k, info in enumerate(tf.pages)]
    return out[::-1] if reverse_pages else out


def iter_pages_across_vertical(files_v):
    """(k, page as a NumPy array) over every page of every file in order (reference
    subpx/mosaic.py:51-58)."""
    k = 0
    for fp in files_v:
        with TiffFile(fp) as tf:
            for j in range(len(tf)):
                yield k, tf.asarray(j)
                k += 1


def iter_pages_canonical(files_v, *, reverse_pages: bool = False):
    """As ``iter_pages_across_vertical``; ``reverse_pages`` numbers the pages from the last page
    of the last file (reference subpx/mosaic.py:61-79)."""
    if not reverse_pages:
        yield from iter_pages_across_vertical(files_v)
        return
    files = [TiffFile(fp) for fp in files_v]
    try:
        order = [(f, j) for f in files for j in range(len(f))][::-1]
        for k_canon, (f, j) in enumerate(order):
            yield k_canon, f.asarray(j)
    finally:
        for f in files:
            f.close()


class PageStream:
    """Pages of a stripe stack as device tensors, read ahead: a reader thread decodes page
    k + 1 into a pinned host buffer while page k is uploaded (copy stream) or processed
    (the consumer's stream). Yields (k_canon, (H, W) uint8/uint16 tensor on ``device``).
    ``keep``: optional predicate on k_canon; pages it refuses are neither read nor uploaded.
    ``depth`` pinned buffers circulate (2 = double buffering)."""

    def __init__(self, files_v, *, reverse_pages: bool = False, device: int = 0, keep=None, depth: int = 2):
        self.files_v = [str(p) for p in files_v]
        self.reverse_pages = bool(reverse_pages)
        self.device = int(device)
        self.keep = keep
        self.depth = max(1, int(depth))
        self.read_seconds = 0.0

    def __iter__(self):
        import time

        import torch

        pages = _page_list(self.files_v, self.reverse_pages)
        work = [(kc, p) for kc, p in enumerate(pages) if self.keep is None or self.keep(kc)]
        if not work:
            return
        dev = torch.device("cuda", self.device)
        nbytes = max(H * W * dt.itemsize for _, (_, _, (H, W, dt)) in work)
        slots = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(self.depth)]
        free: queue.Queue = queue.Queue()
        for i in range(self.depth):
            free.put((i, None))
        ready: queue.Queue = queue.Queue()
        stop = threading.Event()
        stream = self

        def reader():
            files = {}
            try:
                for kc, (fi, j, (H, W, dt)) in work:
                    slot, done = free.get()
                    if stop.is_set():
                        return
                    if done is not None:
                        done.synchronize()  # the slot's previous upload has finished
                    f = files.get(fi)
                    if f is None:
                        f = files[fi] = TiffFile(self.files_v[fi])
                    t0 = time.perf_counter()
                    f.read_into(j, slots[slot].data_ptr(), W * dt.itemsize)
                    stream.read_seconds += time.perf_counter() - t0
                    ready.put((kc, slot, H, W, dt, None))
            except BaseException as e:  # handed to the consumer
                ready.put((None, None, None, None, None, e))
            finally:
                for f in files.values():
                    f.close()

        th = threading.Thread(target=reader, daemon=True)
        th.start()
        copy = torch.cuda.Stream(device=dev)
        try:
            with torch.cuda.device(dev):
                for _ in range(len(work)):
                    kc, slot, H, W, dt, err = ready.get()
                    if err is not None:
                        raise err
                    tdt = torch.uint8 if dt == np.uint8 else torch.uint16
                    host = slots[slot][: H * W * dt.itemsize].view(tdt).view(H, W)
                    img = torch.empty((H, W), dtype=tdt, device=dev)
                    copy.wait_stream(torch.cuda.current_stream(dev))  # img's memory is free
                    with torch.cuda.stream(copy):
                        img.copy_(host, non_blocking=True)
                    done = torch.cuda.Event()
                    done.record(copy)
                    img.record_stream(copy)
                    torch.cuda.current_stream(dev).wait_event(done)
                    free.put((slot, done))
                    yield kc, img
        finally:
            stop.set()
            for i in range(self.depth):
                free.put((i, None))
            th.join()


def map_centers_to_global(centers_xy, stripe_meta: dict, board_meta: dict, *, k_canon: int, H: int,
                          crop_top: int = 0, flipud: bool = False) -> np.ndarray:
    """Page coordinates -> the stripe's global frame (reference subpx/mosaic.py:82-100): y is
    flipped within the page when ``flipud``, offset by k_canon pages of H rows less
    ``crop_top``; only rows inside the stripe's height ``Hprime`` are kept; x is offset by the
    stripe's and the board's x0, y by the stripe's y0."""
    c = np.asarray(centers_xy, dtype=np.float64).reshape(-1, 2)
    if c.shape[0] == 0:
        return c
    y = c[:, 1] if not flipud else (int(H) - 1) - c[:, 1]
    y_stripe = y + float(int(k_canon) * int(H)) - float(crop_top)
    inside = (y_stripe >= 0.0) & (y_stripe < float(int(stripe_meta["Hprime"])))
    if not inside.any():
        return np.zeros((0, 2), dtype=np.float64)
    x_off = float(stripe_meta.get("x0", 0.0)) + float(board_meta.get("x0", 0.0))
    y_off = float(stripe_meta.get("y0", 0.0))
    return np.stack([c[inside, 0] + x_off, y_stripe[inside] + y_off], axis=1)


def true_runs(mask) -> list:
    """[(start, end)) of every run of True values (reference subpx/mosaic.py:103-116)."""
    m = np.asarray(mask, dtype=bool).ravel()
    if m.size == 0:
        return []
    edges = np.diff(np.concatenate([[0], m.astype(np.int8), [0]]))
    starts = np.flatnonzero(edges == 1)
    ends = np.flatnonzero(edges == -1)
    return [(int(a), int(b)) for a, b in zip(starts, ends)]


def cell_runs_to_global_rects(cell_only_mask, stripe_meta: dict, board_meta: dict, *, H: int, crop_top: int = 0,
                              board_width: float | None = None) -> list:
    """Rectangles (x0, y0, x1, y1) in the stripe's frame covered by each run of kept pages,
    clipped to the stripe's height (reference subpx/mosaic.py:119-132)."""
    top = float(stripe_meta.get("y0", 0.0))
    bottom = top + float(stripe_meta["Hprime"])
    left = float(stripe_meta.get("x0", 0.0)) + float(board_meta.get("x0", 0.0))
    width = float(board_meta.get("width", 0.0)) if board_width is None else float(board_width)
    rects = []
    for k0, k1 in true_runs(cell_only_mask):
        y0 = max(top, top + (k0 * int(H) - float(crop_top)))
        y1 = min(bottom, top + (k1 * int(H) - float(crop_top)))
        if y1 > y0:
            rects.append((left, y0, left + width, y1))
    return rects


def process_stripe_board(files_v, stripe_meta: dict, board_meta: dict, cell_only_mask, *, H: int,
                         reverse_pages: bool = False, flipud: bool = False, crop_top: int = 0,
                         detect_kwargs: dict | None = None, margin: float = 5.0, dedupe_eps: float = 1.5,
                         device: int = 0, stats: dict | None = None) -> np.ndarray:
    """Centres of every page of a stripe/board stack whose ``cell_only_mask`` entry is set, in
    the stripe's global frame, pages in canonical order (reference subpx/mosaic.py:135-162):
    ``detect_centers(page, **detect_kwargs)``, ``filter_centers(margin)``,
    ``dedupe_centers(dedupe_eps)``, ``map_centers_to_global``. Pages stream from disk to the GPU
    ahead of the detection (``PageStream``). ``stats`` (a dict) receives the page count and the
    reader's and the whole loop's seconds."""
    import time

    from .centers import dedupe_centers, detect_centers, filter_centers

    kwargs = {} if detect_kwargs is None else dict(detect_kwargs)
    kwargs.setdefault("device", device)
    keep_mask = np.asarray(cell_only_mask, dtype=bool).ravel()

    def keep(kc):
        # the reference indexes the mask by k_canon: a page beyond it raises there as well
        return bool(keep_mask[kc])

    t0 = time.perf_counter()
    found = []
    pages = PageStream(files_v, reverse_pages=reverse_pages, device=device, keep=keep)
    n = 0
    for k_canon, img in pages:
        n += 1
        c = detect_centers(img, **kwargs).centers_xy
        c = filter_centers(c, tuple(img.shape), margin=margin)
        c = dedupe_centers(c, eps=dedupe_eps)
        g = map_centers_to_global(c, stripe_meta, board_meta, k_canon=k_canon, H=H, crop_top=crop_top,
                                  flipud=flipud)
        if g.size:
            found.append(g)
    if stats is not None:
        stats.update(pages=n, read_seconds=pages.read_seconds, seconds=time.perf_counter() - t0)
    if not found:
        return np.zeros((0, 2), dtype=np.float64)
    return np.vstack(found)
