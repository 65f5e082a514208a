"""On-disk pages (CPU): the library's TIFF reader (spx_tiff_*) against Pillow's reader on files
Pillow writes (none / PackBits / Deflate, 8 and 16 bit, predictor, many strips, many pages,
BigTIFF), and against a big-endian tiled file whose bytes this test spells out; the mosaic
helpers (file keys, stack order, runs, global mapping) on cases worked by hand; the centre
files."""

import struct

import numpy as np
import pytest
from PIL import Image

from imageprocessor_new_amd import _lib
from imageprocessor_new_amd.io import load_centers, save_centers
from imageprocessor_new_amd.mosaic import (TiffFile, TiffKey, build_tiff_index, cell_runs_to_global_rects,
                                           iter_pages_across_vertical, iter_pages_canonical,
                                           map_centers_to_global, parse_tiff_key, true_runs)


def _pages(seed, n, H, W, dtype):
    rng = np.random.default_rng(seed)
    hi = 256 if dtype == np.uint8 else 65536
    # smooth ramps plus noise: runs for PackBits, differences for the predictor
    yy, xx = np.mgrid[:H, :W]
    out = []
    for k in range(n):
        base = ((xx * (k + 3) + yy * 7) % hi).astype(np.int64)
        noise = rng.integers(0, 4, size=(H, W))
        out.append(((base + noise) % hi).astype(dtype))
    return out


def _write_pil(path, pages, **kw):
    frames = [Image.fromarray(p) for p in pages]
    frames[0].save(path, save_all=True, append_images=frames[1:], **kw)


def _pil_pages(path):
    im = Image.open(path)
    out = []
    for k in range(im.n_frames):
        im.seek(k)
        out.append(np.array(im))
    return out


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8])
@pytest.mark.parametrize("kw", [{}, {"compression": "packbits"}, {"compression": "tiff_adobe_deflate"},
                                {"compression": "tiff_deflate"},
                                {"compression": "tiff_adobe_deflate", "tiffinfo": {317: 2}},
                                {"big_tiff": True}, {"big_tiff": True, "compression": "packbits"}],
                         ids=["raw", "packbits", "adobe_deflate", "deflate", "deflate_predictor", "bigtiff",
                              "bigtiff_packbits"])
def test_reader_matches_pillow(tmp_path, dtype, kw):
    pages = _pages(1, 3, 301, 517, dtype)
    path = tmp_path / "p.tif"
    _write_pil(path, pages, **kw)
    ref = _pil_pages(path)
    with TiffFile(path) as tf:
        assert len(tf) == 3
        for k in range(3):
            got = tf.asarray(k)
            assert got.dtype == ref[k].dtype and got.shape == ref[k].shape
            assert np.array_equal(got, ref[k]), k
            assert np.array_equal(got, pages[k]), k


def test_reader_into_a_wider_row_pitch(tmp_path):
    page = _pages(2, 1, 40, 33, np.uint16)[0]
    path = tmp_path / "p.tif"
    _write_pil(path, [page], compression="packbits")
    buf = np.full((40, 50), 7, dtype=np.uint16)
    with TiffFile(path) as tf:
        tf.read_into(0, buf.ctypes.data, buf.strides[0])
    assert np.array_equal(buf[:, :33], page)
    assert (buf[:, 33:] == 7).all()


def _packbits(data: bytes) -> bytes:
    # literal runs of at most 128 bytes and repeats of equal bytes (any valid PackBits stream
    # decodes the same; this one uses both forms)
    out, i = bytearray(), 0
    while i < len(data):
        j = i
        while j < len(data) and data[j] == data[i] and j - i < 128:
            j += 1
        if j - i >= 3:
            out += struct.pack("b", -(j - i - 1)) + data[i:i + 1]
            i = j
            continue
        j = i
        while j < len(data) and j - i < 128 and not (j + 2 < len(data) and data[j] == data[j + 1] == data[j + 2]):
            j += 1
        out += struct.pack("B", j - i - 1) + data[i:j]
        i = j
    return bytes(out)


def _big_endian_tiled(path, page, tile, predictor):
    """A big-endian classic TIFF of one 16-bit page in PackBits tiles of `tile` x `tile`
    (overhanging the page at the right and bottom), with the horizontal predictor."""
    H, W = page.shape
    across, down = -(-W // tile), -(-H // tile)
    blobs = []
    for ty in range(down):
        for tx in range(across):
            t = np.zeros((tile, tile), dtype=np.uint16)
            part = page[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile]
            t[:part.shape[0], :part.shape[1]] = part
            if predictor == 2:
                t = np.concatenate([t[:, :1], (np.diff(t.astype(np.int64), axis=1) % 65536).astype(np.uint16)],
                                   axis=1)
            blobs.append(_packbits(t.astype(">u2").tobytes()))
    entries = []
    data_off = 8
    offsets = []
    for b in blobs:
        offsets.append(data_off)
        data_off += len(b)
    off_arr = data_off
    cnt_arr = off_arr + 4 * len(blobs)
    ifd = cnt_arr + 4 * len(blobs)
    entries = [(256, 3, 1, W), (257, 3, 1, H), (258, 3, 1, 16), (259, 3, 1, 32773), (262, 3, 1, 1),
               (277, 3, 1, 1), (317, 3, 1, predictor), (322, 3, 1, tile), (323, 3, 1, tile),
               (324, 4, len(blobs), off_arr), (325, 4, len(blobs), cnt_arr)]
    with open(path, "wb") as f:
        f.write(b"MM" + struct.pack(">HI", 42, ifd))
        for b in blobs:
            f.write(b)
        f.write(struct.pack(f">{len(blobs)}I", *offsets))
        f.write(struct.pack(f">{len(blobs)}I", *[len(b) for b in blobs]))
        f.write(struct.pack(">H", len(entries)))
        for tag, typ, cnt, val in entries:
            if typ == 3 and cnt == 1:
                f.write(struct.pack(">HHIHH", tag, typ, cnt, val, 0))
            else:
                f.write(struct.pack(">HHII", tag, typ, cnt, val))
        f.write(struct.pack(">I", 0))


@pytest.mark.parametrize("predictor", [1, 2])
def test_big_endian_tiles_spelled_out(tmp_path, predictor):
    page = _pages(3, 1, 70, 45, np.uint16)[0]
    path = tmp_path / "be.tif"
    _big_endian_tiled(path, page, 16, predictor)
    with TiffFile(path) as tf:
        assert tf.pages == [(70, 45, np.dtype(np.uint16))]
        assert np.array_equal(tf.asarray(0), page)


def test_reader_refuses_what_it_does_not_read(tmp_path):
    bad = tmp_path / "x.tif"
    bad.write_bytes(b"not a tiff at all")
    with pytest.raises(_lib.SpxError, match="byte-order"):
        TiffFile(bad)
    rgb = tmp_path / "rgb.tif"
    Image.fromarray(np.zeros((8, 8, 3), dtype=np.uint8)).save(rgb)
    with pytest.raises(_lib.SpxError, match="grey page"):
        TiffFile(rgb)
    flt = tmp_path / "f.tif"
    Image.fromarray(np.zeros((8, 8), dtype=np.float32)).save(flt)
    with pytest.raises(_lib.SpxError, match="grey page"):
        TiffFile(flt)
    with pytest.raises(ValueError, match="cannot open"):
        TiffFile(tmp_path / "missing.tif")


def test_keys_and_stack_order(tmp_path):
    assert parse_tiff_key("/a/b/3_12_7.tif") == TiffKey(3, 12, 7)
    assert parse_tiff_key("0_0_1.TIFF") == TiffKey(0, 0, 1)
    for bad in ("3_12.tif", "x3_12_7.tif", "3_12_7.png", "3_12_7.tif.bak"):
        with pytest.raises(ValueError, match="Bad filename"):
            parse_tiff_key(bad)
    idx = build_tiff_index(["1_10_2.tif", "1_2_2.tif", "0_5_2.tif", "1_7_3.tif", "1_3_2.tiff"])
    assert {k: [p.name for p in v] for k, v in idx.items()} == {
        (1, 2): ["1_2_2.tif", "1_3_2.tiff", "1_10_2.tif"], (0, 2): ["0_5_2.tif"], (1, 3): ["1_7_3.tif"]}


def test_page_order_across_files(tmp_path):
    a, b = _pages(4, 2, 20, 30, np.uint16), _pages(5, 3, 20, 30, np.uint16)
    _write_pil(tmp_path / "0_0_0.tif", a)
    _write_pil(tmp_path / "0_1_0.tif", b)
    files = [tmp_path / "0_0_0.tif", tmp_path / "0_1_0.tif"]
    fwd = list(iter_pages_across_vertical(files))
    assert [k for k, _ in fwd] == [0, 1, 2, 3, 4]
    for (_, got), want in zip(fwd, a + b):
        assert np.array_equal(got, want)
    rev = list(iter_pages_canonical(files, reverse_pages=True))
    assert [k for k, _ in rev] == [0, 1, 2, 3, 4]
    for (_, got), want in zip(rev, (a + b)[::-1]):
        assert np.array_equal(got, want)


def test_runs():
    assert true_runs([]) == []
    assert true_runs([0, 0]) == []
    assert true_runs([1, 1, 0, 1, 0, 0, 1, 1, 1]) == [(0, 2), (3, 4), (6, 9)]
    rng = np.random.default_rng(6)
    for _ in range(200):
        m = rng.random(rng.integers(1, 40)) < 0.5
        runs = true_runs(m)
        back = np.zeros_like(m)
        for s, e in runs:
            assert e > s
            back[s:e] = True
        assert np.array_equal(back, m)
        # maximal: runs neither touch nor overlap
        assert all(runs[i][1] < runs[i + 1][0] for i in range(len(runs) - 1))


def test_global_mapping_worked_by_hand():
    stripe = {"Hprime": 2500, "x0": 100.0, "y0": 2000.0}
    board = {"x0": 30.0, "width": 640.0}
    c = np.array([[10.0, 5.0], [20.0, 990.0], [7.5, 599.0], [8.0, 600.0]])
    # page 2 of 1000 rows, 100 rows cropped: y_stripe = y + 1900, kept while below Hprime = 2500
    g = map_centers_to_global(c, stripe, board, k_canon=2, H=1000, crop_top=100)
    assert np.array_equal(g, [[140.0, 3905.0], [137.5, 4499.0]])
    g = map_centers_to_global(c, stripe, board, k_canon=0, H=1000, flipud=True)
    assert np.array_equal(g, [[140.0, 2994.0], [150.0, 2009.0], [137.5, 2400.0], [138.0, 2399.0]])
    assert map_centers_to_global(np.zeros((0, 2)), stripe, board, k_canon=0, H=1000).shape == (0, 2)
    assert map_centers_to_global(c, stripe, board, k_canon=5, H=1000).shape == (0, 2)
    rects = cell_runs_to_global_rects([1, 1, 0, 1, 1, 1], stripe, board, H=1000, crop_top=100)
    # the second run starts below the stripe's bottom (4900 > 4500) and is dropped
    assert rects == [(130.0, 2000.0, 770.0, 3900.0)]
    rects = cell_runs_to_global_rects([0, 1], {"Hprime": 5000}, {"x0": 1.0}, H=1000, board_width=9.0)
    assert rects == [(1.0, 1000.0, 10.0, 2000.0)]


def test_centre_files(tmp_path):
    c = np.random.default_rng(7).random((11, 2)) * 1000
    p = save_centers(tmp_path / "c.npy", c.astype(np.float32))
    got = load_centers(p)
    assert got.dtype == np.float64 and np.array_equal(got, c.astype(np.float32).astype(np.float64))
