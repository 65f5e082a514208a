"""The C++ / OpenMP restatement of the CPU detect path (oracle/cpu_detect.cpp), the CPU
baseline timed beside the GPU, against the NumPy oracle it restates."""

import numpy as np
import pytest
from scipy.spatial import cKDTree

from imageprocessor_new_amd import synthetic
from oracle import cpu_detect, segment


@pytest.mark.parametrize("threads", [1, 3])
def test_otsu_equals_the_oracle(threads):
    rng = np.random.default_rng(7)
    for lo, hi, shape in ((0, 4000, (123, 77)), (1000, 1003, (40, 50)), (0, 65535, (64, 64))):
        img = rng.integers(lo, hi + 1, size=shape).astype(np.uint16)
        assert cpu_detect.otsu_u16(img, threads=threads) == segment.otsu_u16(img)


@pytest.mark.parametrize("threads", [1, 4])
def test_detect_equals_the_oracle(threads):
    arr = synthetic.make_workload("square2048", seed=3, H=400, W=360)
    img = arr.image.copy()
    # a spot cut by the border, a thin line (ill-conditioned fit) and a 3-pixel speck
    yy, xx = np.mgrid[0:8, 96:116]
    img[0:8, 96:116] = np.rint(500 + 3000 * np.exp(-((xx - 106.3) ** 2 + (yy + 1.2) ** 2) / 18.0)).astype(np.uint16)
    img[200, 10:40] = 3000
    img[390:392, 350] = 3000
    for kw in (dict(area_min=20, area_max=5000, pad=3), dict(area_min=1, area_max=50000, pad=1)):
        ref, _, t = segment.detect_centers_logquad(img, **kw)
        got, t2 = cpu_detect.detect_centers_logquad(img, threads=threads, **kw)
        assert t2 == t
        assert got.shape == ref.shape and ref.shape[0] > 50
        d, _ = cKDTree(got).query(ref)
        assert d.max() <= 1e-8, d.max()
