"""Pins of the edge-erf oracle with the reference's own known answers."""

import math

import numpy as np
import pytest
from scipy.special import erf

from imageprocessor_new_amd import synthetic
from oracle import edge_erf


def test_same_edge_found_at_any_position():
    # reference tests/test_gpu_edge_erf.py:24-33
    starts = np.array([0.0, 100.0, 150.0, 500.0, 1300.0, 2000.0])
    k = np.arange(24.0)
    e = np.vectorize(math.erf)
    profile = 15.0 + 70.0 * 0.5 * (1 + e((k - 10.37) / (math.sqrt(2) * 1.3)))
    found, ok = edge_erf.edge_erf(np.tile(profile, (6, 1)), starts, starts + 10.0)
    assert ok.all()
    np.testing.assert_allclose(found - starts, 10.37, atol=1e-6)


@pytest.mark.parametrize("edge,sigma,amp,base,seed_off", [
    (10.0, 1.5, 100.0, 50.0, 0.0), (10.0, 2.5, 100.0, 50.0, 0.0), (10.0, 0.8, 200.0, 20.0, 0.0),
    (10.0, 1.5, -100.0, 150.0, 0.0), (10.0, 1.5, 100.0, 50.0, 0.5), (10.0, 1.5, 100.0, 50.0, 1.5),
    (10.0, 1.5, 100.0, 50.0, -1.0), (10.0, 1.5, 5.0, 50.0, 0.0), (10.0, 5.0, 100.0, 50.0, 0.0),
    (10.0, 0.4, 100.0, 50.0, 0.0)])
def test_reference_diagnostic_profiles(edge, sigma, amp, base, seed_off):
    # reference tests/diag_erf_unit.py:106-123 (ideal profiles, printed there, not asserted).
    # Five Gauss-Newton steps inside a +-5 px window leave wide edges short of the exact
    # position: the bound grows with sigma (measured 1e-3 px at sigma 2.5, 0.023 at 5).
    xs = np.arange(21.0)
    prof = base + amp * 0.5 * (1.0 + erf((xs - edge) / (np.sqrt(2.0) * sigma)))
    found, ok = edge_erf.edge_erf(prof[None], np.array([0.0]), np.array([edge + seed_off]))
    assert ok[0]
    assert abs(found[0] - edge) < (1e-4 if sigma <= 1.5 else 0.05)


def _page(seed=1, amp=150.0, bg=30.0, read=2.0, shape=(256, 256)):
    # 8-bit-like contrast (the reference's pages): the fitter's unitless cond(J^T J) < 1e8
    # gate rejects every edge of a 3000-count hole, whose Jacobian columns differ in scale
    # by the amplitude.
    lat = synthetic.square_lattice(shape[0], shape[1], 32.0)
    rng = np.random.default_rng(seed)
    c = lat + rng.normal(0, 0.3, lat.shape)
    return synthetic.render_hole_array(shape, c, radius=8.0, sigma=1.2, amplitude=amp, background=bg,
                                       read_sigma=read, seed=seed)


def _rows(img):
    from oracle import segment

    t = segment.otsu_u16(img)
    num, labels, stats, cents = segment.connected_components_fast(img > t, 8)
    keep = stats[1:, 4] >= 20
    return np.concatenate([stats[1:, :4], cents[1:]], axis=1)[keep]


def test_refine_centers_on_disks():
    arr = _page()
    c, ok = edge_erf.refine_edge_erf(arr.image.astype(np.float32), _rows(arr.image))
    assert ok.mean() > 0.9
    d = np.hypot(*(c[ok][:, None, :] - arr.centers_xy[None, :, :]).transpose(2, 0, 1)).min(axis=1)
    assert np.median(d) < 0.1


def test_high_amplitude_edges_fail_the_cond_gate():
    arr = _page(amp=3000.0, bg=500.0, read=30.0)
    c, ok = edge_erf.refine_edge_erf(arr.image.astype(np.float32), _rows(arr.image))
    assert not ok.any()


def test_edge_moment_of_symmetric_step_is_its_centre():
    # the |diff|^8 moment of a step symmetric about sample 10.5 is exactly 10.5 (+ base)
    k = np.arange(21.0)
    prof = 40.0 + 90.0 * 0.5 * (1 + erf((k - 10.5) / (np.sqrt(2) * 1.3)))
    x, ok = edge_erf.edge_gradmoment(prof[None], np.array([100.0]))
    assert ok[0]
    assert abs(x[0] - 110.5) < 1e-12


@pytest.mark.parametrize("refine", ["moment", "erf"])
def test_mirror_symmetric_disk_is_centred_exactly(refine):
    # a disk symmetric about (40.5, 37.0): mirrored profiles give mirrored edges, so both
    # refinements return the centre to rounding; a wrong half-sample or base offset fails
    yy, xx = np.mgrid[0:80, 0:84].astype(np.float64)
    rho = np.hypot(xx - 40.5, yy - 37.0)
    img = (30.0 + 150.0 * 0.5 * (1 - erf((rho - 8.0) / (np.sqrt(2) * 1.2)))).astype(np.float32)
    rows = np.array([[32, 29, 18, 17, 40.5, 37.0]], np.float32)
    fn = edge_erf.refine_edge_moment if refine == "moment" else edge_erf.refine_edge_erf
    c, ok = fn(img, rows)
    assert ok[0]
    np.testing.assert_allclose(c[0], [40.5, 37.0], atol=1e-9)
