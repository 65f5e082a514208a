"""Result dataclasses (reference subpx/types.py:9-14, 122-129; ErfFitResult is new)."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np


@dataclass
class CenterResult:
    """Detected centres, reference subpx/types.py:9-14."""

    centers_xy: np.ndarray
    method: str
    backend: str = "cpu"
    meta: dict[str, Any] = field(default_factory=dict)


@dataclass
class ComponentStatsResult:
    """Connected components, reference subpx/types.py:122-129."""

    labels: np.ndarray
    stats: np.ndarray
    centroids: np.ndarray
    num_labels: int
    backend: str = "cpu"
    meta: dict[str, Any] = field(default_factory=dict)


@dataclass
class ErfFitResult:
    """Per-feature erf-disk / erf-ellipse fits.

    ``params`` rows are [x0, y0, r, sigma, A, B, q] in image pixels (q = 1 for disks);
    ``status`` holds SPX_FIT_* codes (0 converged, 1 iteration cap, 2 singular,
    3 diverged, 4 too few pixels, 5 centre left the patch, 6 flat patch). ``centers``
    (K, 2) float64 holds x0, y0 unrounded (``params`` is float32, whose spacing near
    16384 px is 2e-3 px); ``centers_xy`` reads it when present.
    """

    params: np.ndarray
    rms: np.ndarray
    iterations: np.ndarray
    status: np.ndarray
    model: np.ndarray
    backend: str = "gpu"
    meta: dict[str, Any] = field(default_factory=dict)
    centers: np.ndarray | None = None

    @property
    def ok(self) -> np.ndarray:
        return (self.status <= 1) & np.isfinite(self.params[:, :2]).all(axis=1)

    @property
    def centers_xy(self) -> np.ndarray:
        if self.centers is not None:
            return np.asarray(self.centers, dtype=np.float64)
        return self.params[:, :2].astype(np.float64)

    @property
    def cd(self) -> np.ndarray:
        """Hole diameter along x (2 r)."""
        return 2.0 * self.params[:, 2].astype(np.float64)


@dataclass
class OutlineFitResult:
    """Outline fit of one page (``fit_outline``; reference subpx/types.py:156-201): one sharp
    outline and one blur for the page, every hole at its own position, brightness, local level
    and edge offset. Per hole, in seed order: ``centers_xy``, ``cd_px`` (area-equivalent diameter
    of the hole's own outline), ``cd_x_px`` / ``cd_y_px`` (x / y diameters of the equal-area
    ellipse with that outline's spread), ``cd_se_px``, ``edge_offset_px`` (from the good holes'
    mean), ``brightness``, ``level`` (image grey levels), ``rays_px`` (N, n), ``at_clip``, ``good``,
    ``rms``. Per page: ``outline_cd_px``, ``blur_fwhm_px``, ``per_hole_size``, ``converged``,
    ``rounds``, ``blur``, ``theta``, ``template`` and ``stack`` (fine grid, ``upsample`` bins per
    pixel), ``summary``. NumPy arrays, lengths in pixels."""

    centers_xy: np.ndarray
    cd_px: np.ndarray
    cd_x_px: np.ndarray
    cd_y_px: np.ndarray
    cd_se_px: np.ndarray
    edge_offset_px: np.ndarray
    brightness: np.ndarray
    level: np.ndarray
    rays_px: np.ndarray
    at_clip: np.ndarray
    good: np.ndarray
    rms: np.ndarray
    outline_cd_px: float
    blur_fwhm_px: tuple[float, float]
    per_hole_size: bool
    converged: bool
    rounds: int
    blur: str
    theta: np.ndarray
    template: np.ndarray
    stack: np.ndarray
    upsample: int
    summary: dict[str, Any] = field(default_factory=dict)
    pixel_size: float = 1.0
    unit: str = "px"


@dataclass
class OutlineFastResult:
    """Result type of the one-pass outline fit (reference ``fit_outline_fast``,
    subpx/types.py:204-240): one outline and one blur per image, no size per hole. Per hole:
    ``centers_xy``, ``brightness``, ``level``, ``rms``, ``good``; per page: ``outline_cd_px``,
    ``cd_x_px``, ``cd_y_px``, ``rays_px``, ``blur_fwhm_px``, ``theta``, ``blur``, ``stack``,
    ``stack_weight``, ``upsample``, ``summary``. Lengths in pixels; ``pixel_size`` converts
    them to ``unit``."""

    centers_xy: np.ndarray
    brightness: np.ndarray
    level: np.ndarray
    rms: np.ndarray
    good: np.ndarray
    outline_cd_px: float
    cd_x_px: float
    cd_y_px: float
    rays_px: np.ndarray
    blur_fwhm_px: tuple[float, float]
    theta: np.ndarray
    stack: np.ndarray
    stack_weight: np.ndarray
    upsample: int
    blur: str
    summary: dict[str, Any] = field(default_factory=dict)
    pixel_size: float = 1.0
    unit: str = "px"
