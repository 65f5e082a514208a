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
    3 diverged, 4 too few pixels, 5 centre left the patch, 6 flat patch).
    """

    params: np.ndarray
    rms: np.ndarray
    iterations: np.ndarray
    status: np.ndarray
    model: np.ndarray
    backend: str = "gpu"
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def ok(self) -> np.ndarray:
        return (self.status <= 1) & np.isfinite(self.params[:, :2]).all(axis=1)

    @property
    def centers_xy(self) -> np.ndarray:
        return self.params[:, :2].astype(np.float64)

    @property
    def cd(self) -> np.ndarray:
        """Hole diameter along x (2 r)."""
        return 2.0 * self.params[:, 2].astype(np.float64)
