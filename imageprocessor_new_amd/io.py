"""Centre lists on disk (reference subpx/io.py:14-21): ``.npy`` files of (N, 2) float64 [x, y]."""

from __future__ import annotations

from pathlib import Path

import numpy as np


def save_centers(path, centers_xy) -> Path:
    """Write the centres as a float64 ``.npy`` array; returns the path written (NumPy adds
    ``.npy`` when the name has no such suffix, as the reference's np.save does)."""
    path = Path(path)
    np.save(path, np.asarray(centers_xy, dtype=np.float64))
    return path


def load_centers(path) -> np.ndarray:
    """The array ``save_centers`` wrote (no pickled objects are loaded)."""
    return np.asarray(np.load(Path(path), allow_pickle=False))
