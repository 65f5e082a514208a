"""Deprecated aliases kept for drop-in use (reference subpx/legacy.py:93-100)."""

from __future__ import annotations

import warnings

from .centers import detect_centers


def _warn(old: str, new: str) -> None:
    warnings.warn(f"{old} is deprecated; use {new}", DeprecationWarning, stacklevel=3)


def detect_centers_tiled(image, *, tile_h=8192, **kwargs):
    """Alias of ``detect_centers(image, tile_h=...)`` (reference subpx/legacy.py:93-95)."""
    _warn("detect_centers_tiled", "detect_centers(image, tile_h=...)")
    return detect_centers(image, tile_h=tile_h, **kwargs)


def detect_centers_tiled_global_otsu(image, *, tile_h=8192, threshold_mode="global", **kwargs):
    """Alias with a global Otsu threshold (reference subpx/legacy.py:98-100)."""
    _warn("detect_centers_tiled_global_otsu", "detect_centers(image, tile_h=..., threshold_mode='global')")
    return detect_centers(image, tile_h=tile_h, threshold_mode=threshold_mode, **kwargs)
