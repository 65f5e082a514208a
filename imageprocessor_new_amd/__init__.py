"""MI355X-native per-feature subpixel refinement for contact-hole images.

Drop-in for the reference ``subpx`` centre path on ``backend="gpu"``: ``detect_centers``,
``detect_centers_tiled``, ``connected_components_stats``, ``dedupe_centers``,
``refine_centers_edge_erf_gpu``, ``refine_centers_edge_moment_gpu``,
``refine_centers_logquad_gpu_match_cpu``, the batched erf-disk fits and the page outline fit
(``fit_outline``), all backed by hand-written
gfx950 HIP kernels in libsubpx_hip.so (C-ABI: include/subpx_hip.h).
"""

from .types import (CenterResult, ComponentStatsResult, ErfFitResult, OutlineFastResult,  # noqa: F401
                    OutlineFitResult)

__version__ = "0.1.0"


def __getattr__(name):
    # torch-dependent modules load lazily so that the dataclasses import without a device.
    if name in ("detect_centers", "dedupe_centers", "filter_centers", "choose_refine_method_for_bbox"):
        from . import centers

        return getattr(centers, name)
    if name in ("detect_centers_tiled", "detect_centers_tiled_global_otsu"):
        from . import legacy

        return getattr(legacy, name)
    if name == "connected_components_stats":
        from .components import connected_components_stats

        return connected_components_stats
    if name in ("fit_erf_disks", "fit_erf_disks_device"):
        from . import fit

        return getattr(fit, name)
    if name in ("refine_centers_edge_erf_gpu", "refine_centers_edge_moment_gpu"):
        from . import edge_erf

        return getattr(edge_erf, name)
    if name in ("fit_outline", "ellipse_widths", "outline_rays", "fit_outline_fast"):
        from . import outline

        return getattr(outline, name)
    if name in ("fit_holes", "HolePage"):
        from . import holes

        return getattr(holes, name)
    if name in ("TiffKey", "parse_tiff_key", "build_tiff_index", "iter_pages_across_vertical", "iter_pages_canonical",
                "map_centers_to_global", "true_runs", "cell_runs_to_global_rects", "process_stripe_board",
                "PageStream"):
        from . import mosaic

        return getattr(mosaic, name)
    if name in ("save_centers", "load_centers"):
        from . import io as centers_io

        return getattr(centers_io, name)
    if name == "refine_centers_logquad_gpu_match_cpu":
        from . import refine

        return refine.refine_centers_logquad_gpu_match_cpu
    raise AttributeError(name)
